// Forward 2-D DWT levels for CDNA4 (encoder side): reversible 5/3 (int32,
// floor semantics) and irreversible 9/7 (FP32), one LDS tile per work-group,
// and a first level fused with the DC level shift and the forward RCT/ICT.
//
// Reference behaviour: the sliding-window forward kernels TC/dwt_f53.cl:184-210
// and TC/dwt_f97.cl:220-248 (lifting primitives TC/dwt_transform_buffer.cl,
// driven level by level by TC/DWTForward53.cpp:20-43 / TC/DWTForward97.cpp),
// and rct_kernel TC/preprocess_rct.cl:18, ict_kernel TC/preprocess_ict.cl:20,
// fdc_level_shift_kernel TC/preprocess_dc_level_shift.cl:16.  Repaired here:
// floor instead of C's truncating '/', T.800 F.4 lifting with the K scaling on
// both axes, symmetric extension with the origin parity of every level and the
// length-1 rule (F.4.2), and the T.800 G.3 ICT instead of YUV-style constants.
//
// Design.  The mirror image of the inverse kernels: a 256-thread work-group
// owns a TXxTY tile of the resolution, loads it with a halo of H samples
// (2 for 5/3, 4 for 9/7) mirrored at the resolution borders, lifts every
// column (VER_SD) and then every row (HOR_SD) in LDS, and writes the four
// quarter-tiles to the subbands with consecutive threads on consecutive band
// elements.  HBM traffic per level is one read of the resolution and one write
// of its subbands; the first level reads the 8/16-bit image itself.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../tc_kernels.h"

namespace tc {
namespace {

constexpr int NT = 256;

// Tile shapes: the same four as the inverse launches.
__host__ __device__ constexpr int fshape_w(int s) { return s == 1 ? 128 : (s == 2 ? 32 : 64); }
__host__ __device__ constexpr int fshape_h(int s) { return s == 1 ? 16 : (s == 3 ? 64 : 32); }

// T.800 Table F.4 lifting parameters.
constexpr float kAlpha = -1.586134342059924f;
constexpr float kBeta = -0.052980118572961f;
constexpr float kGamma = 0.882911075530934f;
constexpr float kDelta = 0.443506852043971f;
constexpr float kK = 1.230174104914001f;
constexpr float kInvK = (float)(1.0 / 1.230174104914001);

// Periodic symmetric extension of index i into [i0, i1), i1 - i0 >= 2.
__device__ __forceinline__ int pse(int i, int i0, int i1) {
    int per = 2 * (i1 - i0 - 1);
    int m = (i - i0) % per;
    if (m < 0) m += per;
    if (m >= i1 - i0) m = per - m;
    return i0 + m;
}

__device__ __forceinline__ int floor_half(int v) { return v >> 1; }
__device__ __forceinline__ int ceil_half(int v) { return (v + 1) >> 1; }

// Work-group ids go to the 8 XCDs round robin, and each XCD has its own L2.
// This bijection of [0, gridDim.x) gives XCD k (ids k, k+8, ...) one
// contiguous run of work items, so neighbouring tiles, which share halo
// lines, are fetched through the same L2.
__device__ __forceinline__ uint32_t xcd_run_index() {
    const uint32_t n = gridDim.x, b = blockIdx.x, k = b & 7u, q = n >> 3, r = n & 7u;
    return k * q + (k < r ? k : r) + (b >> 3);
}

template <bool IS97>
struct FLift;

// 1D_FILTD 5/3 (T.800 F.4.8.1) over LDS lines whose element 0 has parity par0.
// The sweeps take NC component tiles at once (tile c at base + c * cs), so a
// fused first level lifts every component between the same barriers.
template <>
struct FLift<false> {
    static constexpr int H = 2;
    typedef int32_t T;
    template <int NC = 1>
    __device__ static void steps(int32_t* base, int sp, int sl, int nlines, int n, int par0, int tid, int cs = 0) {
        const int half = (n + 1) / 2;
        const int all = NC * nlines;
        // STEP1: odd -= floor((even_l + even_r) / 2)
        for (int i = tid; i < all * half; i += NT) {
            const int L = i % all, j = i / all;
            const int c = NC == 1 ? 0 : L / nlines, line = NC == 1 ? L : L % nlines;
            const int pos = 2 * j + 1 - (par0 & 1);
            if (pos >= 1 && pos < n - 1) {
                int32_t* p = base + c * cs + line * sl + pos * sp;
                p[0] -= (p[-sp] + p[sp]) >> 1;
            }
        }
        __syncthreads();
        // STEP2: even += floor((odd_l + odd_r + 2) / 4)
        for (int i = tid; i < all * half; i += NT) {
            const int L = i % all, j = i / all;
            const int c = NC == 1 ? 0 : L / nlines, line = NC == 1 ? L : L % nlines;
            const int pos = 2 * j + (par0 & 1);
            if (pos >= 2 && pos < n - 2) {
                int32_t* p = base + c * cs + line * sl + pos * sp;
                p[0] += (p[-sp] + p[sp] + 2) >> 2;
            }
        }
        __syncthreads();
    }
};

// 1D_FILTD 9/7 (T.800 F.4.8.2) STEP1..4; the K scaling of STEP5/6 is applied
// when the subbands are written.
template <>
struct FLift<true> {
    static constexpr int H = 4;
    typedef float T;
    template <int NC>
    __device__ static void step(float* base, int sp, int sl, int nlines, int n, int par, int lo, float k, int tid,
                                int cs) {
        const int half = (n + 1) / 2;
        const int all = NC * nlines;
        for (int i = tid; i < all * half; i += NT) {
            const int L = i % all, j = i / all;
            const int c = NC == 1 ? 0 : L / nlines, line = NC == 1 ? L : L % nlines;
            const int pos = 2 * j + par;
            if (pos >= lo && pos < n - lo) {
                float* p = base + c * cs + line * sl + pos * sp;
                p[0] = fmaf(k, p[-sp] + p[sp], p[0]);
            }
        }
        __syncthreads();
    }
    template <int NC = 1>
    __device__ static void steps(float* base, int sp, int sl, int nlines, int n, int par0, int tid, int cs = 0) {
        const int pe = par0 & 1, po = 1 - pe;
        step<NC>(base, sp, sl, nlines, n, po, 1, kAlpha, tid, cs);
        step<NC>(base, sp, sl, nlines, n, pe, 2, kBeta, tid, cs);
        step<NC>(base, sp, sl, nlines, n, po, 3, kGamma, tid, cs);
        step<NC>(base, sp, sl, nlines, n, pe, 4, kDelta, tid, cs);
    }
};

// Gain of one axis for a sample of parity p after the lifting: K / 1/K for
// 9/7, and the length-1 rule (an odd-origin single sample is doubled).
template <bool IS97>
__device__ __forceinline__ float axis_gain(bool len1, int org, int p) {
    if (len1) return (org & 1) ? 2.0f : 1.0f;
    return IS97 ? (p ? kK : kInvK) : 1.0f;
}

// One line of N samples lifted in registers (T.800 F.4.8 STEP1..4 for 9/7,
// STEP1..2 for 5/3) over the same ranges as the LDS sweeps above: STEP s
// updates positions [s, N - s) of its parity.  P: parity of position 0.  The
// line is read from and written back to LDS at `stride`.
template <bool IS97, int N, int P>
__device__ __forceinline__ void lift_line(typename FLift<IS97>::T* p, int stride) {
    typedef typename FLift<IS97>::T T;
    T x[N];
#pragma unroll
    for (int i = 0; i < N; i++) x[i] = p[i * stride];
    if (IS97) {
        const float k[4] = {kAlpha, kBeta, kGamma, kDelta};
#pragma unroll
        for (int s = 1; s <= 4; s++)
#pragma unroll
            for (int i = s; i < N - s; i++)
                if (((i + P) & 1) == (s & 1)) x[i] = fmaf(k[s - 1], x[i - 1] + x[i + 1], x[i]);
    } else {
#pragma unroll
        for (int i = 1; i < N - 1; i++)
            if ((i + P) & 1) x[i] -= ((int)x[i - 1] + (int)x[i + 1]) >> 1;
#pragma unroll
        for (int i = 2; i < N - 2; i++)
            if (!((i + P) & 1)) x[i] += ((int)x[i - 1] + (int)x[i + 1] + 2) >> 2;
    }
#pragma unroll
    for (int i = 0; i < N; i++) p[i * stride] = x[i];
}

// Lifts the loaded tiles of NC components (tile c at tile + c * cs): every
// column, then every row, one line per thread in registers (no barrier inside
// a sweep), unless an axis has length 1.
template <bool IS97, int TX, int TY, int NC = 1>
__device__ __forceinline__ void tile_fdwt(typename FLift<IS97>::T* tile, int pitch, int ox, int oy, bool h1,
                                          bool v1x, int cs = 0) {
    constexpr int H = FLift<IS97>::H;
    constexpr int W = TX + 2 * H, HH = TY + 2 * H;
    const int tid = threadIdx.x;
    if (!v1x) {
        const bool odd = ((oy - H) & 1) != 0;
        for (int L = tid; L < NC * W; L += NT) {
            typename FLift<IS97>::T* p = tile + (L / W) * cs + (L % W);
            if (odd)
                lift_line<IS97, HH, 1>(p, pitch);
            else
                lift_line<IS97, HH, 0>(p, pitch);
        }
        __syncthreads();
    }
    if (!h1) {
        const bool odd = ((ox - H) & 1) != 0;
        for (int L = tid; L < NC * HH; L += NT) {
            typename FLift<IS97>::T* p = tile + (L / HH) * cs + (L % HH) * pitch;
            if (odd)
                lift_line<IS97, W, 1>(p, 1);
            else
                lift_line<IS97, W, 0>(p, 1);
        }
        __syncthreads();
    }
}

// Writes the tile's four quarter-tiles to the subbands (consecutive threads on
// consecutive elements of one band row).
template <bool IS97, int TX, int TY>
__device__ __forceinline__ void store_bands(const typename FLift<IS97>::T* tile, int pitch,
                                            typename FLift<IS97>::T* __restrict__ out, uint64_t ll, uint64_t hl,
                                            uint64_t lh, uint64_t hh, uint32_t llp, uint32_t bp, int u0, int u1,
                                            int v0, int v1, int ox, int oy) {
    typedef typename FLift<IS97>::T T;
    constexpr int H = FLift<IS97>::H;
    constexpr int QX = TX / 2, QY = TY / 2;
    const bool h1 = (u1 - u0) == 1, v1x = (v1 - v0) == 1;
    const int cu0 = ceil_half(u0), fu0 = floor_half(u0), cv0 = ceil_half(v0), fv0 = floor_half(v0);
    for (int i = threadIdx.x; i < 4 * QX * QY; i += NT) {
        const int b = i / (QX * QY), rem = i % (QX * QY);
        const int pu = b & 1, pv = b >> 1;  // 0 LL, 1 HL, 2 LH, 3 HH
        const int lx = 2 * (rem % QX) + (pu ^ (ox & 1)), ly = 2 * (rem / QX) + (pv ^ (oy & 1));
        const int u = ox + lx, v = oy + ly;
        if (u >= u1 || v >= v1) continue;
        const int cx = pu ? floor_half(u) - fu0 : (u >> 1) - cu0;
        const int cy = pv ? floor_half(v) - fv0 : (v >> 1) - cv0;
        T val = tile[(ly + H) * pitch + lx + H];
        if (IS97) {
            val *= axis_gain<true>(h1, u0, pu) * axis_gain<true>(v1x, v0, pv);
        } else {
            if (h1 && (u0 & 1)) val *= 2;
            if (v1x && (v0 & 1)) val *= 2;
        }
        const uint64_t o = b == 0 ? ll + (uint64_t)cy * llp + cx
                                  : (b == 1 ? hl : (b == 2 ? lh : hh)) + (uint64_t)cy * bp + cx;
        out[o] = val;
    }
}

template <bool IS97, int TX, int TY>
__global__ __launch_bounds__(NT) void fdwt_level_kernel(const tc_fdwt_job* __restrict__ jobs,
                                                       const uint32_t* __restrict__ wg_job,
                                                       typename FLift<IS97>::T* __restrict__ base) {
    typedef typename FLift<IS97>::T T;
    constexpr int H = FLift<IS97>::H;
    constexpr int W = TX + 2 * H, HH = TY + 2 * H;
    constexpr int PITCH = W + 1;
    __shared__ T tile[HH * PITCH];
    const uint32_t wg = xcd_run_index();
    const tc_fdwt_job& J = jobs[wg_job[wg]];
    const int t = (int)wg - J.wg_base;
    const int ox = J.u0 + (t % J.tiles_x) * TX, oy = J.v0 + (t / J.tiles_x) * TY;
    const bool h1 = (J.u1 - J.u0) == 1, v1x = (J.v1 - J.v0) == 1;
    if (ox - H >= J.u0 && ox + TX + H <= J.u1 && oy - H >= J.v0 && oy + TY + H <= J.v1) {
        // interior tile: the extended window needs no mirroring
        const T* src = base + J.in_off + (uint64_t)(oy - H - J.v0) * J.in_pitch + (ox - H - J.u0);
        for (int i = threadIdx.x; i < W * HH; i += NT) {
            const int ly = i / W, lx = i % W;
            tile[ly * PITCH + lx] = src[(uint64_t)ly * J.in_pitch + lx];
        }
    } else {
        for (int i = threadIdx.x; i < W * HH; i += NT) {
            const int ly = i / W, lx = i % W;
            const int u = ox - H + lx, v = oy - H + ly;
            const int mu = h1 ? J.u0 : pse(u, J.u0, J.u1);
            const int mv = v1x ? J.v0 : pse(v, J.v0, J.v1);
            tile[ly * PITCH + lx] = base[J.in_off + (uint64_t)(mv - J.v0) * J.in_pitch + (mu - J.u0)];
        }
    }
    __syncthreads();
    tile_fdwt<IS97, TX, TY>(tile, PITCH, ox, oy, h1, v1x);
    store_bands<IS97, TX, TY>(tile, PITCH, base, J.ll_off, J.hl_off, J.lh_off, J.hh_off, J.ll_pitch, J.band_pitch,
                              J.u0, J.u1, J.v0, J.v1, ox, oy);
}

template <bool IS97, int NC, typename I, int TX, int TY>
__global__ __launch_bounds__(NT) void fdwt_first_kernel(const tc_fdwt_first_job* __restrict__ jobs,
                                                       const uint32_t* __restrict__ wg_job,
                                                       const I* __restrict__ img,
                                                       typename FLift<IS97>::T* __restrict__ base) {
    typedef typename FLift<IS97>::T T;
    constexpr int H = FLift<IS97>::H;
    constexpr int W = TX + 2 * H, HH = TY + 2 * H;
    constexpr int PITCH = W + 1;
    constexpr int TS = HH * PITCH;
    __shared__ T tile[NC * TS];
    const uint32_t wg = xcd_run_index();
    const tc_fdwt_first_job& J = jobs[wg_job[wg]];
    const int t = (int)wg - J.wg_base;
    const int ox = J.u0 + (t % J.tiles_x) * TX, oy = J.v0 + (t / J.tiles_x) * TY;
    const bool h1 = (J.u1 - J.u0) == 1, v1x = (J.v1 - J.v0) == 1;
    const bool interior = ox - H >= J.u0 && ox + TX + H <= J.u1 && oy - H >= J.v0 && oy + TY + H <= J.v1;
    for (int i = threadIdx.x; i < W * HH; i += NT) {
        const int ly = i / W, lx = i % W;
        const int u = ox - H + lx, v = oy - H + ly;
        // interior tiles need no mirroring
        const int mu = interior ? u : (h1 ? J.u0 : pse(u, J.u0, J.u1));
        const int mv = interior ? v : (v1x ? J.v0 : pse(v, J.v0, J.v1));
        const I* px = img + J.in_off + (uint64_t)(mv - J.v0) * J.in_pitch + (uint64_t)(mu - J.u0) * J.in_cstride;
        int s[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) s[c] = (int)px[c] - J.shift[c];
        if (NC == 3 && J.mct) {
            if (IS97) {
                // forward ICT, T.800 G.3 (the matrix of TC/preprocess_ict.cl:29-31)
                const float r = (float)s[0], g = (float)s[1], b = (float)s[2];
                tile[ly * PITCH + lx] = fmaf(0.299f, r, fmaf(0.587f, g, 0.114f * b));
                tile[TS + ly * PITCH + lx] = fmaf(-0.16875f, r, fmaf(-0.33126f, g, 0.5f * b));
                tile[2 * TS + ly * PITCH + lx] = fmaf(0.5f, r, fmaf(-0.41869f, g, -0.08131f * b));
                continue;
            } else {
                // forward RCT, T.800 G.2 (TC/preprocess_rct.cl:18-31)
                const int y = (s[0] + 2 * s[1] + s[2]) >> 2, cb = s[2] - s[1], cr = s[0] - s[1];
                s[0] = y;
                s[1] = cb;
                s[2] = cr;
            }
        }
#pragma unroll
        for (int c = 0; c < NC; c++) tile[c * TS + ly * PITCH + lx] = (T)s[c];
    }
    __syncthreads();
    if (J.levels == 0) {
        // no DWT: the tile's own samples go to the LL plane
        for (int i = threadIdx.x; i < TX * TY; i += NT) {
            const int ly = i / TX, lx = i % TX;
            const int u = ox + lx, v = oy + ly;
            if (u >= J.u1 || v >= J.v1) continue;
#pragma unroll
            for (int c = 0; c < NC; c++)
                base[J.ll_off[c] + (uint64_t)(v - J.v0) * J.ll_pitch[c] + (u - J.u0)] =
                    tile[c * TS + (ly + H) * PITCH + lx + H];
        }
        return;
    }
    tile_fdwt<IS97, TX, TY, NC>(tile, PITCH, ox, oy, h1, v1x, TS);
#pragma unroll
    for (int c = 0; c < NC; c++)
        store_bands<IS97, TX, TY>(tile + c * TS, PITCH, base, J.ll_off[c], J.hl_off[c], J.lh_off[c], J.hh_off[c],
                                  J.ll_pitch[c], J.band_pitch[c], J.u0, J.u1, J.v0, J.v1, ox, oy);
}

template <int S>
static void flevel_launch(hipStream_t s, bool is97, const tc_fdwt_job* jobs, const uint32_t* wg_job, int nwg,
                          void* base) {
    constexpr int TX = fshape_w(S), TY = fshape_h(S);
    if (is97)
        hipLaunchKernelGGL((fdwt_level_kernel<true, TX, TY>), dim3(nwg), dim3(NT), 0, s, jobs, wg_job, (float*)base);
    else
        hipLaunchKernelGGL((fdwt_level_kernel<false, TX, TY>), dim3(nwg), dim3(NT), 0, s, jobs, wg_job,
                           (int32_t*)base);
}

template <int S>
static hipError_t ffirst_launch(hipStream_t s, bool is97, int nc, int in_bytes, const tc_fdwt_first_job* jobs,
                                const uint32_t* wg_job, int nwg, const void* img, void* base) {
    constexpr int TX = fshape_w(S), TY = fshape_h(S);
#define TC_FFIRST(I97, NCC, IT)                                                                            \
    hipLaunchKernelGGL((fdwt_first_kernel<I97, NCC, IT, TX, TY>), dim3(nwg), dim3(NT), 0, s, jobs, wg_job, \
                       (const IT*)img, (typename FLift<I97>::T*)base)
    if (nc == 3) {
        if (is97) {
            if (in_bytes == 1) TC_FFIRST(true, 3, uint8_t); else TC_FFIRST(true, 3, uint16_t);
        } else {
            if (in_bytes == 1) TC_FFIRST(false, 3, uint8_t); else TC_FFIRST(false, 3, uint16_t);
        }
    } else if (nc == 1) {
        if (is97) {
            if (in_bytes == 1) TC_FFIRST(true, 1, uint8_t); else TC_FFIRST(true, 1, uint16_t);
        } else {
            if (in_bytes == 1) TC_FFIRST(false, 1, uint8_t); else TC_FFIRST(false, 1, uint16_t);
        }
    } else {
        return hipErrorInvalidValue;
    }
#undef TC_FFIRST
    return hipSuccess;
}

}  // namespace

hipError_t launch_fdwt_level(hipStream_t s, int shape, bool is97, const tc_fdwt_job* jobs, const uint32_t* wg_job,
                             int nwg, void* base) {
    if (nwg <= 0) return hipSuccess;
    switch (shape) {
        case 0: flevel_launch<0>(s, is97, jobs, wg_job, nwg, base); break;
        case 1: flevel_launch<1>(s, is97, jobs, wg_job, nwg, base); break;
        case 2: flevel_launch<2>(s, is97, jobs, wg_job, nwg, base); break;
        case 3: flevel_launch<3>(s, is97, jobs, wg_job, nwg, base); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fdwt_first(hipStream_t s, int shape, bool is97, int nc, int in_bytes,
                             const tc_fdwt_first_job* jobs, const uint32_t* wg_job, int nwg, const void* img,
                             void* base) {
    if (nwg <= 0) return hipSuccess;
    hipError_t e;
    switch (shape) {
        case 0: e = ffirst_launch<0>(s, is97, nc, in_bytes, jobs, wg_job, nwg, img, base); break;
        case 1: e = ffirst_launch<1>(s, is97, nc, in_bytes, jobs, wg_job, nwg, img, base); break;
        case 2: e = ffirst_launch<2>(s, is97, nc, in_bytes, jobs, wg_job, nwg, img, base); break;
        case 3: e = ffirst_launch<3>(s, is97, nc, in_bytes, jobs, wg_job, nwg, img, base); break;
        default: return hipErrorInvalidValue;
    }
    return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace tc
