// Inverse 2-D DWT levels for CDNA4: reversible 5/3 (int32, floor semantics)
// and irreversible 9/7 (FP32), one LDS tile per work-group.
//
// Reference behaviour: the sliding-window kernels TC/dwt_r53.cl:183-211 and
// TC/dwt_r97.cl:229-256 with the lifting primitives of
// TC/dwt_transform_buffer.cl:102-556 and the mirroring of TC/dwt_io.cl:7-38,
// driven level by level by TC/DWTReverse53.cpp:23-43 / TC/DWTReverse97.cpp:24-44.
// Repaired here: floor instead of C's truncating '/', FP32 neighbours and
// FP32 K scaling on both axes (ITU-T T.800 F.3.8), symmetric extension with the
// origin parity of every level and the length-1 rule (F.3.7).
//
// Design.  A launch runs one level for a list of jobs (tile-components, and in
// the batched form many images).  Each 256-thread work-group owns a TXxTY
// output tile: it gathers the (TX+2H)x(TY+2H) interleaved input straight from
// the four subbands (mirroring at the resolution borders, K-scaling folded into
// the load), lifts every row, then every column, in LDS, and writes the tile
// once with coalesced rows.  HBM traffic is one read of the subbands and one
// write of the output per level; the LL copy-back of the reference
// (TC/DWTKernel.cpp:127-154) does not exist.  The final level also applies the
// inverse RCT/ICT, the DC level shift, rounding and clamping, and packs the
// interleaved 8/16-bit output (the work of TC/Preprocessor.cpp:44-157).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../tc_kernels.h"

namespace tc {
namespace {

constexpr int NT = 256;

// Output tile shapes (TX x TY) a launch can use; the planner's job lists and
// the launch must agree on the shape.
constexpr int kShapes = 4;
__host__ __device__ constexpr int shape_w(int s) { return s == 1 ? 128 : (s == 2 ? 32 : 64); }
__host__ __device__ constexpr int shape_h(int s) { return s == 1 ? 16 : (s == 3 ? 64 : 32); }

// T.800 Table F.4 lifting parameters (the reference's TC/dwt_r97.cl:5-8 and
// TC/dwt_common.cl:10-11 hold the same values, applied as x += c*(l+r)).
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

// A work-group-uniform 64-bit value (a band offset) as a scalar register
// value.  The lane reads keep the compiler from choosing between the band
// offsets through a dynamically addressed copy of Src in scratch memory.
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    return (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v) |
           ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32)) << 32);
}

// Work-group ids go to the 8 XCDs round robin, and each XCD has its own L2.
// This bijection of [0, gridDim.x) gives XCD k (ids k, k+8, ...) one
// contiguous run of work items, so neighbouring tiles, which share halo
// lines, are fetched through the same L2.
__device__ __forceinline__ uint32_t xcd_run_index() {
    const uint32_t n = gridDim.x, b = blockIdx.x, k = b & 7u, q = n >> 3, r = n & 7u;
    return k * q + (k < r ? k : r) + (b >> 3);
}

// Source element of interleaved position (u, v) of the level described by
// (u0,u1,v0,v1): picks LL/HL/LH/HH by parity and returns its element offset.
struct Src {
    uint64_t ll, hl, lh, hh;
    uint32_t llp, bp;
};

__device__ __forceinline__ uint64_t src_index(const Src& s, int u, int v, int u0, int v0) {
    int cx = (u & 1) ? (floor_half(u) - floor_half(u0)) : ((u >> 1) - ceil_half(u0));
    int cy = (v & 1) ? (floor_half(v) - floor_half(v0)) : ((v >> 1) - ceil_half(v0));
    if (!(u & 1) && !(v & 1)) return s.ll + (uint64_t)cy * s.llp + cx;
    if ((u & 1) && !(v & 1)) return s.hl + (uint64_t)cy * s.bp + cx;
    if (!(u & 1)) return s.lh + (uint64_t)cy * s.bp + cx;
    return s.hh + (uint64_t)cy * s.bp + cx;
}

template <bool IS97>
struct Lift;

template <>
struct Lift<false> {
    static constexpr int H = 2;
    typedef int32_t T;
    // 1D_SR 5/3 over the LDS line a[0..n) whose element 0 has parity par0.
    __device__ static void steps(int32_t* base, int stride_pos, int stride_line, int nlines, int n, int par0,
                                 int tid) {
        int half = (n + 1) / 2;
        // STEP1: even -= floor((odd_l + odd_r + 2) / 4)
        for (int i = tid; i < nlines * half; i += NT) {
            int line = i % nlines, j = i / nlines;
            int pos = 2 * j + (par0 & 1);  // absolute-even positions
            if (pos >= 1 && pos < n - 1) {
                int32_t* p = base + line * stride_line + pos * stride_pos;
                p[0] -= (p[-stride_pos] + p[stride_pos] + 2) >> 2;
            }
        }
        __syncthreads();
        // STEP2: odd += floor((even_l + even_r) / 2)
        for (int i = tid; i < nlines * half; i += NT) {
            int line = i % nlines, j = i / nlines;
            int pos = 2 * j + 1 - (par0 & 1);
            if (pos >= 2 && pos < n - 2) {
                int32_t* p = base + line * stride_line + pos * stride_pos;
                p[0] += (p[-stride_pos] + p[stride_pos]) >> 1;
            }
        }
        __syncthreads();
    }
};

template <>
struct Lift<true> {
    static constexpr int H = 4;
    typedef float T;
    __device__ static void step(float* base, int stride_pos, int stride_line, int nlines, int n, int par,
                                int lo, float c, int tid) {
        int half = (n + 1) / 2;
        for (int i = tid; i < nlines * half; i += NT) {
            int line = i % nlines, j = i / nlines;
            int pos = 2 * j + par;
            if (pos >= lo && pos < n - lo) {
                float* p = base + line * stride_line + pos * stride_pos;
                p[0] = fmaf(-c, p[-stride_pos] + p[stride_pos], p[0]);
            }
        }
        __syncthreads();
    }
    // STEP3..6 of 1D_SR 9/7 (the K scaling of STEP1/2 is applied at load).
    __device__ static void steps(float* base, int stride_pos, int stride_line, int nlines, int n, int par0,
                                 int tid) {
        int pe = par0 & 1, po = 1 - pe;  // first even / odd position index
        step(base, stride_pos, stride_line, nlines, n, pe, 1, kDelta, tid);
        step(base, stride_pos, stride_line, nlines, n, po, 2, kGamma, tid);
        step(base, stride_pos, stride_line, nlines, n, pe, 3, kBeta, tid);
        step(base, stride_pos, stride_line, nlines, n, po, 4, kAlpha, tid);
    }
};

// One line of N samples lifted in registers: 1D_SR STEP3..6 for 9/7 (the K
// scaling is applied at load), STEP1..2 for 5/3, over the same ranges as the
// LDS sweeps above (step s updates positions [s, N - s) of its parity).  P:
// parity of position 0.  Read from and written back to LDS at `stride`.
template <bool IS97, int N, int P>
__device__ __forceinline__ void lift_line_inv(typename Lift<IS97>::T* p, int stride) {
    typedef typename Lift<IS97>::T T;
    T x[N];
#pragma unroll
    for (int i = 0; i < N; i++) x[i] = p[i * stride];
    if (IS97) {
        const float k[4] = {kDelta, kGamma, kBeta, kAlpha};
#pragma unroll
        for (int s = 1; s <= 4; s++)
#pragma unroll
            for (int i = s; i < N - s; i++)
                if (((i + P) & 1) == ((s + 1) & 1)) x[i] = fmaf(-k[s - 1], x[i - 1] + x[i + 1], x[i]);
    } else {
#pragma unroll
        for (int i = 1; i < N - 1; i++)
            if (!((i + P) & 1)) x[i] -= ((int)x[i - 1] + (int)x[i + 1] + 2) >> 2;
#pragma unroll
        for (int i = 2; i < N - 2; i++)
            if ((i + P) & 1) x[i] += ((int)x[i - 1] + (int)x[i + 1]) >> 1;
    }
#pragma unroll
    for (int i = 0; i < N; i++) p[i * stride] = x[i];
}

// Lifts NC loaded tiles (tile c at tile + c * cs): every row, then every
// column, one line per thread in registers, unless an axis has length 1.
template <bool IS97, int TX, int TY, int NC>
__device__ __forceinline__ void tile_lift_inv(typename Lift<IS97>::T* tile, int pitch, int ox, int oy, bool h1,
                                              bool v1x, int cs) {
    constexpr int H = Lift<IS97>::H;
    constexpr int W = TX + 2 * H, HH = TY + 2 * H;
    const int tid = threadIdx.x;
    if (!h1) {
        const bool odd = ((ox - H) & 1) != 0;
        for (int L = tid; L < NC * HH; L += NT) {
            typename Lift<IS97>::T* p = tile + (L / HH) * cs + (L % HH) * pitch;
            if (odd)
                lift_line_inv<IS97, W, 1>(p, 1);
            else
                lift_line_inv<IS97, W, 0>(p, 1);
        }
        __syncthreads();
    }
    if (!v1x) {
        const bool odd = ((oy - H) & 1) != 0;
        for (int L = tid; L < NC * W; L += NT) {
            typename Lift<IS97>::T* p = tile + (L / W) * cs + (L % W);
            if (odd)
                lift_line_inv<IS97, HH, 1>(p, pitch);
            else
                lift_line_inv<IS97, HH, 0>(p, pitch);
        }
        __syncthreads();
    }
}

// Loads the extended tile and, with `lift`, lifts rows then columns.  Returns
// with the reconstructed samples at tile[H+ty][H+tx].
template <bool IS97, int TX, int TY>
__device__ void tile_idwt(typename Lift<IS97>::T* tile, int pitch, const typename Lift<IS97>::T* __restrict__ src,
                          const Src& s_in, int u0, int u1, int v0, int v1, int ox, int oy, bool copy_only,
                          bool lift = true) {
    typedef typename Lift<IS97>::T T;
    constexpr int H = Lift<IS97>::H;
    constexpr int W = TX + 2 * H, HH = TY + 2 * H;
    const int tid = threadIdx.x;
    const bool h1 = (u1 - u0) == 1, v1x = (v1 - v0) == 1;
    const uint64_t bll = uni64(s_in.ll), bhl = uni64(s_in.hl), blh = uni64(s_in.lh), bhh = uni64(s_in.hh);
    const uint32_t llp = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_in.llp);
    const uint32_t bp = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_in.bp);
    const Src s{bll, bhl, blh, bhh, llp, bp};
    // interior tile: the extended window needs no mirroring, and every load of
    // a thread is issued before the first is used
    if (!copy_only && ox - H >= u0 && ox + TX + H <= u1 && oy - H >= v0 && oy + TY + H <= v1) {
        constexpr int NIT = (W * HH + NT - 1) / NT;
        const int cu0 = ceil_half(u0), fu0 = floor_half(u0), cv0 = ceil_half(v0), fv0 = floor_half(v0);
        T vals[NIT];
#pragma unroll
        for (int k = 0; k < NIT; k++) {
            const int i = tid + k * NT;
            if (i < W * HH) {
                const int ly = i / W, lx = i % W;
                const int u = ox - H + lx, v = oy - H + ly;
                const int ou = u & 1, ov = v & 1;
                const int cx = ou ? floor_half(u) - fu0 : (u >> 1) - cu0;
                const int cy = ov ? floor_half(v) - fv0 : (v >> 1) - cv0;
                const uint64_t b = ov ? (ou ? bhh : blh) : (ou ? bhl : bll);
                const uint32_t bpitch = (ou | ov) ? bp : llp;
                vals[k] = src[b + (uint64_t)cy * bpitch + cx];
            }
        }
#pragma unroll
        for (int k = 0; k < NIT; k++) {
            const int i = tid + k * NT;
            if (i < W * HH) {
                const int ly = i / W, lx = i % W;
                T val = vals[k];
                if (IS97) {
                    const int ou = (ox - H + lx) & 1, ov = (oy - H + ly) & 1;
                    float f = (float)val;
                    f *= ou ? kInvK : kK;
                    f *= ov ? kInvK : kK;
                    val = (T)f;
                }
                tile[ly * pitch + lx] = val;
            }
        }
        __syncthreads();
        if (lift) tile_lift_inv<IS97, TX, TY, 1>(tile, pitch, ox, oy, false, false, 0);
        return;
    }
    for (int i = tid; i < W * HH; i += NT) {
        int ly = i / W, lx = i % W;
        int u = ox - H + lx, v = oy - H + ly;
        T val;
        if (copy_only) {
            int cu = u < u0 ? u0 : (u >= u1 ? u1 - 1 : u);
            int cv = v < v0 ? v0 : (v >= v1 ? v1 - 1 : v);
            val = src[s.ll + (uint64_t)(cv - v0) * s.llp + (cu - u0)];
        } else {
            int mu = h1 ? u0 : pse(u, u0, u1);
            int mv = v1x ? v0 : pse(v, v0, v1);
            val = src[src_index(s, mu, mv, u0, v0)];
            if (IS97) {
                float f = (float)val;
                if (!h1) f *= (mu & 1) ? kInvK : kK;
                if (!v1x) f *= (mv & 1) ? kInvK : kK;
                val = (T)f;
            }
            // length-1 signal at an odd origin: X = Y / 2 (T.800 F.3.7)
            if (h1 && (u0 & 1)) val = IS97 ? (T)(val * (T)0.5f) : (T)((int32_t)val / 2);
            if (v1x && (v0 & 1)) val = IS97 ? (T)(val * (T)0.5f) : (T)((int32_t)val / 2);
        }
        tile[ly * pitch + lx] = val;
    }
    __syncthreads();
    if (copy_only || !lift) return;
    tile_lift_inv<IS97, TX, TY, 1>(tile, pitch, ox, oy, h1, v1x, 0);
}

template <bool IS97, int TX, int TY>
__global__ __launch_bounds__(NT) void idwt_level_kernel(const tc_dwt_job* __restrict__ jobs,
                                                       const uint32_t* __restrict__ wg_job,
                                                       typename Lift<IS97>::T* __restrict__ base) {
    typedef typename Lift<IS97>::T T;
    constexpr int H = Lift<IS97>::H;
    constexpr int PITCH = TX + 2 * H + 1;
    __shared__ T tile[(TY + 2 * H) * PITCH];
    const uint32_t wg = xcd_run_index();
    const tc_dwt_job& J = jobs[wg_job[wg]];
    int t = (int)wg - J.wg_base;
    int ox = J.u0 + (t % J.tiles_x) * TX, oy = J.v0 + (t / J.tiles_x) * TY;
    Src s{J.ll_off, J.hl_off, J.lh_off, J.hh_off, J.ll_pitch, J.band_pitch};
    tile_idwt<IS97, TX, TY>(tile, PITCH, base, s, J.u0, J.u1, J.v0, J.v1, ox, oy, false);
    for (int i = threadIdx.x; i < TX * TY; i += NT) {
        int ly = i / TX, lx = i % TX;
        int u = ox + lx, v = oy + ly;
        if (u < J.u1 && v < J.v1)
            base[J.out_off + (uint64_t)(v - J.v0) * J.out_pitch + (u - J.u0)] = tile[(ly + H) * PITCH + lx + H];
    }
}

template <typename O>
__device__ __forceinline__ O clamp_out(int v, int lo, int hi) {
    return (O)(v < lo ? lo : (v > hi ? hi : v));
}

template <bool IS97, int NC, typename O, int TX, int TY>
__global__ __launch_bounds__(NT) void idwt_final_kernel(const tc_final_job* __restrict__ jobs,
                                                       const uint32_t* __restrict__ wg_job,
                                                       const typename Lift<IS97>::T* __restrict__ base,
                                                       O* __restrict__ outimg) {
    typedef typename Lift<IS97>::T T;
    constexpr int H = Lift<IS97>::H;
    constexpr int PITCH = TX + 2 * H + 1;
    constexpr int TS = (TY + 2 * H) * PITCH;
    __shared__ T tile[NC * TS];
    const uint32_t wg = xcd_run_index();
    const tc_final_job& J = jobs[wg_job[wg]];
    int t = (int)wg - J.wg_base;
    int ox = J.wx0 + (t % J.tiles_x) * TX, oy = J.wy0 + (t / J.tiles_x) * TY;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        Src s{J.ll_off[c], J.hl_off[c], J.lh_off[c], J.hh_off[c], J.ll_pitch[c], J.band_pitch[c]};
        tile_idwt<IS97, TX, TY>(tile + c * TS, PITCH, base, s, J.u0, J.u1, J.v0, J.v1, ox, oy, J.levels == 0, false);
    }
    // every component's rows, then every component's columns, between the same barriers
    if (J.levels != 0)
        tile_lift_inv<IS97, TX, TY, NC>(tile, PITCH, ox, oy, (J.u1 - J.u0) == 1, (J.v1 - J.v0) == 1, TS);
    __syncthreads();
    for (int i = threadIdx.x; i < TX * TY; i += NT) {
        int ly = i / TX, lx = i % TX;
        int u = ox + lx, v = oy + ly;
        if (u >= J.wx1 || v >= J.wy1) continue;
        int li = (ly + H) * PITCH + lx + H;
        uint64_t o = J.out_off + (uint64_t)(v - J.wy0) * J.out_pitch + (uint64_t)(u - J.wx0) * J.out_cstride;
        if (NC == 3 && J.mct) {
            if (IS97) {
                // inverse ICT, T.800 G.3 (not the YUV constants of TC/preprocess_ict_inverse.cl:31-35)
                float y0 = (float)tile[li], y1 = (float)tile[TS + li], y2 = (float)tile[2 * TS + li];
                float r = fmaf(1.402f, y2, y0);
                float g = fmaf(-0.71414f, y2, fmaf(-0.34413f, y1, y0));
                float b = fmaf(1.772f, y1, y0);
                outimg[o] = clamp_out<O>((int)rintf(r) + J.shift[0], J.minv[0], J.maxv[0]);
                outimg[o + 1] = clamp_out<O>((int)rintf(g) + J.shift[1], J.minv[1], J.maxv[1]);
                outimg[o + 2] = clamp_out<O>((int)rintf(b) + J.shift[2], J.minv[2], J.maxv[2]);
            } else {
                // inverse RCT, T.800 G.2 (TC/preprocess_rct_inverse.cl:18-36)
                int y0 = (int)tile[li], y1 = (int)tile[TS + li], y2 = (int)tile[2 * TS + li];
                int g = y0 - ((y1 + y2) >> 2);
                outimg[o] = clamp_out<O>(y2 + g + J.shift[0], J.minv[0], J.maxv[0]);
                outimg[o + 1] = clamp_out<O>(g + J.shift[1], J.minv[1], J.maxv[1]);
                outimg[o + 2] = clamp_out<O>(y1 + g + J.shift[2], J.minv[2], J.maxv[2]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < NC; c++) {
                int iv = IS97 ? (int)rintf((float)tile[c * TS + li]) : (int)tile[c * TS + li];
                outimg[o + c] = clamp_out<O>(iv + J.shift[c], J.minv[c], J.maxv[c]);
            }
        }
    }
}

}  // namespace

template <int S>
static void level_launch(hipStream_t s, bool is97, const tc_dwt_job* jobs, const uint32_t* wg_job, int nwg,
                         void* base) {
    constexpr int TX = shape_w(S), TY = shape_h(S);
    if (is97)
        hipLaunchKernelGGL((idwt_level_kernel<true, TX, TY>), dim3(nwg), dim3(NT), 0, s, jobs, wg_job, (float*)base);
    else
        hipLaunchKernelGGL((idwt_level_kernel<false, TX, TY>), dim3(nwg), dim3(NT), 0, s, jobs, wg_job,
                           (int32_t*)base);
}

hipError_t launch_idwt_level(hipStream_t s, int shape, bool is97, const tc_dwt_job* jobs, const uint32_t* wg_job,
                             int nwg, void* base) {
    if (nwg <= 0) return hipSuccess;
    switch (shape) {
        case 0: level_launch<0>(s, is97, jobs, wg_job, nwg, base); break;
        case 1: level_launch<1>(s, is97, jobs, wg_job, nwg, base); break;
        case 2: level_launch<2>(s, is97, jobs, wg_job, nwg, base); break;
        case 3: level_launch<3>(s, is97, jobs, wg_job, nwg, base); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int S>
static hipError_t final_launch(hipStream_t s, bool is97, int nc, int out_bytes, const tc_final_job* jobs,
                               const uint32_t* wg_job, int nwg, const void* base, void* out) {
    constexpr int TX = shape_w(S), TY = shape_h(S);
#define TC_FINAL(I97, NCC, OT)                                                                            \
    hipLaunchKernelGGL((idwt_final_kernel<I97, NCC, OT, TX, TY>), dim3(nwg), dim3(NT), 0, s, jobs, wg_job, \
                       (const typename Lift<I97>::T*)base, (OT*)out)
    if (nc == 3) {
        if (is97) {
            if (out_bytes == 1) TC_FINAL(true, 3, uint8_t); else TC_FINAL(true, 3, uint16_t);
        } else {
            if (out_bytes == 1) TC_FINAL(false, 3, uint8_t); else TC_FINAL(false, 3, uint16_t);
        }
    } else if (nc == 1) {
        if (is97) {
            if (out_bytes == 1) TC_FINAL(true, 1, uint8_t); else TC_FINAL(true, 1, uint16_t);
        } else {
            if (out_bytes == 1) TC_FINAL(false, 1, uint8_t); else TC_FINAL(false, 1, uint16_t);
        }
    } else {
        return hipErrorInvalidValue;
    }
#undef TC_FINAL
    return hipSuccess;
}

hipError_t launch_idwt_final(hipStream_t s, int shape, bool is97, int nc, int out_bytes, const tc_final_job* jobs,
                             const uint32_t* wg_job, int nwg, const void* base, void* out) {
    if (nwg <= 0) return hipSuccess;
    hipError_t e;
    switch (shape) {
        case 0: e = final_launch<0>(s, is97, nc, out_bytes, jobs, wg_job, nwg, base, out); break;
        case 1: e = final_launch<1>(s, is97, nc, out_bytes, jobs, wg_job, nwg, base, out); break;
        case 2: e = final_launch<2>(s, is97, nc, out_bytes, jobs, wg_job, nwg, base, out); break;
        case 3: e = final_launch<3>(s, is97, nc, out_bytes, jobs, wg_job, nwg, base, out); break;
        default: return hipErrorInvalidValue;
    }
    return e != hipSuccess ? e : hipGetLastError();
}

int idwt_shape_count() { return kShapes; }
int idwt_tile_w(int shape) { return shape_w(shape); }
int idwt_tile_h(int shape) { return shape_h(shape); }

}  // namespace tc
