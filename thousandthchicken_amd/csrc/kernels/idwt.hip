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
// the load) into LDS.  Lifting runs on every thread: a row is split into
// segments of 8 outputs, each lifted in registers together with the halo it
// shares with its neighbours (the halo is computed twice, which costs less
// than a barrier per lifting step) and written back in place; then each
// thread lifts one segment of one column the same way and writes its outputs
// straight to HBM, a wave covering 64 consecutive columns.  HBM traffic is one
// read of the subbands and one write of the output per level; the LL copy-back of the reference
// (TC/DWTKernel.cpp:127-154) does not exist.  The final level also applies the
// inverse RCT/ICT, the DC level shift, rounding and clamping, and packs the
// interleaved 8/16-bit output (the work of TC/Preprocessor.cpp:44-157).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

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
    static constexpr int H = 2;  // halo: one sample per lifting step
    typedef int32_t T;
};

template <>
struct Lift<true> {
    static constexpr int H = 4;
    typedef float T;
};

// 1D_SR over N samples in registers, x[0] at a position of parity P:
// STEP3..6 for 9/7 (the K scaling is applied at load), STEP1..2 for 5/3.
// Step s updates the positions [s, N - s) of its parity, so with N = K + 2H
// the middle K samples come out exact and the H on either side are the halo
// the segment shares with its neighbours.
template <bool IS97, int N, int P>
__device__ __forceinline__ void lift_regs(typename Lift<IS97>::T (&x)[N]) {
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
}

// Segment lengths: a row of the window is lifted as TX/KR segments of KR
// outputs, and the columns as TX * TY/KC segments of KC outputs, at most one
// per thread.  At least 8 outputs per segment keep the recomputed halo (2H
// samples) from outweighing the work.
constexpr int kRowSeg = 8;
__host__ __device__ constexpr int col_seg(int tx, int ty) { return tx * ty / NT > 8 ? tx * ty / NT : 8; }

// Mirror of i into [i0, i1) for a window reaching at most one halo past the
// ends (the length exceeds the halo); indices further out, which only feed
// samples outside the rectangle, are clamped.
__device__ __forceinline__ int reflect1(int i, int i0, int i1) {
    int m = i < i0 ? 2 * i0 - i : (i >= i1 ? 2 * (i1 - 1) - i : i);
    return m < i0 ? i0 : (m >= i1 ? i1 - 1 : m);
}

// Loads the extended (TX + 2H) x (TY + 2H) window of one tile-component into
// LDS (pitch `pitch`): every sample straight from its subband, mirrored at the
// resolution borders, with the 9/7 K scaling folded in.  Thread t owns window
// column t % W and the rows t / W + R * m (R = NT / W), so the horizontal
// mirroring, band, source column and scale are worked out once per thread and
// each sample costs only its row's part of the address.  All of a thread's
// loads are issued before the first is used.
template <bool IS97, int TX, int TY>
__device__ __forceinline__ void tile_load(typename Lift<IS97>::T* tile, int pitch,
                                          const typename Lift<IS97>::T* __restrict__ src, const Src& s_in, int u0,
                                          int u1, int v0, int v1, int ox, int oy, bool copy_only) {
    typedef typename Lift<IS97>::T T;
    constexpr int H = Lift<IS97>::H;
    constexpr int W = TX + 2 * H, HH = TY + 2 * H;
    constexpr int R = NT / W;                  // rows in flight per pass over the window
    constexpr int MR = (HH + R - 1) / R;       // rows per thread
    static_assert(R >= 1, "window wider than the work-group");
    const int tid = threadIdx.x;
    if (tid >= R * W) return;
    const int lx = tid % W, ty = tid / W;
    const bool h1 = (u1 - u0) == 1, v1x = (v1 - v0) == 1;
    const uint64_t bll = uni64(s_in.ll), bhl = uni64(s_in.hl), blh = uni64(s_in.lh), bhh = uni64(s_in.hh);
    const uint32_t llp = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_in.llp);
    const uint32_t bp = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_in.bp);
    // 0: no mirroring needed; 1: one reflection at most; 2: short signals (general PSE)
    const bool interior = ox - H >= u0 && ox + TX + H <= u1 && oy - H >= v0 && oy + TY + H <= v1;
    const int mode = interior ? 0 : ((u1 - u0 > H && v1 - v0 > H) ? 1 : 2);
    auto mirror = [&](int i, int i0, int i1, bool one) -> int {
        if (one) return i0;
        return mode == 0 ? i : (mode == 1 ? reflect1(i, i0, i1) : pse(i, i0, i1));
    };
    // this thread's column: its band side, source column and base addresses
    const int u = ox - H + lx;
    const int mu = copy_only ? (u < u0 ? u0 : (u >= u1 ? u1 - 1 : u)) : mirror(u, u0, u1, h1);
    const int ou = copy_only ? 0 : (mu & 1);
    const int cx = copy_only ? mu - u0 : (ou ? floor_half(mu) - floor_half(u0) : (mu >> 1) - ceil_half(u0));
    const uint64_t col_lo = (ou ? bhl : bll) + (uint64_t)(int64_t)cx;  // rows of even v
    const uint64_t col_hi = (ou ? bhh : blh) + (uint64_t)(int64_t)cx;  // rows of odd v
    const uint32_t pitch_lo = ou ? bp : llp;
    const int cv0 = ceil_half(v0), fv0 = floor_half(v0);
    float sc_lo = 1.0f, sc_hi = 1.0f;  // K scaling of the two vertical bands
    if (IS97 && !copy_only) {
        const float hs = h1 ? 1.0f : (ou ? kInvK : kK);
        sc_lo = v1x ? hs : hs * kK;
        sc_hi = v1x ? hs : hs * kInvK;
    }
    T vals[MR];
    if (mode == 0 && !copy_only) {
        // Interior: window rows ty + R*m of one parity are consecutive rows of
        // one vertical band, so each load is a base plus a fixed multiple of
        // the band pitch.  Row m = 0 fixes one parity, m = 1 (R odd) the other.
        const int va = oy - H + ty, qa = va & 1, qb = qa ^ (R & 1);
        const int vb = va + R;
        const uint64_t base_a = (qa ? col_hi : col_lo) +
                                (uint64_t)(int64_t)((qa ? floor_half(va) - fv0 : (va >> 1) - cv0)) * (qa ? bp : pitch_lo);
        const uint64_t base_b = (qb ? col_hi : col_lo) +
                                (uint64_t)(int64_t)((qb ? floor_half(vb) - fv0 : (vb >> 1) - cv0)) * (qb ? bp : pitch_lo);
        const uint64_t pa = qa ? bp : pitch_lo, pb = qb ? bp : pitch_lo;
#pragma unroll
        for (int m = 0; m < MR; m++) {
            if (ty + R * m < HH) {
                // rows m and m' share a parity when R * (m - m') is even; R * m / 2 band rows past row 0 (or
                // R * (m - 1) / 2 past row 1)
                if (((R * m) & 1) == 0)
                    vals[m] = src[base_a + (uint64_t)(R * m / 2) * pa];
                else
                    vals[m] = src[base_b + (uint64_t)(R * (m - 1) / 2) * pb];
            }
        }
    } else {
#pragma unroll
    for (int m = 0; m < MR; m++) {
        const int ly = ty + R * m;
        if (ly < HH) {
            const int v = oy - H + ly;
            if (copy_only) {
                const int cv = v < v0 ? v0 : (v >= v1 ? v1 - 1 : v);
                vals[m] = src[col_lo + (uint64_t)(cv - v0) * llp];
            } else {
                const int mv = mirror(v, v0, v1, v1x);
                const int ov = mv & 1;
                const int cy = ov ? floor_half(mv) - fv0 : (mv >> 1) - cv0;
                vals[m] = src[(ov ? col_hi : col_lo) + (uint64_t)cy * (ov ? bp : pitch_lo)];
            }
        }
    }
    }
#pragma unroll
    for (int m = 0; m < MR; m++) {
        const int ly = ty + R * m;
        if (ly < HH) {
            T val = vals[m];
            if (!copy_only) {
                // parity survives the mirroring, so the band follows from v
                const int ov = (oy - H + ly) & 1;
                if (IS97) val = (T)((float)val * (ov ? sc_hi : sc_lo));
                // length-1 signal at an odd origin: X = Y / 2 (T.800 F.3.7)
                if (h1 && (u0 & 1)) val = IS97 ? (T)(val * (T)0.5f) : (T)((int32_t)val / 2);
                if (v1x && (v0 & 1)) val = IS97 ? (T)(val * (T)0.5f) : (T)((int32_t)val / 2);
            }
            tile[ly * pitch + lx] = val;
        }
    }
}

// Lifts every row of NC loaded windows (window c at tile + c * cs) in place,
// over the TX output columns: each thread takes a KR-sample segment of a row
// with its halo into registers.  Rows go in rounds of NT / (TX / KR); within a
// round every segment is read before any is written, and rounds touch
// different rows.  P: parity of window column 0.
template <bool IS97, int TX, int TY, int NC, int P>
__device__ __forceinline__ void rows_pass(typename Lift<IS97>::T* tile, int pitch, int cs) {
    typedef typename Lift<IS97>::T T;
    constexpr int H = Lift<IS97>::H, HH = TY + 2 * H;
    constexpr int KR = kRowSeg, SEG = TX / KR, N = KR + 2 * H;
    constexpr int ROWS = NC * HH, RPR = NT / SEG;
    static_assert(NT % SEG == 0, "row segments must tile the work-group");
    const int tid = threadIdx.x;
    const int seg = tid / RPR, rr = tid % RPR;  // consecutive lanes take consecutive rows (odd pitch: no conflicts)
#pragma unroll
    for (int r0 = 0; r0 < ROWS; r0 += RPR) {
        const int row = r0 + rr < ROWS ? r0 + rr : ROWS - 1;
        T* const p = tile + (row / HH) * cs + (row % HH) * pitch + seg * KR;
        T x[N];
#pragma unroll
        for (int j = 0; j < N; j++) x[j] = p[j];
        __syncthreads();
        if (r0 + rr < ROWS) {
            lift_regs<IS97, N, P>(x);
#pragma unroll
            for (int j = 0; j < KR; j++) p[H + j] = x[H + j];
        }
    }
    __syncthreads();
}

// The column segment of thread tid of window c after the column pass: KC
// samples of column H + col, rows H + seg * KC + j, lifted unless v1x.  P:
// parity of window row 0.
template <bool IS97, int TX, int TY, int P>
__device__ __forceinline__ void column_segment(const typename Lift<IS97>::T* tile, int pitch, int col, int seg,
                                               bool v1x,
                                               typename Lift<IS97>::T (&x)[col_seg(TX, TY) + 2 * Lift<IS97>::H]) {
    constexpr int H = Lift<IS97>::H, KC = col_seg(TX, TY), N = KC + 2 * H;
    const typename Lift<IS97>::T* const p = tile + seg * KC * pitch + H + col;
#pragma unroll
    for (int j = 0; j < N; j++) x[j] = p[j * pitch];
    if (!v1x) lift_regs<IS97, N, P>(x);
}

template <bool IS97, int TX, int TY>
__global__ __launch_bounds__(NT) void idwt_level_kernel(const tc_dwt_job* __restrict__ jobs,
                                                       const uint32_t* __restrict__ wg_job,
                                                       typename Lift<IS97>::T* __restrict__ base) {
    typedef typename Lift<IS97>::T T;
    constexpr int H = Lift<IS97>::H;
    constexpr int PITCH = TX + 2 * H + 1;
    constexpr int KC = col_seg(TX, TY);
    __shared__ T tile[(TY + 2 * H) * PITCH];
    const uint32_t wg = xcd_run_index();
    const tc_dwt_job& J = jobs[wg_job[wg]];
    const int t = (int)wg - J.wg_base;
    const int ox = J.u0 + (t % J.tiles_x) * TX, oy = J.v0 + (t / J.tiles_x) * TY;
    const Src s{J.ll_off, J.hl_off, J.lh_off, J.hh_off, J.ll_pitch, J.band_pitch};
    const bool h1 = (J.u1 - J.u0) == 1, v1x = (J.v1 - J.v0) == 1;
    tile_load<IS97, TX, TY>(tile, PITCH, base, s, J.u0, J.u1, J.v0, J.v1, ox, oy, false);
    __syncthreads();
    // the window's parities pick one of two compiled forms of each pass
    if (!h1) {
        if ((ox - H) & 1) {
            asm volatile("; odd window columns");
            rows_pass<IS97, TX, TY, 1, 1>(tile, PITCH, 0);
        } else {
            rows_pass<IS97, TX, TY, 1, 0>(tile, PITCH, 0);
        }
    }
    const int col = threadIdx.x % TX, seg = threadIdx.x / TX;
    if (seg >= TY / KC) return;
    // each parity form runs its column pass and its stores on its own (a join
    // between them makes the compiler merge the register arrays through
    // dynamically indexed moves)
    auto finish = [&](auto parity) {
        T x[KC + 2 * H];
        column_segment<IS97, TX, TY, decltype(parity)::value>(tile, PITCH, col, seg, v1x, x);
        const int u = ox + col;
        if (u >= J.u1) return;
        const int vs = oy + seg * KC;
        const int nrows = J.v1 - vs;  // rows of the segment inside the rectangle
        T* o = base + J.out_off + (u - J.u0) + (uint64_t)(vs - J.v0) * J.out_pitch;
        if (nrows >= KC) {
#pragma unroll
            for (int j = 0; j < KC; j++, o += J.out_pitch) *o = x[H + j];
        } else {
#pragma unroll
            for (int j = 0; j < KC; j++, o += J.out_pitch)
                if (j < nrows) *o = x[H + j];
        }
    };
    if ((oy - H) & 1) {
        asm volatile("; odd window rows");
        finish(std::integral_constant<int, 1>());
    } else {
        finish(std::integral_constant<int, 0>());
    }
}

// min of max: one v_med3_i32
template <typename O>
__device__ __forceinline__ O clamp_out(int v, int lo, int hi) {
    return (O)min(max(v, lo), hi);
}

// Column pass of the final level for this thread's segment of every
// component, then the inverse colour transform, DC shift, rounding, clamping
// and packing of its KC output pixels.  P: parity of window row 0.
template <bool IS97, int NC, typename O, int TX, int TY, int P>
__device__ __forceinline__ void final_columns(const tc_final_job& J, const typename Lift<IS97>::T* tile, int ox,
                                              int oy, int col, int seg, bool v1x, O* __restrict__ outimg) {
    typedef typename Lift<IS97>::T T;
    constexpr int H = Lift<IS97>::H;
    constexpr int PITCH = TX + 2 * H + 1;
    constexpr int TS = (TY + 2 * H) * PITCH;
    constexpr int KC = col_seg(TX, TY);
    T x[NC][KC + 2 * H];
#pragma unroll
    for (int c = 0; c < NC; c++) column_segment<IS97, TX, TY, P>(tile + c * TS, PITCH, col, seg, v1x, x[c]);
    const int u = ox + col;
    if (u >= J.wx1) return;
    const int vs = oy + seg * KC;
    const int nrows = J.wy1 - vs;  // rows of the segment inside the window (>= KC for all but the last tiles)
    O* const o0 = outimg + J.out_off + (uint64_t)(u - J.wx0) * J.out_cstride + (uint64_t)(vs - J.wy0) * J.out_pitch;
    // one copy of the store loop per (colour transform, whole segment) case,
    // so that neither test runs per pixel
    auto emit = [&](auto mct, auto whole) {
        O* o = o0;
#pragma unroll
        for (int j = 0; j < KC; j++, o += J.out_pitch) {
            if (!decltype(whole)::value && j >= nrows) continue;
            if (decltype(mct)::value) {
                if (IS97) {
                    // inverse ICT, T.800 G.3 (not the YUV constants of TC/preprocess_ict_inverse.cl:31-35)
                    const float y0 = (float)x[0][H + j], y1 = (float)x[NC > 1 ? 1 : 0][H + j],
                                y2 = (float)x[NC > 2 ? 2 : 0][H + j];
                    const float r = fmaf(1.402f, y2, y0);
                    const float g = fmaf(-0.71414f, y2, fmaf(-0.34413f, y1, y0));
                    const float b = fmaf(1.772f, y1, y0);
                    o[0] = clamp_out<O>((int)rintf(r) + J.shift[0], J.minv[0], J.maxv[0]);
                    o[1] = clamp_out<O>((int)rintf(g) + J.shift[1], J.minv[1], J.maxv[1]);
                    o[2] = clamp_out<O>((int)rintf(b) + J.shift[2], J.minv[2], J.maxv[2]);
                } else {
                    // inverse RCT, T.800 G.2 (TC/preprocess_rct_inverse.cl:18-36)
                    const int y0 = (int)x[0][H + j], y1 = (int)x[NC > 1 ? 1 : 0][H + j],
                              y2 = (int)x[NC > 2 ? 2 : 0][H + j];
                    const int g = y0 - ((y1 + y2) >> 2);
                    o[0] = clamp_out<O>(y2 + g + J.shift[0], J.minv[0], J.maxv[0]);
                    o[1] = clamp_out<O>(g + J.shift[1], J.minv[1], J.maxv[1]);
                    o[2] = clamp_out<O>(y1 + g + J.shift[2], J.minv[2], J.maxv[2]);
                }
            } else {
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    const int iv = IS97 ? (int)rintf((float)x[c][H + j]) : (int)x[c][H + j];
                    o[c] = clamp_out<O>(iv + J.shift[c], J.minv[c], J.maxv[c]);
                }
            }
        }
    };
    typedef std::integral_constant<bool, true> yes;
    typedef std::integral_constant<bool, false> no;
    if (NC == 3 && J.mct) {
        if (nrows >= KC)
            emit(std::integral_constant<bool, NC == 3>(), yes());
        else
            emit(std::integral_constant<bool, NC == 3>(), no());
    } else {
        if (nrows >= KC)
            emit(no(), yes());
        else
            emit(no(), no());
    }
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
    constexpr int KC = col_seg(TX, TY);
    __shared__ T tile[NC * TS];
    const uint32_t wg = xcd_run_index();
    const tc_final_job& J = jobs[wg_job[wg]];
    const int t = (int)wg - J.wg_base;
    const int ox = J.wx0 + (t % J.tiles_x) * TX, oy = J.wy0 + (t / J.tiles_x) * TY;
    const bool copy = J.levels == 0;
    const bool h1 = copy || (J.u1 - J.u0) == 1, v1x = copy || (J.v1 - J.v0) == 1;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const Src s{J.ll_off[c], J.hl_off[c], J.lh_off[c], J.hh_off[c], J.ll_pitch[c], J.band_pitch[c]};
        tile_load<IS97, TX, TY>(tile + c * TS, PITCH, base, s, J.u0, J.u1, J.v0, J.v1, ox, oy, copy);
    }
    __syncthreads();
    // every component's rows in the same rounds
    if (!h1) {
        if ((ox - H) & 1) {
            asm volatile("; odd window columns");
            rows_pass<IS97, TX, TY, NC, 1>(tile, PITCH, TS);
        } else {
            rows_pass<IS97, TX, TY, NC, 0>(tile, PITCH, TS);
        }
    }
    const int col = threadIdx.x % TX, seg = threadIdx.x / TX;
    if (seg >= TY / KC) return;
    // each parity form runs its column pass and its stores on its own: with a
    // join between them the compiler merges the two register arrays through
    // dynamically indexed moves
    if ((oy - H) & 1) {
        asm volatile("; odd window rows");
        final_columns<IS97, NC, O, TX, TY, 1>(J, tile, ox, oy, col, seg, v1x, outimg);
    } else {
        final_columns<IS97, NC, O, TX, TY, 0>(J, tile, ox, oy, col, seg, v1x, outimg);
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
