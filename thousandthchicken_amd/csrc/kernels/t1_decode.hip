// Tier-1 EBCOT decoding (MQ arithmetic decoder + significance propagation,
// magnitude refinement and cleanup passes) for CDNA4, with the reversible
// shift / irreversible dequantization fused into the final store.
//
// Reference behaviour: g_decode, TC/coefficient_coder.cl:778-819 (pass order),
// the MQ decoder TC/coefficient_coder.cl:105-252, context formation
// :330-469 and the three pass functors :471-586, with the by-value no-op
// setters repaired and the pass count honoured (ITU-T T.800 Annex C, D).
//
// Mapping.  One work-group is one 64-lane wavefront.  Each active lane decodes
// one code-block ("G blocks per wave"; G is the packing the host picks), so a
// wave advances G independent MQ streams per instruction.  Per-block state is
// held as bit masks over the columns of one row (bit x = column x):
//   sig (significance), sgn (sign), pi (coded in the current plane's
//   significance pass) and ref (refined at least once).
// The masks of all rows live in LDS laid out [row][lane], so lanes never
// conflict on a bank; while a lane works on a 4-row stripe it keeps the six
// significance/sign rows it needs (stripe +-1 row) in registers and finds the
// next column to visit with bit arithmetic (find-first-set over candidate
// masks) instead of walking every coefficient.  Magnitude and sign bits are
// posted to the zero-filled coefficient plane with atomic ORs (no return), and
// after the decode the wave converts its blocks' sign-magnitude words to
// two's complement integers (5/3) or dequantized floats (9/7) in place, reading
// whole rows across the 64 lanes.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../tc_kernels.h"
#include "t1_common.hpp"

namespace tc {
namespace {

struct Mq {
    uint32_t A, C;
    int CT;
    uint32_t bp, len, cur;
    const uint8_t* data;
};

__device__ __forceinline__ uint32_t mq_byte(const Mq& m, uint32_t i) {
    return i < m.len ? (uint32_t)m.data[i] : 0xFFu;
}

// BYTEIN, T.800 Figure C.19; past the codeword end the decoder sees 0xFF 0xFF.
__device__ __forceinline__ void mq_bytein(Mq& m) {
    if (m.cur == 0xFF) {
        uint32_t b1 = mq_byte(m, m.bp + 1);
        if (b1 > 0x8F) {
            m.C += 0xFF00;
            m.CT = 8;
        } else {
            m.bp++;
            m.cur = b1;
            m.C += b1 << 9;
            m.CT = 7;
        }
    } else {
        m.bp++;
        m.cur = mq_byte(m, m.bp);
        m.C += m.cur << 8;
        m.CT = 8;
    }
}

__device__ __forceinline__ void mq_init(Mq& m, const uint8_t* data, uint32_t len) {
    m.data = data;
    m.len = len;
    m.bp = 0;
    m.cur = mq_byte(m, 0);
    m.C = m.cur << 16;
    mq_bytein(m);
    m.C <<= 7;
    m.CT -= 7;
    m.A = 0x8000;
}

// DECODE (T.800 Figures C.15-C.18).  trans[I] holds, for state I, the packed
// successor states (Qe|I) after an MPS (low word) and after an LPS (high word,
// bit 22 = SWITCH).  RENORMD shifts by whole runs up to the next BYTEIN.
__device__ __forceinline__ uint32_t mq_decode(Mq& m, uint32_t& st, const uint64_t* trans) {
    uint32_t qe = st & 0xFFFF;
    uint32_t mps = (st >> 22) & 1;
    m.A -= qe;
    uint32_t chi = m.C >> 16;
    if (chi >= qe && (m.A & 0x8000)) {
        m.C -= qe << 16;
        return mps;
    }
    bool lps_path = chi < qe;
    if (!lps_path) m.C -= qe << 16;
    bool small = m.A < qe;
    bool take_mps = lps_path ? small : !small;
    uint64_t e = trans[(st >> 16) & 63];
    uint32_t d;
    if (take_mps) {
        d = mps;
        st = ((uint32_t)e & 0x3FFFFF) | (mps << 22);
    } else {
        d = mps ^ 1;
        uint32_t e2 = (uint32_t)(e >> 32);
        mps ^= (e2 >> 22) & 1;
        st = (e2 & 0x3FFFFF) | (mps << 22);
    }
    if (lps_path) m.A = qe;
    int n = __clz(m.A) - 16;
    do {
        if (m.CT == 0) mq_bytein(m);
        int s = n < m.CT ? n : m.CT;
        m.A <<= s;
        m.C <<= s;
        m.CT -= s;
        n -= s;
    } while (n > 0);
    return d;
}

constexpr int kNumLdsCtx = 17;  // labels 0..16 in LDS; run-length and uniform in registers

template <typename M, int MAXH, int G>
struct Lds {
    uint64_t trans[48];
    uint8_t zc[3][256];
    uint8_t sc[256];
    uint32_t ctx[kNumLdsCtx][G];
    M rows[4][MAXH][G];  // sig, sgn, pi, ref
};

template <typename M, int MAXH, int G>
__device__ void decode_block(Lds<M, MAXH, G>& L, int lane, const tc_cblk_desc& d,
                             const uint8_t* __restrict__ arena, uint32_t* __restrict__ coef) {
    const int w = d.w, h = d.h;
    const int cls = d.cls;
    const int numbps = d.numbps;
    const int npasses = d.npasses;
    uint32_t* out = coef + d.out_off;
    const uint32_t pitch = d.out_pitch;
    constexpr int NB = (int)sizeof(M) * 8;
    const M VM = (w >= NB) ? ~(M)0 : (((M)1 << w) - 1);

    for (int y = 0; y < h; y++)
        for (int t = 0; t < 4; t++) L.rows[t][y][lane] = 0;
    for (int c = 0; c < kNumLdsCtx; c++) L.ctx[c][lane] = pack_state(c == 0 ? 4 : 0, 0);
    uint32_t st_run = pack_state(3, 0), st_uni = pack_state(46, 0);

    Mq m;
    mq_init(m, arena + d.data_off, d.data_len);
    const uint64_t* trans = L.trans;

    for (int k = 0; k < npasses; k++) {
        const int p = numbps - 1 - (k + 2) / 3;
        if (p < 0) break;
        const int type = k == 0 ? 2 : (k - 1) % 3;  // 0 sigprop, 1 magref, 2 cleanup
        const uint32_t pbit = 1u << p;
        for (int y0 = 0; y0 < h; y0 += 4) {
            M S[6], X[6], P[4], R[4], VR[4];
            S[0] = y0 > 0 ? L.rows[0][y0 - 1][lane] : (M)0;
            X[0] = y0 > 0 ? L.rows[1][y0 - 1][lane] : (M)0;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                bool ok = y0 + r < h;
                VR[r] = ok ? ~(M)0 : (M)0;
                S[r + 1] = ok ? L.rows[0][y0 + r][lane] : (M)0;
                X[r + 1] = ok ? L.rows[1][y0 + r][lane] : (M)0;
                P[r] = ok ? L.rows[2][y0 + r][lane] : (M)0;
                R[r] = ok ? L.rows[3][y0 + r][lane] : (M)0;
            }
            S[5] = y0 + 4 < h ? L.rows[0][y0 + 4][lane] : (M)0;
            X[5] = y0 + 4 < h ? L.rows[1][y0 + 4][lane] : (M)0;

            // Decodes the sign of (x, row r) and marks it significant.
            auto make_sig = [&](int x, int r) {
                M bit = (M)1 << x;
                uint32_t si = 0;
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (q == r) {
                        uint32_t sl = x > 0 ? bitat(S[q + 1], x - 1) : 0u;
                        uint32_t gl = x > 0 ? bitat(X[q + 1], x - 1) : 0u;
                        uint32_t sr = bitat(S[q + 1], x + 1 < NB ? x + 1 : 0) & (x + 1 < NB ? 1u : 0u);
                        uint32_t gr = bitat(X[q + 1], x + 1 < NB ? x + 1 : 0) & (x + 1 < NB ? 1u : 0u);
                        si = sl | (gl << 1) | (sr << 2) | (gr << 3) | (bitat(S[q], x) << 4) |
                             (bitat(X[q], x) << 5) | (bitat(S[q + 2], x) << 6) | (bitat(X[q + 2], x) << 7);
                    }
                uint32_t e = L.sc[si];
                uint32_t lab = e & 15u;
                uint32_t sg = mq_decode(m, L.ctx[lab][lane], trans) ^ (e >> 4);
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (q == r) {
                        S[q + 1] |= bit;
                        if (sg) X[q + 1] |= bit;
                    }
                atomicOr(out + (uint32_t)(y0 + r) * pitch + x, pbit | (sg << 31));
            };

            if (type == 0) {
                // significance propagation (T.800 D.3.1)
                M cand = 0;
#pragma unroll
                for (int r = 0; r < 4; r++) cand |= ~S[r + 1] & spread<M>(S[r] | S[r + 1] | S[r + 2]) & VR[r];
                cand &= VM;
                int x = -1;
                bool forced = false;
                for (;;) {
                    if (!forced) {
                        M rem = (x < 0) ? cand : ((x + 1 >= NB) ? (M)0 : (cand & (~(M)0 << (x + 1))));
                        if (!rem) break;
                        x = (int)__builtin_ctzll((unsigned long long)rem);
                    } else {
                        x = x + 1;
                    }
                    forced = false;
                    M bit = (M)1 << x;
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        if ((VR[r] & bit) && !(S[r + 1] & bit)) {
                            uint32_t nb = nbh8<M>(S[r], S[r + 1], S[r + 2], x);
                            if (nb) {
                                uint32_t lab = L.zc[cls][nb];
                                if (mq_decode(m, L.ctx[lab][lane], trans)) {
                                    make_sig(x, r);
                                    if (x + 1 < w) forced = true;
                                }
                                P[r] |= bit;
                            }
                        }
                    }
                }
            } else if (type == 1) {
                // magnitude refinement (T.800 D.3.3)
                M cand = 0;
#pragma unroll
                for (int r = 0; r < 4; r++) cand |= S[r + 1] & ~P[r] & VR[r];
                while (cand) {
                    int x = (int)__builtin_ctzll((unsigned long long)cand);
                    M bit = (M)1 << x;
                    cand &= ~bit;
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        if ((S[r + 1] & bit) && !(P[r] & bit) && (VR[r] & bit)) {
                            uint32_t lab;
                            if (R[r] & bit)
                                lab = 16;
                            else
                                lab = nbh8<M>(S[r], S[r + 1], S[r + 2], x) ? 15 : 14;
                            if (mq_decode(m, L.ctx[lab][lane], trans))
                                atomicOr(out + (uint32_t)(y0 + r) * pitch + x, pbit);
                            R[r] |= bit;
                        }
                    }
                }
            } else {
                // cleanup with run-length mode (T.800 D.3.4)
                M c0 = ~S[1] & ~P[0] & VR[0] & VM, c1 = ~S[2] & ~P[1] & VR[1] & VM;
                M c2 = ~S[3] & ~P[2] & VR[2] & VM, c3 = ~S[4] & ~P[3] & VR[3] & VM;
                M cand = c0 | c1 | c2 | c3;
                M all4 = c0 & c1 & c2 & c3;  // empty unless the stripe has 4 rows
                while (cand) {
                    int x = (int)__builtin_ctzll((unsigned long long)cand);
                    M bit = (M)1 << x;
                    cand &= ~bit;
                    int r0 = 0;
                    if (all4 & bit) {
                        M any = S[0] | S[1] | S[2] | S[3] | S[4] | S[5];
                        if (!(spread<M>(any) & bit)) {
                            if (!mq_decode(m, st_run, trans)) continue;
                            uint32_t rr = mq_decode(m, st_uni, trans) << 1;
                            rr |= mq_decode(m, st_uni, trans);
                            make_sig(x, (int)rr);
                            r0 = (int)rr + 1;
                        }
                    }
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        M cr = r == 0 ? c0 : (r == 1 ? c1 : (r == 2 ? c2 : c3));
                        if (r >= r0 && (cr & bit)) {
                            uint32_t nb = nbh8<M>(S[r], S[r + 1], S[r + 2], x);
                            uint32_t lab = L.zc[cls][nb];
                            if (mq_decode(m, L.ctx[lab][lane], trans)) make_sig(x, r);
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (y0 + r < h) {
                    L.rows[0][y0 + r][lane] = S[r + 1];
                    L.rows[1][y0 + r][lane] = X[r + 1];
                    L.rows[2][y0 + r][lane] = (type == 2) ? (M)0 : P[r];
                    L.rows[3][y0 + r][lane] = R[r];
                }
            }
        }
    }
}

template <typename M, int MAXH, int G>
__global__ __launch_bounds__(64) void t1_decode_kernel(const tc_cblk_desc* __restrict__ descs,
                                                       const uint32_t* __restrict__ order, int n,
                                                       const uint8_t* __restrict__ arena,
                                                       uint32_t* __restrict__ coef) {
    __shared__ Lds<M, MAXH, G> L;
    const int lane = threadIdx.x;
    for (int i = lane; i < 47; i += 64) L.trans[i] = trans_entry(i);
    for (int i = lane; i < 3 * 256; i += 64) L.zc[i >> 8][i & 255] = (uint8_t)zc_label(i >> 8, i & 255);
    for (int i = lane; i < 256; i += 64) L.sc[i] = (uint8_t)sc_entry(i);
    __syncthreads();

    const int base = blockIdx.x * G;
    if (lane < G && base + lane < n) {
        uint32_t id = order ? order[base + lane] : (uint32_t)(base + lane);
        tc_cblk_desc d = descs[id];
        if (d.numbps > 0 && d.npasses > 0) decode_block<M, MAXH, G>(L, lane, d, arena, coef);
    }
    // make every lane's posted ORs visible, and drop stale L1 lines, before the
    // cooperative read-back
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
    __syncthreads();
    for (int j = 0; j < G && base + j < n; j++) {
        uint32_t id = order ? order[base + j] : (uint32_t)(base + j);
        tc_cblk_desc d = descs[id];
        convert_block(lane, d, coef);
    }
}

__global__ void fill32_kernel(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

}  // namespace

template <typename M, int MAXH, int G>
static size_t t1_lds_bytes() {
    return sizeof(Lds<M, MAXH, G>);
}

template <typename M, int MAXH, int G>
static hipError_t launch_t1_variant(hipStream_t s, const tc_cblk_desc* descs, const uint32_t* order, int n,
                                    const uint8_t* arena, uint32_t* coef) {
    if (n <= 0) return hipSuccess;
    int grid = (n + G - 1) / G;
    hipLaunchKernelGGL((t1_decode_kernel<M, MAXH, G>), dim3(grid), dim3(64), 0, s, descs, order, n, arena, coef);
    return hipGetLastError();
}

// Dispatch over (size class, blocks per wave).  cls32: every block fits 32x32.
hipError_t launch_t1(hipStream_t s, bool cls32, int g, const tc_cblk_desc* descs, const uint32_t* order, int n,
                     const uint8_t* arena, uint32_t* coef) {
    if (cls32) {
        switch (g) {
        case 8: return launch_t1_variant<uint32_t, 32, 8>(s, descs, order, n, arena, coef);
        case 16: return launch_t1_variant<uint32_t, 32, 16>(s, descs, order, n, arena, coef);
        case 32: return launch_t1_variant<uint32_t, 32, 32>(s, descs, order, n, arena, coef);
        case 64: return launch_t1_variant<uint32_t, 32, 64>(s, descs, order, n, arena, coef);
        default: return hipErrorInvalidValue;
        }
    }
    switch (g) {
    case 8: return launch_t1_variant<uint64_t, 64, 8>(s, descs, order, n, arena, coef);
    case 16: return launch_t1_variant<uint64_t, 64, 16>(s, descs, order, n, arena, coef);
    case 32: return launch_t1_variant<uint64_t, 64, 32>(s, descs, order, n, arena, coef);
    case 64: return launch_t1_variant<uint64_t, 64, 64>(s, descs, order, n, arena, coef);
    default: return hipErrorInvalidValue;
    }
}

size_t t1_lds_size(bool cls32, int g) {
    if (cls32) {
        switch (g) {
        case 8: return t1_lds_bytes<uint32_t, 32, 8>();
        case 16: return t1_lds_bytes<uint32_t, 32, 16>();
        case 32: return t1_lds_bytes<uint32_t, 32, 32>();
        case 64: return t1_lds_bytes<uint32_t, 32, 64>();
        default: return 0;
        }
    }
    switch (g) {
    case 8: return t1_lds_bytes<uint64_t, 64, 8>();
    case 16: return t1_lds_bytes<uint64_t, 64, 16>();
    case 32: return t1_lds_bytes<uint64_t, 64, 32>();
    case 64: return t1_lds_bytes<uint64_t, 64, 64>();
    default: return 0;
    }
}

hipError_t launch_fill32(hipStream_t s, uint32_t* p, uint32_t v, uint64_t n) {
    if (!n) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

}  // namespace tc
