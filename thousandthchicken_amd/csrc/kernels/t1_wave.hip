// Tier-1 decoding with one wavefront per code-block, for launches with few
// code-blocks (one image at a time), where the serial MQ symbol chain of each
// block bounds the run time.
//
// Reference behaviour as in t1_lockstep.hip (g_decode, TC/coefficient_coder.cl:
// 778-819 with the repaired setters and the pass count honoured).
//
// Mapping.  The MQ registers, the position in the block and every decision
// are wave-uniform, so the decoder runs on the scalar unit.  Block state that
// needs indexing by a value -- the 19 context states, the per-row bit masks
// (significance, sign, coded-this-plane, refined, and the current plane's
// 1-bits), the zero-coding/sign-coding tables, the MQ transition table and a
// 256-byte window of the codeword -- lives in vector registers used as
// 64-entry arrays (lane i holds entry i): a read is one v_readlane, a write one
// v_writelane, with no LDS or global round trip on the symbol path.  Each
// plane's 1-bit masks go to LDS when the plane ends; at the end of the block
// the 64 lanes rebuild and dequantize whole rows and store them once.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../tc_kernels.h"
#include "t1_common.hpp"

namespace tc {
namespace {

__device__ __forceinline__ uint32_t rl(uint32_t v, uint32_t i) { return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)i); }
// write of entry i: a lane compare and a select (v_cmp + v_cndmask)
__device__ __forceinline__ uint32_t wl(uint32_t v, uint32_t i, uint32_t val) { return threadIdx.x == i ? val : v; }

// A 64-entry array of row masks, entry y in lane y.
template <typename M>
struct Rows;
template <>
struct Rows<uint32_t> {
    uint32_t v = 0;
    __device__ __forceinline__ uint32_t get(uint32_t y) const { return rl(v, y); }
    __device__ __forceinline__ void set(uint32_t y, uint32_t m) { v = wl(v, y, m); }
    __device__ __forceinline__ uint64_t mine() const { return v; }
};
template <>
struct Rows<uint64_t> {
    uint32_t lo = 0, hi = 0;
    __device__ __forceinline__ uint64_t get(uint32_t y) const { return (uint64_t)rl(lo, y) | ((uint64_t)rl(hi, y) << 32); }
    __device__ __forceinline__ void set(uint32_t y, uint64_t m) {
        lo = wl(lo, y, (uint32_t)m);
        hi = wl(hi, y, (uint32_t)(m >> 32));
    }
    __device__ __forceinline__ uint64_t mine() const { return (uint64_t)lo | ((uint64_t)hi << 32); }
};

struct MqS {
    uint32_t A, C;
    int CT;
    uint32_t bp, len, cur;
    uint32_t win;    // lane j: codeword bytes [wbase + 4j, wbase + 4j + 4)
    uint32_t wbase;  // first byte of the window (multiple of 4)
    const uint32_t* words;  // codeword start, 4-byte aligned
    uint32_t nwords;        // dwords readable from words
};

__device__ __forceinline__ uint32_t wbyte(MqS& m, uint32_t i) {
    if (i >= m.len) return 0xFFu;
    uint32_t rel = i - m.wbase;
    if (rel >= 256u) {
        m.wbase = i & ~3u;
        uint32_t widx = (m.wbase >> 2) + threadIdx.x;
        m.win = widx < m.nwords ? m.words[widx] : 0xFFFFFFFFu;
        rel = i - m.wbase;
    }
    return (rl(m.win, rel >> 2) >> ((rel & 3u) * 8u)) & 0xFFu;
}

__device__ __forceinline__ void mqs_bytein(MqS& m) {
    if (m.cur == 0xFF) {
        uint32_t b1 = wbyte(m, m.bp + 1);
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
        m.cur = wbyte(m, m.bp);
        m.C += m.cur << 8;
        m.CT = 8;
    }
}

__device__ __forceinline__ uint32_t mqs_decode(MqS& m, uint32_t& st, uint32_t trlo, uint32_t trhi) {
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
    uint32_t I = (st >> 16) & 63;
    uint32_t d;
    if (take_mps) {
        d = mps;
        st = (rl(trlo, I) & 0x3FFFFF) | (mps << 22);
    } else {
        d = mps ^ 1;
        uint32_t e2 = rl(trhi, I);
        mps ^= (e2 >> 22) & 1;
        st = (e2 & 0x3FFFFF) | (mps << 22);
    }
    if (lps_path) m.A = qe;
    int n = __clz(m.A) - 16;
    do {
        if (m.CT == 0) mqs_bytein(m);
        int s = n < m.CT ? n : m.CT;
        m.A <<= s;
        m.C <<= s;
        m.CT -= s;
        n -= s;
    } while (n > 0);
    return d;
}

// Block parameters, every one wave-uniform (held in scalar registers).
struct BlockArgs {
    uint32_t data_off, data_len;
    uint64_t out_off;
    uint32_t out_pitch;
    int w, h, numbps, npasses, mode;
    float delta;
};

__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

__device__ __forceinline__ BlockArgs load_args(const tc_cblk_desc* __restrict__ dp, uint32_t* cls) {
    const uint32_t* q = reinterpret_cast<const uint32_t*>(dp);  // 8 dwords, layout of tc_cblk_desc
    BlockArgs a;
    a.data_off = uni(q[0]);
    a.data_len = uni(q[1]);
    a.out_off = (uint64_t)uni(q[2]) | ((uint64_t)uni(q[3]) << 32);
    a.out_pitch = uni(q[4]);
    uint32_t wh = uni(q[5]), pk = uni(q[6]);
    a.w = (int)(wh & 0xFFFF);
    a.h = (int)(wh >> 16);
    a.numbps = (int)(pk & 0xFF);
    a.npasses = (int)((pk >> 8) & 0xFF);
    *cls = (pk >> 16) & 0xFF;
    a.mode = (int)((pk >> 24) & 0xFF);
    a.delta = __uint_as_float(uni(q[7]));
    return a;
}

// 3-bit window of columns x-1, x, x+1 of row mask m (x-1 in bit 0).
template <typename M>
__device__ __forceinline__ uint32_t w3(M m, int x) {
    return (uint32_t)((x == 0 ? (m << 1) : (m >> (x - 1))) & (M)7);
}

template <typename M>
__device__ __forceinline__ uint32_t bit_of(M m, int x) {
    return (uint32_t)(m >> x) & 1u;
}

// Scalar decoder state of one block (all fields wave-uniform except the
// lane-indexed arrays).
template <typename M>
struct WaveBlock {
    Rows<M> sig, sgn, pip, ref, ones, sig0;  // sig0: significance when the plane began
    uint32_t ctxv;                           // context states, lane = label 0..16
    uint32_t st_run, st_uni;
    MqS m;
    uint32_t zcv, scv, trlo, trhi;
    int w, h;
    M VM;

    __device__ __forceinline__ uint32_t zc_of(uint32_t nb9) const {
        return (rl(zcv, nb9 >> 3) >> ((nb9 & 7u) * 4u)) & 15u;
    }
    __device__ __forceinline__ uint32_t dec(uint32_t lab) {
        uint32_t st = rl(ctxv, lab);
        uint32_t b = mqs_decode(m, st, trlo, trhi);
        ctxv = wl(ctxv, lab, st);
        return b;
    }
    // per-column significance windows of rows y0-1..y0+4 (lane = column):
    // bits 3i..3i+2 hold columns x-1, x, x+1 of stripe row i-1
    __device__ __forceinline__ uint32_t slices(const M (&S)[6]) const {
        const int x = threadIdx.x;
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 6; i++) v |= w3<M>(S[i], x) << (3 * i);
        return v;
    }
    __device__ __forceinline__ void load_sig(int y0, int nr, M (&S)[6]) const {
        S[0] = y0 > 0 ? (M)sig.get(y0 - 1) : (M)0;
        S[5] = y0 + 4 < h ? (M)sig.get(y0 + 4) : (M)0;
#pragma unroll
        for (int r = 0; r < 4; r++) S[r + 1] = r < nr ? (M)sig.get(y0 + r) : (M)0;
    }
    // decodes the sign of (x, stripe row r) and marks it significant
    __device__ __forceinline__ void make_sig(int x, int r, uint32_t& sl, uint32_t& carry, M (&S)[6], M (&X)[4],
                                             M xtop, M xbot) {
        constexpr int NB = (int)sizeof(M) * 8;
        const M bit = (M)1 << x;
        const uint32_t sh = 3 * r;
        uint32_t xl = 0, xr = 0, xt = 0, xb = 0;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (q == r) {
                xl = x > 0 ? bit_of(X[q], x - 1) : 0u;
                xr = x + 1 < NB ? bit_of(X[q], x + 1) : 0u;
                xt = q == 0 ? bit_of(xtop, x) : bit_of(X[q > 0 ? q - 1 : 0], x);
                xb = q == 3 ? bit_of(xbot, x) : bit_of(X[q < 3 ? q + 1 : 3], x);
            }
        uint32_t si = ((sl >> (sh + 3)) & 1u) | (xl << 1) | (((sl >> (sh + 5)) & 1u) << 2) | (xr << 3) |
                      (((sl >> (sh + 1)) & 1u) << 4) | (xt << 5) | (((sl >> (sh + 7)) & 1u) << 6) | (xb << 7);
        uint32_t e = (rl(scv, si >> 2) >> ((si & 3u) * 8u)) & 0xFFu;
        uint32_t sgv = dec(e & 15u) ^ (e >> 4);
        sl |= 1u << (sh + 4);     // own centre bit
        carry |= 1u << (sh + 3);  // left-neighbour bit of column x+1
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (q == r) {
                S[q + 1] |= bit;
                if (sgv) X[q] |= bit;
            }
    }

    // significance propagation over one stripe (T.800 D.3.1)
    __device__ __forceinline__ void spp(int y0, int nr) {
        constexpr int NB = (int)sizeof(M) * 8;
        M S[6];
        load_sig(y0, nr, S);
        M cand = 0;
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (r < nr) cand |= ~S[r + 1] & spread<M>(S[r] | S[r + 1] | S[r + 2]);
        cand &= VM;
        if (!cand) return;  // pi stays clear
        M X[4], P[4] = {0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 4; r++) X[r] = r < nr ? (M)sgn.get(y0 + r) : (M)0;
        const M xtop = y0 > 0 ? (M)sgn.get(y0 - 1) : (M)0;
        const M xbot = y0 + 4 < h ? (M)sgn.get(y0 + 4) : (M)0;
        const uint32_t slice = slices(S);
        int x = -1;
        uint32_t carry = 0;
        for (;;) {
            int nx;
            if (carry) {
                nx = x + 1;
            } else {
                M rem = (x < 0) ? cand : ((x + 1 >= NB) ? (M)0 : (cand & (~(M)0 << (x + 1))));
                if (!rem) break;
                nx = (int)__builtin_ctzll((unsigned long long)rem);
            }
            uint32_t sl = rl(slice, nx) | (nx == x + 1 ? carry : 0u);
            x = nx;
            carry = 0;
            const M bit = (M)1 << x;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (r < nr && !((sl >> (3 * r + 4)) & 1u)) {
                    uint32_t nb = (sl >> (3 * r)) & 0x1FFu;
                    if (nb & 0x1EFu) {
                        if (dec(zc_of(nb))) make_sig(x, r, sl, carry, S, X, xtop, xbot);
                        P[r] |= bit;
                    }
                }
            }
            if (x + 1 >= w) carry = 0;
        }
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (r < nr) {
                sig.set(y0 + r, S[r + 1]);
                sgn.set(y0 + r, X[r]);
                pip.set(y0 + r, P[r]);
            }
    }

    // magnitude refinement over one stripe (T.800 D.3.3)
    __device__ __forceinline__ void mrp(int y0, int nr) {
        M S[6];
        load_sig(y0, nr, S);
        M P[4], cand = 0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            P[r] = r < nr ? (M)pip.get(y0 + r) : (M)0;
            cand |= S[r + 1] & ~P[r];
        }
        cand &= VM;
        if (!cand) return;
        M R[4], O[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            R[r] = r < nr ? (M)ref.get(y0 + r) : (M)0;
            O[r] = 0;
        }
        const uint32_t slice = slices(S);
        while (cand) {
            int x = (int)__builtin_ctzll((unsigned long long)cand);
            const M bit = (M)1 << x;
            cand &= cand - 1;
            uint32_t sl = rl(slice, x);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if ((S[r + 1] & bit) && !(P[r] & bit)) {
                    uint32_t lab = (R[r] & bit) ? 16u : (((sl >> (3 * r)) & 0x1EFu) ? 15u : 14u);
                    if (dec(lab)) O[r] |= bit;
                    R[r] |= bit;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (r < nr) {
                ref.set(y0 + r, R[r]);
                if (O[r]) ones.set(y0 + r, (M)ones.get(y0 + r) | O[r]);
            }
    }

    // cleanup with run-length mode over one stripe (T.800 D.3.4)
    __device__ __forceinline__ void cup(int y0, int nr) {
        M S[6];
        load_sig(y0, nr, S);
        M P[4], cand = 0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            P[r] = r < nr ? (M)pip.get(y0 + r) : (M)0;
            if (r < nr) cand |= ~S[r + 1] & ~P[r];
        }
        cand &= VM;
        M Pany = P[0] | P[1] | P[2] | P[3];
        if (cand) {
            M X[4];
#pragma unroll
            for (int r = 0; r < 4; r++) X[r] = r < nr ? (M)sgn.get(y0 + r) : (M)0;
            const M xtop = y0 > 0 ? (M)sgn.get(y0 - 1) : (M)0;
            const M xbot = y0 + 4 < h ? (M)sgn.get(y0 + 4) : (M)0;
            const uint32_t slice = slices(S);
            int lastx = -2;
            uint32_t carry = 0;
            while (cand) {
                int x = (int)__builtin_ctzll((unsigned long long)cand);
                const M bit = (M)1 << x;
                cand &= cand - 1;
                uint32_t sl = rl(slice, x) | (x == lastx + 1 ? carry : 0u);
                lastx = x;
                carry = 0;
                int r0 = 0;
                if (nr == 4 && sl == 0 && !(Pany & bit)) {
                    if (!mqs_decode(m, st_run, trlo, trhi)) continue;
                    uint32_t rr = mqs_decode(m, st_uni, trlo, trhi) << 1;
                    rr |= mqs_decode(m, st_uni, trlo, trhi);
                    make_sig(x, (int)rr, sl, carry, S, X, xtop, xbot);
                    r0 = (int)rr + 1;
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    if (r >= r0 && r < nr && !((sl >> (3 * r + 4)) & 1u) && !(P[r] & bit)) {
                        if (dec(zc_of((sl >> (3 * r)) & 0x1FFu))) make_sig(x, r, sl, carry, S, X, xtop, xbot);
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (r < nr) {
                    sig.set(y0 + r, S[r + 1]);
                    sgn.set(y0 + r, X[r]);
                }
        }
        if (Pany) {
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (r < nr) pip.set(y0 + r, (M)0);
        }
    }
};

template <typename M>
__device__ __forceinline__ void wave_decode_block(const BlockArgs d, const uint32_t* __restrict__ arena_w, uint32_t arena_words,
                                  uint32_t* __restrict__ coef, uint64_t (*planes)[64], uint64_t* sgnrows,
                                  uint32_t zcv, uint32_t scv, uint32_t trlo, uint32_t trhi) {
    const int lane = threadIdx.x;
    const int w = d.w, h = d.h;
    const int numbps = d.numbps, npasses = d.npasses;
    constexpr int NB = (int)sizeof(M) * 8;

    WaveBlock<M> B;
    B.w = w;
    B.h = h;
    B.VM = (w >= NB) ? ~(M)0 : (((M)1 << w) - 1);
    B.ctxv = lane < 17 ? pack_state(lane == 0 ? 4 : 0, 0) : 0u;
    B.st_run = pack_state(3, 0);
    B.st_uni = pack_state(46, 0);
    B.zcv = zcv;
    B.scv = scv;
    B.trlo = trlo;
    B.trhi = trhi;
    MqS& m = B.m;
    m.words = arena_w + (d.data_off >> 2);
    m.nwords = arena_words - (d.data_off >> 2);
    m.len = d.data_len;
    m.wbase = 0;
    m.win = lane < (int)m.nwords ? m.words[lane] : 0xFFFFFFFFu;
    m.bp = 0;
    m.cur = wbyte(m, 0);
    m.C = m.cur << 16;
    mqs_bytein(m);
    m.C <<= 7;
    m.CT -= 7;
    m.A = 0x8000;

    for (int k = 0; k < npasses; k++) {
        const int p = numbps - 1 - (k + 2) / 3;
        if (p < 0) break;
        const int type = k == 0 ? 2 : (k - 1) % 3;  // 0 sigprop, 1 magref, 2 cleanup
        if (type == 0 || k == 0) B.sig0 = B.sig;   // a plane begins
        for (int y0 = 0; y0 < h; y0 += 4) {
            const int nr = h - y0 < 4 ? h - y0 : 4;
            if (type == 0)
                B.spp(y0, nr);
            else if (type == 1)
                B.mrp(y0, nr);
            else
                B.cup(y0, nr);
        }
        // a plane ends after its cleanup pass, or when the passes run out:
        // its 1-bits are the refinement 1-bits plus every new significance
        if (type == 2 || k == npasses - 1) {
            planes[p][lane] = (uint64_t)(B.ones.mine() | (B.sig.mine() & ~B.sig0.mine()));
            B.ones = Rows<M>();
        }
    }
    sgnrows[lane] = B.sgn.mine();
    __syncthreads();

    // rebuild, dequantize and store the block's rows
    int kl = npasses - 1;
    int plast = numbps - 1 - (kl + 2) / 3;
    if (plast < 0) plast = 0;
    int tlast = kl == 0 ? 2 : (kl - 1) % 3;
    const uint32_t sigonly = 1u << plast;
    int span = 1;
    while (span < w) span <<= 1;
    int rows_per = 64 / span;
    int lx = lane % span, ly = lane / span;
    uint32_t* out = coef + d.out_off;
    if (lx < w) {
        for (int y = ly; y < h; y += rows_per) {
            uint32_t mag = 0;
            for (int pp = plast; pp < numbps; pp++) mag |= (uint32_t)((planes[pp][y] >> lx) & 1u) << pp;
            uint32_t v = 0;
            if (mag) {
                bool neg = ((sgnrows[y] >> lx) & 1u) != 0;
                int q = (tlast == 0 && mag != sigonly) ? plast + 1 : plast;
                if ((d.mode & TC_MODE_IRREVERSIBLE) == 0) {
                    int32_t iv = (int32_t)(mag + (q > 0 ? (1u << (q - 1)) : 0u));
                    v = (uint32_t)(neg ? -iv : iv);
                } else {
                    float f = ((float)mag + ldexpf(0.5f, q)) * d.delta;
                    v = __float_as_uint(neg ? -f : f);
                }
            }
            out[(uint32_t)y * d.out_pitch + lx] = v;
        }
    }
    __syncthreads();
}

template <typename M>
__global__ __launch_bounds__(64) void t1_wave_kernel(const tc_cblk_desc* __restrict__ descs,
                                                    const uint32_t* __restrict__ order, int n,
                                                    const uint32_t* __restrict__ arena_w, uint32_t arena_words,
                                                    uint32_t* __restrict__ coef) {
    __shared__ uint64_t planes[32][64];
    __shared__ uint64_t sgnrows[64];
    const int lane = threadIdx.x;
    // lane-indexed tables: zero-coding labels of 4 patterns per lane (per class),
    // sign entries of 4 patterns per lane, MQ successor states
    uint32_t zc3[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        uint32_t v = 0;
        for (int b = 0; b < 8; b++) {
            uint32_t i9 = (uint32_t)(lane * 8 + b);
            uint32_t n8 = (i9 & 0xFu) | ((i9 >> 5) << 4);  // drop the centre bit
            v |= zc_label(c, n8) << (4 * b);
        }
        zc3[c] = v;
    }
    uint32_t scv = 0;
    for (int b = 0; b < 4; b++) scv |= sc_entry((uint32_t)(lane * 4 + b)) << (8 * b);
    uint64_t te = lane < 47 ? trans_entry(lane) : 0;
    uint32_t trlo = (uint32_t)te, trhi = (uint32_t)(te >> 32);
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        uint32_t id = order ? uni(order[b]) : (uint32_t)b;
        uint32_t cls;
        const BlockArgs d = load_args(descs + id, &cls);
        if (d.numbps <= 0 || d.npasses <= 0) continue;
        uint32_t zcv = cls == 0 ? zc3[0] : (cls == 1 ? zc3[1] : zc3[2]);
        wave_decode_block<M>(d, arena_w, arena_words, coef, planes, sgnrows, zcv, scv, trlo, trhi);
    }
}

}  // namespace

hipError_t launch_t1_wave(hipStream_t s, bool narrow, int grid, const tc_cblk_desc* descs, const uint32_t* order,
                          int n, const uint8_t* arena, uint32_t arena_bytes, uint32_t* coef) {
    if (n <= 0) return hipSuccess;
    if (grid > n) grid = n;
    if (narrow)
        hipLaunchKernelGGL(t1_wave_kernel<uint32_t>, dim3(grid), dim3(64), 0, s, descs, order, n,
                           (const uint32_t*)arena, arena_bytes / 4, coef);
    else
        hipLaunchKernelGGL(t1_wave_kernel<uint64_t>, dim3(grid), dim3(64), 0, s, descs, order, n,
                           (const uint32_t*)arena, arena_bytes / 4, coef);
    return hipGetLastError();
}

}  // namespace tc
