// Device helpers shared by the tier-1 kernels: MQ tables (T.800 Table C.2),
// packed context states, the zero-coding and sign-coding label functions
// (T.800 Tables D.1-D.3) and bit-mask neighbourhood extraction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tc {
namespace {

// MQ probability estimation, T.800 Table C.2 (same literals as
// TC/coefficient_coder.cl:36-60).
__constant__ uint16_t kQe[47] = {
    0x5601, 0x3401, 0x1801, 0x0AC1, 0x0521, 0x0221, 0x5601, 0x5401, 0x4801, 0x3801,
    0x3001, 0x2401, 0x1C01, 0x1601, 0x5601, 0x5401, 0x5101, 0x4801, 0x3801, 0x3401,
    0x3001, 0x2801, 0x2401, 0x2201, 0x1C01, 0x1801, 0x1601, 0x1401, 0x1201, 0x1101,
    0x0AC1, 0x09C1, 0x08A1, 0x0521, 0x0441, 0x02A1, 0x0221, 0x0141, 0x0111, 0x0085,
    0x0049, 0x0025, 0x0015, 0x0009, 0x0005, 0x0001, 0x5601};
__constant__ uint8_t kNmps[47] = {1,  2,  3,  4,  5,  38, 7,  8,  9,  10, 11, 12, 13, 29, 15, 16,
                                  17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32,
                                  33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 45, 46};
__constant__ uint8_t kNlps[47] = {1,  6,  9,  12, 29, 33, 6,  14, 14, 14, 17, 18, 20, 21, 14, 14,
                                  15, 16, 17, 18, 19, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29,
                                  30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 46};

// A context state is packed as Qe (bits 0-15) | index I (16-21) | MPS (22), so
// the common no-renormalisation path needs no table lookup.
__device__ __forceinline__ uint32_t pack_state(uint32_t I, uint32_t mps) {
    return (uint32_t)kQe[I] | (I << 16) | (mps << 22);
}

// Zero-coding label from the 8-neighbour significance pattern
// (bit order tl,t,tr,l,r,bl,b,br), T.800 Table D.1.
__device__ uint32_t zc_label(int cls, uint32_t n) {
    int h = ((n >> 3) & 1) + ((n >> 4) & 1);
    int v = ((n >> 1) & 1) + ((n >> 6) & 1);
    int d = (n & 1) + ((n >> 2) & 1) + ((n >> 5) & 1) + ((n >> 7) & 1);
    if (cls == 1) {
        int t = h;
        h = v;
        v = t;
    }
    if (cls < 2) {
        if (h == 2) return 8;
        if (h == 1) return v ? 7 : (d ? 6 : 5);
        if (v == 2) return 4;
        if (v == 1) return 3;
        return d >= 2 ? 2 : d;
    }
    int hv = h + v;
    if (d >= 3) return 8;
    if (d == 2) return hv ? 7 : 6;
    if (d == 1) return hv >= 2 ? 5 : (hv ? 4 : 3);
    return hv >= 2 ? 2 : hv;
}

// Sign-coding label (9..13) in bits 0-3 and the XOR bit in bit 4, T.800
// Tables D.2/D.3.  Index bits: sig_l, sgn_l, sig_r, sgn_r, sig_u, sgn_u,
// sig_d, sgn_d.
__device__ uint32_t sc_entry(uint32_t i) {
    int l = (i & 1) ? ((i & 2) ? -1 : 1) : 0;
    int r = (i & 4) ? ((i & 8) ? -1 : 1) : 0;
    int u = (i & 16) ? ((i & 32) ? -1 : 1) : 0;
    int d = (i & 64) ? ((i & 128) ? -1 : 1) : 0;
    int H = l + r, V = u + d;
    H = H > 0 ? 1 : (H < 0 ? -1 : 0);
    V = V > 0 ? 1 : (V < 0 ? -1 : 0);
    uint32_t x = 0;
    if (H < 0) {
        H = -H;
        V = -V;
        x = 1;
    }
    if (H == 0 && V < 0) {
        V = 1;
        x = 1;
    }
    uint32_t lab = H ? (V > 0 ? 13 : (V == 0 ? 12 : 11)) : (V ? 10 : 9);
    return lab | (x << 4);
}

template <typename M>
__device__ __forceinline__ M spread(M m) {
    return m | (m << 1) | (m >> 1);
}

// Bits x-1, x, x+1 of m as a 3-bit value (x-1 in bit 0).
template <typename M>
__device__ __forceinline__ uint32_t win3(M m, int x) {
    uint32_t lo = x > 0 ? (uint32_t)(m >> (x - 1)) & 1u : 0u;
    return lo | (((uint32_t)(m >> x) & 3u) << 1);
}

template <typename M>
__device__ __forceinline__ uint32_t nbh8(M up, M row, M dn, int x) {
    uint32_t a = win3(up, x), b = win3(row, x), c = win3(dn, x);
    return a | ((b & 1u) << 3) | ((b >> 2) << 4) | (c << 5);
}

template <typename M>
__device__ __forceinline__ uint32_t bitat(M m, int x) {
    return (uint32_t)(m >> x) & 1u;
}

// MPS/LPS successor of state i, packed (Qe|I) in the low / high word, the high
// word carrying SWITCH in bit 22.
__device__ __forceinline__ uint64_t trans_entry(int i) {
    uint64_t lo = pack_state(kNmps[i], 0) & 0x3FFFFF;
    uint64_t hi = (pack_state(kNlps[i], 0) & 0x3FFFFF) | ((i == 0 || i == 6 || i == 14) ? (1u << 22) : 0u);
    return lo | (hi << 32);
}

// Converts one decoded block from sign-magnitude to its coefficient values.
// Midpoint reconstruction follows the usual decoder convention: a significant
// coefficient is rebuilt at mag + 2^(q-1) (reversible, integer) or
// (mag + 0.5*2^q)*Delta_b (irreversible), q being the lowest bit-plane coded
// for it.  q is the plane of the last pass, one higher for coefficients that
// were significant before a final significance-propagation pass and so were
// not refined in it.
__device__ void convert_block(int lane, const tc_cblk_desc& d, uint32_t* __restrict__ coef) {
    int numbps = d.numbps, np = d.npasses;
    if (numbps <= 0 || np <= 0) return;
    int kl = np - 1;
    int plast = numbps - 1 - (kl + 2) / 3;
    if (plast < 0) plast = 0;
    int tlast = kl == 0 ? 2 : (kl - 1) % 3;
    uint32_t* out = coef + d.out_off;
    const int w = d.w, h = d.h;
    const uint32_t sigonly = 1u << plast;
    // the 64 lanes cover 64/span rows per step, span = w rounded up to a power of two
    int span = 1;
    while (span < w) span <<= 1;
    int rows_per = 64 / span;
    int lx = lane % span, ly = lane / span;
    if (lx >= w) return;
    for (int y = ly; y < h; y += rows_per) {
        uint32_t* p = out + (uint32_t)y * d.out_pitch + lx;
        uint32_t v = *p;
        if (!v) continue;
        uint32_t mag = v & 0x7FFFFFFFu;
        int q = (tlast == 0 && mag != sigonly) ? plast + 1 : plast;
        bool neg = (v >> 31) != 0;
        if ((d.mode & TC_MODE_IRREVERSIBLE) == 0) {
            int32_t iv = (int32_t)(mag + (q > 0 ? (1u << (q - 1)) : 0u));
            *p = (uint32_t)(neg ? -iv : iv);
        } else {
            float f = ((float)mag + ldexpf(0.5f, q)) * d.delta;
            *p = __float_as_uint(neg ? -f : f);
        }
    }
}

}  // namespace
}  // namespace tc
