// Tier-1 decoding, one wavefront per code-block, stripe-column form.
//
// Reference behaviour: g_decode (TC/coefficient_coder.cl:778-819) with its
// passes (:255-700) as T.800 Annex D defines them, the pass count honoured and
// the MQ decoder of Annex C (C.3).
//
// Mapping.  The MQ decoder and the scan run on the scalar unit: registers,
// context states in use, the run-length and uniform contexts and the loop
// state are wave-uniform.  Block state is column-major in vector registers:
// lane x holds 64-bit masks of column x (bit y = row y) for significance,
// sign, visited-in-this-plane, refined-before and the current plane's 1-bits.
// For each stripe and pass the 64 lanes build, in parallel, per-column words
// (the 3-column x 6-row significance and sign windows, candidate rows,
// zero-coding or refinement labels); one scalar loop written in assembly
// then decodes the whole stripe: it visits only the columns that hold
// candidates (and the column after new significance), reads each column's
// words with v_readlane ahead of use, carries new significance from column x
// to x+1 in scalar registers, and leaves its per-column results in one lane
// write.  After the stripe the lanes fold the results back into the masks.
// The zero-coding and sign-coding tables, the 17 zero/sign/refinement
// context states and the MQ successor table are 64-entry lane arrays.  The
// MQ decoder is one block of scalar instructions (T.800 C.3 semantics,
// bit-exact) inside each loop, so its registers never move between paths;
// one wave pays about 6 cycles per scalar instruction and 14-21 per branch,
// so the loops are laid out for few branches on the common paths.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../tc_kernels.h"
#include "t1_common.hpp"

namespace tc {
namespace {

__device__ int lane_write_i32(int val, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

__device__ __forceinline__ uint32_t rl(uint32_t v, uint32_t i) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)i);
}
__device__ __forceinline__ uint32_t wl(uint32_t v, uint32_t i, uint32_t val) {
    return (uint32_t)lane_write_i32((int)val, (int)i, (int)v);
}
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// context state: Qe << 16 | I << 1 | MPS
__device__ __forceinline__ uint32_t st_pack(uint32_t I, uint32_t mps) {
    return ((uint32_t)kQe[I] << 16) | (I << 1) | mps;
}

struct Mq {
    uint32_t C;      // the C register of C.3
    uint32_t A;      // the A register << 16
    uint32_t CT;     // shifts left before BYTEIN
    uint32_t cur;    // the byte at BP
    uint32_t np;     // BP + 1
    uint32_t wc;     // cached codeword dword
    uint32_t wci;    // its dword index (0xFFFFFFFF: none)
    uint32_t len;    // codeword bytes; later bytes read as 0xFF
    const uint32_t* words;
};

__device__ __forceinline__ uint32_t mq_byte(Mq& m, uint32_t i) {
    if (i >= m.len) return 0xFFu;
    const uint32_t wi = i >> 2;
    if (wi != m.wci) {
        m.wc = m.words[wi];
        m.wci = wi;
    }
    return (m.wc >> ((i & 3u) * 8u)) & 0xFFu;
}

// BYTEIN (C.3.4)
__device__ __forceinline__ void mq_bytein(Mq& m) {
    const uint32_t b1 = mq_byte(m, m.np);
    if (m.cur == 0xFFu) {
        if (b1 > 0x8Fu) {
            m.C += 0xFF00u;
            m.CT = 8;
        } else {
            m.np++;
            m.cur = b1;
            m.C += b1 << 9;
            m.CT = 7;
        }
    } else {
        m.np++;
        m.cur = b1;
        m.C += b1 << 8;
        m.CT = 8;
    }
}

// INITDEC (C.3.5)
__device__ __forceinline__ void mq_init(Mq& m, const uint32_t* words, uint32_t len) {
    m.words = words;
    m.len = len;
    m.wci = 0xFFFFFFFFu;
    m.cur = mq_byte(m, 0);
    m.np = 1;
    m.C = m.cur << 16;
    mq_bytein(m);
    m.C <<= 7;
    m.CT -= 7;
    m.A = 0x80000000u;
}

// RENORMD across a byte boundary (CT < n; CT already reduced by n, so it
// holds CT - n), common case in one straight run: the byte at BP is in the
// cached dword, the byte before it is not 0xFF and one byte covers the rest
// of the shift (n - CT <= 8).  BYTEIN's C += B << 8 lands before the last
// n - CT shifts, so C = (C << n) + (B << (8 + n - CT)) and CT = 8 - (n - CT).
// Anything else continues at P "slow" (CT restored there).
#define TC_MQ_REFILL(P, TAIL, EXIT)                  \
    "L%=" P "refill:\n\t"                             \
    "s_cmp_ge_u32 %[np], %[len]\n\t"                  \
    "s_cbranch_scc1 L%=" P "slow\n\t"                 \
    "s_cmp_eq_u32 %[cur], 0xff\n\t"                   \
    "s_cbranch_scc1 L%=" P "slow\n\t"                 \
    "s_lshr_b32 %[t1], %[np], 2\n\t"                  \
    "s_cmp_eq_u32 %[t1], %[wci]\n\t"                  \
    "s_cbranch_scc0 L%=" P "slow\n\t"                 \
    "s_add_u32 %[t1], %[CT], 8\n\t"                   \
    "s_cmp_lt_i32 %[t1], 0\n\t"                       \
    "s_cbranch_scc1 L%=" P "slow\n\t"                 \
    "s_mov_b32 %[CT], %[t1]\n\t"                      \
    "s_and_b32 %[t1], %[np], 3\n\t"                   \
    "s_lshl_b32 %[t1], %[t1], 3\n\t"                  \
    "s_lshr_b32 %[t0], %[wc], %[t1]\n\t"              \
    "s_and_b32 %[cur], %[t0], 0xff\n\t"               \
    "s_lshl_b32 %[C], %[C], %[n]\n\t"                 \
    "s_sub_u32 %[t1], 16, %[CT]\n\t"                  \
    "s_lshl_b32 %[t0], %[cur], %[t1]\n\t"             \
    "s_add_u32 %[C], %[C], %[t0]\n\t"                 \
    "s_lshl_b32 %[A], %[A], %[n]\n\t"                 \
    "s_add_u32 %[np], %[np], 1\n\t"                   \
    TAIL "s_branch " EXIT "\n"

// DECODE (C.3.2) with the conditional exchanges (C.3.3), the successor state
// and RENORMD/BYTEIN (C.3.3/C.3.4), RENORMD shifting min(needed, CT) bits per
// step (the same bits as one at a time), over the operands of the asm block
// that uses it: label prefix P, context state operand ST (Qe << 16 | I << 1 |
// MPS, updated in place), the decision in d.  A decision without
// renormalisation branches to FAST; every other path runs TAIL and branches
// to EXIT.  trm/trl: lane I holds the packed successor state of I after an MPS /
// LPS (the LPS one with SWITCH in bit 0); both are read first so their
// latency overlaps the interval arithmetic.  CT is reduced first and restored
// on the refill path; the codeword dword load waits for itself.
#define TC_MQ_DECIDE(P, ST, FAST, TAIL, EXIT)                \
    "s_bfe_u32 %[t0], %[" ST "], 0x60001\n\t"             \
    "s_and_b32 %[q], %[" ST "], 0xffff0000\n\t"           \
    "s_and_b32 %[d], %[" ST "], 1\n\t"                    \
    "s_sub_u32 %[A], %[A], %[q]\n\t"                  \
    "v_readlane_b32 %[t2], %[trm], %[t0]\n\t"         \
    "v_readlane_b32 %[t3], %[trl], %[t0]\n\t"         \
    "s_cmp_lt_u32 %[C], %[q]\n\t"                     \
    "s_cbranch_scc1 L%=" P "lower\n\t"                \
    "s_sub_u32 %[C], %[C], %[q]\n\t"                  \
    "s_cmp_lt_i32 %[A], 0\n\t"                        \
    "s_cbranch_scc1 " FAST "\n\t"                     \
    "s_cmp_lt_u32 %[A], %[q]\n\t"                     \
    "s_cselect_b32 %[t2], %[t3], %[t2]\n\t"           \
    "s_cselect_b32 %[t1], 1, 0\n"                     \
    "L%=" P "renorm:\n\t"                             \
    "s_xor_b32 %[" ST "], %[t2], %[d]\n\t"                \
    "s_xor_b32 %[d], %[d], %[t1]\n\t"                 \
    "s_flbit_i32_b32 %[n], %[A]\n\t"                  \
    "s_sub_u32 %[CT], %[CT], %[n]\n\t"                \
    "s_cbranch_scc1 L%=" P "refill\n\t"               \
    "s_lshl_b32 %[C], %[C], %[n]\n\t"                 \
    "s_lshl_b32 %[A], %[A], %[n]\n\t"                 \
    TAIL "s_branch " EXIT "\n"                             \
    "L%=" P "lower:\n\t"                              \
    "s_cmp_lt_u32 %[A], %[q]\n\t"                     \
    "s_cselect_b32 %[t2], %[t2], %[t3]\n\t"           \
    "s_cselect_b32 %[t1], 0, 1\n\t"                   \
    "s_mov_b32 %[A], %[q]\n\t"                        \
    "s_xor_b32 %[" ST "], %[t2], %[d]\n\t"            \
    "s_xor_b32 %[d], %[d], %[t1]\n\t"                 \
    "s_flbit_i32_b32 %[n], %[A]\n\t"                  \
    "s_sub_u32 %[CT], %[CT], %[n]\n\t"                \
    "s_cbranch_scc1 L%=" P "refill\n\t"               \
    "s_lshl_b32 %[C], %[C], %[n]\n\t"                 \
    "s_lshl_b32 %[A], %[A], %[n]\n\t"                 \
    TAIL "s_branch " EXIT "\n"                             \
    TC_MQ_REFILL(P, TAIL, EXIT)                             \
    "L%=" P "slow:\n\t"                               \
    "s_add_u32 %[CT], %[CT], %[n]\n"                  \
    "L%=" P "loop:\n\t"                               \
    "s_cmp_eq_u32 %[CT], 0\n\t"                       \
    "s_cbranch_scc0 L%=" P "shift\n\t"                \
    "s_cmp_ge_u32 %[np], %[len]\n\t"                  \
    "s_mov_b32 %[t0], 0xff\n\t"                       \
    "s_cbranch_scc1 L%=" P "got\n\t"                  \
    "s_lshr_b32 %[t1], %[np], 2\n\t"                  \
    "s_cmp_eq_u32 %[t1], %[wci]\n\t"                  \
    "s_cbranch_scc1 L%=" P "cached\n\t"               \
    "s_mov_b32 %[wci], %[t1]\n\t"                     \
    "s_lshl_b32 %[t1], %[t1], 2\n\t"                  \
    "s_load_dword %[wc], %[wp], %[t1]\n\t"            \
    "s_waitcnt lgkmcnt(0)\n"                          \
    "L%=" P "cached:\n\t"                             \
    "s_and_b32 %[t1], %[np], 3\n\t"                   \
    "s_lshl_b32 %[t1], %[t1], 3\n\t"                  \
    "s_lshr_b32 %[t0], %[wc], %[t1]\n\t"              \
    "s_and_b32 %[t0], %[t0], 0xff\n"                  \
    "L%=" P "got:\n\t"                                \
    "s_cmp_eq_u32 %[cur], 0xff\n\t"                   \
    "s_cbranch_scc0 L%=" P "plain\n\t"                \
    "s_cmp_gt_u32 %[t0], 0x8f\n\t"                    \
    "s_cbranch_scc0 L%=" P "stuff\n\t"                \
    "s_add_u32 %[C], %[C], 0xff00\n\t"                \
    "s_mov_b32 %[CT], 8\n\t"                          \
    "s_branch L%=" P "shift\n"                        \
    "L%=" P "stuff:\n\t"                              \
    "s_add_u32 %[np], %[np], 1\n\t"                   \
    "s_mov_b32 %[cur], %[t0]\n\t"                     \
    "s_lshl_b32 %[t0], %[t0], 9\n\t"                  \
    "s_add_u32 %[C], %[C], %[t0]\n\t"                 \
    "s_mov_b32 %[CT], 7\n\t"                          \
    "s_branch L%=" P "shift\n"                        \
    "L%=" P "plain:\n\t"                              \
    "s_add_u32 %[np], %[np], 1\n\t"                   \
    "s_mov_b32 %[cur], %[t0]\n\t"                     \
    "s_lshl_b32 %[t0], %[t0], 8\n\t"                  \
    "s_add_u32 %[C], %[C], %[t0]\n\t"                 \
    "s_mov_b32 %[CT], 8\n"                            \
    "L%=" P "shift:\n\t"                              \
    "s_min_u32 %[t1], %[n], %[CT]\n\t"                \
    "s_lshl_b32 %[C], %[C], %[t1]\n\t"                \
    "s_lshl_b32 %[A], %[A], %[t1]\n\t"                \
    "s_sub_u32 %[CT], %[CT], %[t1]\n\t"               \
    "s_sub_u32 %[n], %[n], %[t1]\n\t"                 \
    "s_cmp_eq_u32 %[n], 0\n\t"                        \
    "s_cbranch_scc0 L%=" P "loop\n\t"                 \
    TAIL "s_branch " EXIT "\n"
#define TC_MQ_BODY(P, ST, EXIT) TC_MQ_DECIDE(P, ST, EXIT, "", EXIT)

// The refinement decisions of one stripe (T.800 D.3.3) in one scalar loop.
// cols: the columns holding candidates (not empty); lane x of stream: column
// x's candidate labels in scan order, 4 bits each (label - 14), ended by 15,
// and, when all of them are the same label c, c | 4 << 24 | (count - 1) << 28.
// Column x's decisions go to lane x of res, the first one highest.  A label
// other than the one used last swaps the context states through the lane
// array ctxv (cl/cs as in Blk; cl is held as label - 14 inside); a column
// whose candidates all use the context used last runs a counted loop without
// label tests (after swapping that label in when it is not the one used
// last).  The next column's label word is read while the current column
// decodes.
#define TC_MRP_BIT "s_lshl1_add_u32 %[bits], %[bits], %[d]\n\t"
#define TC_MRP_COUNT TC_MRP_BIT "s_sub_u32 %[cnt], %[cnt], 1\n\ts_cbranch_scc0 L%=_fast\n\t"
__device__ __forceinline__ void mq_mrp_stripe(Mq& m, uint32_t& cs, uint32_t& cl, uint32_t& ctxv, uint64_t cols,
                                              uint32_t stream, uint32_t& res, uint32_t trm, uint32_t trl) {
    uint32_t x, xn, iw, iwn, bits, code, cnt, d, q, t0, t1, t2, t3, n;
    asm volatile(
        "s_sub_u32 %[cl], %[cl], 14\n\t"
        "s_ff1_i32_b64 %[x], %[cols]\n\t"
        "v_readlane_b32 %[iw], %[stream], %[x]\n"
        "L%=_col:\n\t"
        "s_bitset0_b64 %[cols], %[x]\n\t"
        "s_ff1_i32_b64 %[xn], %[cols]\n\t"
        "s_max_i32 %[xn], %[xn], 0\n\t"
        "s_mov_b32 %[bits], 0\n\t"
        "v_readlane_b32 %[iwn], %[stream], %[xn]\n\t"
        "s_bfe_u32 %[t0], %[iw], 0x30018\n\t"
        "s_or_b32 %[t1], %[cl], 4\n\t"
        "s_cmp_eq_u32 %[t0], %[t1]\n\t"
        "s_cbranch_scc1 L%=_cnt\n\t"
        "s_bitcmp1_b32 %[iw], 26\n\t"
        "s_cbranch_scc0 L%=_next\n\t"
        // one label for the whole column, not the one used last: swap it in
        "s_and_b32 %[t0], %[t0], 3\n\t"
        "s_add_u32 %[t1], %[cl], 14\n\t"
        "s_mov_b32 m0, %[t1]\n\t"
        "v_writelane_b32 %[ctxv], %[cs], m0\n\t"
        "s_mov_b32 %[cl], %[t0]\n\t"
        "s_add_u32 %[t1], %[t0], 14\n\t"
        "v_readlane_b32 %[cs], %[ctxv], %[t1]\n"
        "L%=_cnt:\n\t"
        "s_bfe_u32 %[cnt], %[iw], 0x2001c\n"
        "L%=_fast:\n\t"
        TC_MQ_DECIDE("f", "cs", "L%=ffx", TC_MRP_COUNT, "L%=_colend")
        "L%=ffx:\n\t"
        TC_MRP_COUNT
        "s_branch L%=_colend\n"
        "L%=_next:\n\t"
        "s_and_b32 %[code], %[iw], 15\n\t"
        "s_lshr_b32 %[iw], %[iw], 4\n\t"
        "s_cmp_eq_u32 %[code], %[cl]\n\t"
        "s_cbranch_scc0 L%=_other\n"
        "L%=_dec:\n\t"
        TC_MQ_DECIDE("g", "cs", "L%=gfx", TC_MRP_BIT, "L%=_next")
        "L%=gfx:\n\t"
        TC_MRP_BIT
        "s_branch L%=_next\n"
        // another label, or the column's end
        "L%=_other:\n\t"
        "s_cmp_eq_u32 %[code], 15\n\t"
        "s_cbranch_scc1 L%=_colend\n\t"
        "s_add_u32 %[t0], %[cl], 14\n\t"
        "s_mov_b32 m0, %[t0]\n\t"
        "v_writelane_b32 %[ctxv], %[cs], m0\n\t"
        "s_add_u32 %[t0], %[code], 14\n\t"
        "s_mov_b32 %[cl], %[code]\n\t"
        "v_readlane_b32 %[cs], %[ctxv], %[t0]\n\t"
        "s_branch L%=_dec\n"
        "L%=_colend:\n\t"
        "s_mov_b32 m0, %[x]\n\t"
        "v_writelane_b32 %[res], %[bits], m0\n\t"
        "s_cmp_eq_u64 %[cols], 0\n\t"
        "s_cbranch_scc1 L%=_end\n\t"
        "s_mov_b32 %[x], %[xn]\n\t"
        "s_mov_b32 %[iw], %[iwn]\n\t"
        "s_branch L%=_col\n"
        // no lane read may still be in flight when the registers are handed back
        "L%=_end:\n\t"
        "s_or_b32 %[t2], %[t2], %[t3]\n\t"
        "s_or_b32 %[t2], %[t2], %[iwn]\n\t"
        "s_add_u32 %[cl], %[cl], 14"
        : [A] "+s"(m.A), [C] "+s"(m.C), [CT] "+s"(m.CT), [cs] "+s"(cs), [cl] "+s"(cl), [cur] "+s"(m.cur),
          [np] "+s"(m.np), [wc] "+s"(m.wc), [wci] "+s"(m.wci), [cols] "+s"(cols), [ctxv] "+v"(ctxv),
          [res] "+v"(res), [x] "=&s"(x), [xn] "=&s"(xn), [iw] "=&s"(iw), [iwn] "=&s"(iwn), [bits] "=&s"(bits),
          [code] "=&s"(code), [cnt] "=&s"(cnt), [d] "=&s"(d), [q] "=&s"(q), [t0] "=&s"(t0), [t1] "=&s"(t1),
          [t2] "=&s"(t2), [t3] "=&s"(t3), [n] "=&s"(n)
        : [len] "s"(m.len), [wp] "s"(m.words), [stream] "v"(stream), [trm] "v"(trm), [trl] "v"(trl)
        : "scc", "m0");
}

// The sign of stripe row r (r3 = 3r), which just became significant (T.800
// D.3.2): the sign-coding entry (label | XOR bit << 4) of the 8-bit key that
// pairs (sign, significance) of the upper, left, right and lower neighbours
// from sl and gl, the context swapped in through ctxv as in Blk; then
// row r joins sl/gl, nsig/csgn and the next column's carry/gcarry, and the
// row below joins rows when it is in `below`.
#define TC_SIGN_UPDATE(P, EXIT)                      \
    "s_lshr_b32 %[u0], %[sl], %[r3]\n\t"              \
    "s_and_b32 %[u0], %[u0], 0xaa\n\t"                \
    "s_lshr_b32 %[u1], %[gl], %[r3]\n\t"              \
    "s_lshr_b32 %[u1], %[u1], 1\n\t"                  \
    "s_and_b32 %[u1], %[u1], 0x55\n\t"                \
    "s_or_b32 %[u0], %[u0], %[u1]\n\t"                \
    "s_lshr_b32 %[u1], %[u0], 2\n\t"                  \
    "v_readlane_b32 %[lab], %[scv], %[u1]\n\t"        \
    "s_and_b32 %[u0], %[u0], 3\n\t"                   \
    "s_lshl_b32 %[u0], %[u0], 3\n\t"                  \
    "s_lshr_b32 %[lab], %[lab], %[u0]\n\t"            \
    "s_bfe_u32 %[xr], %[lab], 0x10004\n\t"            \
    "s_and_b32 %[lab], %[lab], 15\n\t"                \
    "s_cmp_eq_u32 %[lab], %[cl]\n\t"                  \
    "s_cbranch_scc1 L%=" P "dec\n\t"                  \
    "s_mov_b32 m0, %[cl]\n\t"                         \
    "v_writelane_b32 %[ctxv], %[cs], m0\n\t"          \
    "s_mov_b32 %[cl], %[lab]\n\t"                     \
    "v_readlane_b32 %[cs], %[ctxv], %[lab]\n"         \
    "L%=" P "dec:\n\t"                                \
    TC_MQ_BODY(P "m", "cs", "L%=" P "res")            \
    "L%=" P "res:\n\t"                                \
    "s_xor_b32 %[d], %[d], %[xr]\n\t"                 \
    "s_add_u32 %[u0], %[r3], 4\n\t"                   \
    "s_bitset1_b32 %[sl], %[u0]\n\t"                  \
    "s_lshl_b32 %[u1], %[d], %[u0]\n\t"               \
    "s_or_b32 %[gl], %[gl], %[u1]\n\t"                \
    "s_add_u32 %[u0], %[r3], 3\n\t"                   \
    "s_bitset1_b32 %[carry], %[u0]\n\t"               \
    "s_lshl_b32 %[u1], %[d], %[u0]\n\t"               \
    "s_or_b32 %[gcarry], %[gcarry], %[u1]\n\t"        \
    "s_bitset1_b32 %[nsig], %[r]\n\t"                 \
    "s_lshl_b32 %[u1], %[d], %[r]\n\t"                \
    "s_or_b32 %[csgn], %[csgn], %[u1]\n\t"            \
    "s_lshl_b32 %[u1], 2, %[r]\n\t"                   \
    "s_and_b32 %[u1], %[u1], %[below]\n\t"            \
    "s_or_b32 %[rows], %[rows], %[u1]\n\t"            \
    "s_branch " EXIT "\n"

// The zero-coding decisions of one stripe column (T.800 D.3.1 / D.3.4), rows
// not empty: each row of rows in order (into vis), its label from w1 (labels
// of rows 0..3 as of the stripe's start, 4 bits each from bit 14) unless its
// 3x3 window in sl differs from the one in w0 (bits 0..17), then from the
// zero-coding table zcv; a decision of 1 decodes the sign (TC_SIGN_UPDATE).
// Ends with a branch to EXIT.
#define TC_ZC_ROWS(P, EXIT)                          \
    "L%=" P "row:\n\t"                                \
    "s_ff1_i32_b32 %[r], %[rows]\n\t"                 \
    "s_bitset0_b32 %[rows], %[r]\n\t"                 \
    "s_bitset1_b32 %[vis], %[r]\n\t"                  \
    "s_lshl1_add_u32 %[r3], %[r], %[r]\n\t"           \
    "s_xor_b32 %[u0], %[sl], %[w0]\n\t"               \
    "s_add_u32 %[u1], %[r3], 0x90000\n\t"             \
    "s_bfe_u32 %[u0], %[u0], %[u1]\n\t"               \
    "s_cbranch_scc1 L%=" P "zc\n\t"                   \
    "s_lshl2_add_u32 %[u1], %[r], 0x4000e\n\t"        \
    "s_bfe_u32 %[lab], %[w1], %[u1]\n"                \
    "L%=" P "lab:\n\t"                                \
    "s_cmp_eq_u32 %[lab], %[cl]\n\t"                  \
    "s_cbranch_scc0 L%=" P "sw\n"                     \
    "L%=" P "dec:\n\t"                                \
    TC_MQ_BODY(P "z", "cs", "L%=" P "res")            \
    "L%=" P "res:\n\t"                                \
    "s_cmp_eq_u32 %[d], 0\n\t"                        \
    "s_cbranch_scc0 L%=" P "sig\n"                    \
    "L%=" P "next:\n\t"                               \
    "s_cmp_lg_u32 %[rows], 0\n\t"                     \
    "s_cbranch_scc1 L%=" P "row\n\t"                  \
    "s_branch " EXIT "\n"                             \
    "L%=" P "zc:\n\t"                                 \
    "s_bfe_u32 %[u0], %[sl], %[u1]\n\t"               \
    "s_lshr_b32 %[u1], %[u0], 3\n\t"                  \
    "v_readlane_b32 %[lab], %[zcv], %[u1]\n\t"        \
    "s_and_b32 %[u0], %[u0], 7\n\t"                   \
    "s_lshl_b32 %[u0], %[u0], 2\n\t"                  \
    "s_lshr_b32 %[lab], %[lab], %[u0]\n\t"            \
    "s_and_b32 %[lab], %[lab], 15\n\t"                \
    "s_branch L%=" P "lab\n"                          \
    "L%=" P "sw:\n\t"                                 \
    "s_mov_b32 m0, %[cl]\n\t"                         \
    "v_writelane_b32 %[ctxv], %[cs], m0\n\t"          \
    "s_mov_b32 %[cl], %[lab]\n\t"                     \
    "v_readlane_b32 %[cs], %[ctxv], %[lab]\n\t"       \
    "s_branch L%=" P "dec\n"                          \
    "L%=" P "sig:\n\t"                                \
    TC_SIGN_UPDATE(P "g", "L%=" P "next")

// Operands shared by the two stripe loops below.
#define TC_STRIPE_OPERANDS                                                                                          \
    : [A] "+s"(m.A), [C] "+s"(m.C), [CT] "+s"(m.CT), [cs] "+s"(cs), [cl] "+s"(cl), [cur] "+s"(m.cur),               \
      [np] "+s"(m.np), [wc] "+s"(m.wc), [wci] "+s"(m.wci), [cols] "+s"(cols), [ctxv] "+v"(ctxv), [res] "+v"(res),   \
      [x] "=&s"(x), [xp] "=&s"(xp), [carry] "=&s"(carry), [gcarry] "=&s"(gcarry), [w0] "=&s"(w0), [w1] "=&s"(w1),  \
      [pw0] "=&s"(pw0), [pw1] "=&s"(pw1), [pw2] "=&s"(pw2), [sl] "=&s"(sl), [gl] "=&s"(gl), [rows] "=&s"(rows),    \
      [vis] "=&s"(vis), [nsig] "=&s"(nsig), [csgn] "=&s"(csgn), [below] "=&s"(below), [r] "=&s"(r),                 \
      [r3] "=&s"(r3), [u0] "=&s"(u0), [u1] "=&s"(u1), [lab] "=&s"(lab), [xr] "=&s"(xr), [d] "=&s"(d), [q] "=&s"(q), \
      [t0] "=&s"(t0), [t1] "=&s"(t1), [t2] "=&s"(t2), [t3] "=&s"(t3), [n] "=&s"(n)

// Significance propagation over one stripe (T.800 D.3.1) in one scalar loop.
// Columns are visited in order: the next candidate column of cols, or x + 1
// (inside the block, w columns) after x gained significance.  Lane x of vw0,
// vw1, vw2: column x's words (Blk::column_words); the next candidate column's
// words are read while the current one decodes.  rowm: the stripe's rows.
// Column x's result goes to lane x of res: coded rows | new significance << 4
// | their signs << 8.  A column's window takes the rows that turned
// significant in the column to its left (carry, gcarry); rows next to them
// join its candidates.
__device__ __forceinline__ void spp_stripe(Mq& m, uint32_t& cs, uint32_t& cl, uint32_t& ctxv, uint64_t cols,
                                           uint32_t vw0, uint32_t vw1, uint32_t vw2, uint32_t& res, uint32_t rowm,
                                           uint32_t w, uint32_t zcv, uint32_t scv, uint32_t trm, uint32_t trl) {
    uint32_t x, xp, carry, gcarry, w0, w1, pw0, pw1, pw2, sl, gl, rows, vis, nsig, csgn, below;
    uint32_t r, r3, u0, u1, lab, xr, d, q, t0, t1, t2, t3, n;
    asm volatile(
        "s_mov_b32 %[x], -1\n\t"
        "s_mov_b32 %[nsig], 0\n\t"
        "s_mov_b32 %[carry], 0\n\t"
        "s_mov_b32 %[gcarry], 0\n\t"
        "s_ff1_i32_b64 %[xp], %[cols]\n\t"
        "s_max_i32 %[u0], %[xp], 0\n\t"
        "v_readlane_b32 %[pw0], %[vw0], %[u0]\n\t"
        "v_readlane_b32 %[pw1], %[vw1], %[u0]\n\t"
        "v_readlane_b32 %[pw2], %[vw2], %[u0]\n"
        "L%=col:\n\t"
        "s_add_u32 %[u0], %[x], 1\n\t"
        "s_cmp_lt_u32 %[u0], %[w]\n\t"
        "s_cselect_b32 %[u1], %[nsig], 0\n\t"
        "s_cmp_lg_u32 %[u1], 0\n\t"
        "s_cselect_b32 %[r], %[u0], %[xp]\n\t"
        "s_cmp_lt_i32 %[r], 0\n\t"
        "s_cbranch_scc1 L%=end\n\t"
        "s_cmp_eq_u32 %[r], %[u0]\n\t"
        "s_cselect_b32 %[lab], %[nsig], 0\n\t"
        "s_cselect_b32 %[t0], %[carry], 0\n\t"
        "s_cselect_b32 %[t1], %[gcarry], 0\n\t"
        "s_mov_b32 %[x], %[r]\n\t"
        "s_cmp_eq_u32 %[x], %[xp]\n\t"
        "s_cbranch_scc1 L%=pf\n\t"
        "v_readlane_b32 %[w0], %[vw0], %[x]\n\t"
        "v_readlane_b32 %[w1], %[vw1], %[x]\n\t"
        "v_readlane_b32 %[t2], %[vw2], %[x]\n\t"
        "s_branch L%=got\n"
        "L%=pf:\n\t"
        "s_mov_b32 %[w0], %[pw0]\n\t"
        "s_mov_b32 %[w1], %[pw1]\n\t"
        "s_mov_b32 %[t2], %[pw2]\n"
        "L%=got:\n\t"
        "s_bitset0_b64 %[cols], %[x]\n\t"
        "s_ff1_i32_b64 %[xp], %[cols]\n\t"
        "s_max_i32 %[u0], %[xp], 0\n\t"
        "v_readlane_b32 %[pw0], %[vw0], %[u0]\n\t"
        "v_readlane_b32 %[pw1], %[vw1], %[u0]\n\t"
        "v_readlane_b32 %[pw2], %[vw2], %[u0]\n\t"
        "s_bfe_u32 %[t3], %[w0], 0x40016\n\t"
        "s_andn2_b32 %[below], %[rowm], %[t3]\n\t"
        "s_and_b32 %[sl], %[w0], 0x3ffff\n\t"
        "s_or_b32 %[sl], %[sl], %[t0]\n\t"
        "s_or_b32 %[gl], %[t2], %[t1]\n\t"
        "s_lshl_b32 %[u0], %[lab], 1\n\t"
        "s_lshr_b32 %[u1], %[lab], 1\n\t"
        "s_or_b32 %[u0], %[u0], %[u1]\n\t"
        "s_or_b32 %[u0], %[u0], %[lab]\n\t"
        "s_and_b32 %[u0], %[u0], %[below]\n\t"
        "s_bfe_u32 %[rows], %[w0], 0x40012\n\t"
        "s_or_b32 %[rows], %[rows], %[u0]\n\t"
        "s_mov_b32 %[vis], 0\n\t"
        "s_mov_b32 %[nsig], 0\n\t"
        "s_mov_b32 %[csgn], 0\n\t"
        "s_mov_b32 %[carry], 0\n\t"
        "s_mov_b32 %[gcarry], 0\n\t"
        "s_cmp_eq_u32 %[rows], 0\n\t"
        "s_cbranch_scc1 L%=cend\n"
        TC_ZC_ROWS("p", "L%=cend")
        "L%=cend:\n\t"
        "s_cmp_eq_u32 %[vis], 0\n\t"
        "s_cbranch_scc1 L%=col\n\t"
        "s_lshl_b32 %[u0], %[nsig], 4\n\t"
        "s_or_b32 %[u0], %[u0], %[vis]\n\t"
        "s_lshl_b32 %[u1], %[csgn], 8\n\t"
        "s_or_b32 %[u0], %[u0], %[u1]\n\t"
        "s_mov_b32 m0, %[x]\n\t"
        "v_writelane_b32 %[res], %[u0], m0\n\t"
        "s_branch L%=col\n"
        // no lane read may still be in flight when the registers are handed back
        "L%=end:\n\t"
        "s_or_b32 %[u0], %[t2], %[t3]\n\t"
        "s_or_b32 %[u0], %[pw0], %[pw1]\n\t"
        "s_or_b32 %[u0], %[u0], %[pw2]"
        TC_STRIPE_OPERANDS
        : [len] "s"(m.len), [wp] "s"(m.words), [rowm] "s"(rowm), [w] "s"(w), [vw0] "v"(vw0), [vw1] "v"(vw1),
          [vw2] "v"(vw2), [zcv] "v"(zcv), [scv] "v"(scv), [trm] "v"(trm), [trl] "v"(trl)
        : "scc", "m0");
}

// Cleanup over one stripe (T.800 D.3.4) in one scalar loop over the candidate
// columns of cols (columns of vw0/vw1/vw2 and res as in spp_stripe; the
// result is new significance | their signs << 8).  A column of four
// candidate rows with an empty window starts in run-length mode (states strl
// and stuni for the run-length and uniform contexts): a 0 ends the column, a
// 1 is followed by the row (two uniform decisions) and its sign, and the rows
// after it continue as zero-coding rows.
__device__ __forceinline__ void cup_stripe(Mq& m, uint32_t& cs, uint32_t& cl, uint32_t& ctxv, uint32_t& strl,
                                           uint32_t& stuni, uint64_t cols, uint32_t vw0, uint32_t vw1, uint32_t vw2,
                                           uint32_t& res, uint32_t zcv, uint32_t scv, uint32_t trm, uint32_t trl) {
    uint32_t x, xp, carry, gcarry, w0, w1, pw0, pw1, pw2, sl, gl, rows, vis, nsig, csgn, below;
    uint32_t r, r3, u0, u1, lab, xr, d, q, t0, t1, t2, t3, n;
    asm volatile(
        "s_mov_b32 %[x], -2\n\t"
        "s_mov_b32 %[below], 0\n\t"
        "s_mov_b32 %[carry], 0\n\t"
        "s_mov_b32 %[gcarry], 0\n\t"
        "s_ff1_i32_b64 %[xp], %[cols]\n\t"
        "s_max_i32 %[u0], %[xp], 0\n\t"
        "v_readlane_b32 %[pw0], %[vw0], %[u0]\n\t"
        "v_readlane_b32 %[pw1], %[vw1], %[u0]\n\t"
        "v_readlane_b32 %[pw2], %[vw2], %[u0]\n"
        "L%=col:\n\t"
        "s_cmp_lt_i32 %[xp], 0\n\t"
        "s_cbranch_scc1 L%=end\n\t"
        "s_add_u32 %[u0], %[x], 1\n\t"
        "s_cmp_eq_u32 %[xp], %[u0]\n\t"
        "s_cselect_b32 %[t0], %[carry], 0\n\t"
        "s_cselect_b32 %[t1], %[gcarry], 0\n\t"
        "s_mov_b32 %[x], %[xp]\n\t"
        "s_mov_b32 %[w0], %[pw0]\n\t"
        "s_mov_b32 %[w1], %[pw1]\n\t"
        "s_mov_b32 %[t2], %[pw2]\n\t"
        "s_bitset0_b64 %[cols], %[x]\n\t"
        "s_ff1_i32_b64 %[xp], %[cols]\n\t"
        "s_max_i32 %[u0], %[xp], 0\n\t"
        "v_readlane_b32 %[pw0], %[vw0], %[u0]\n\t"
        "v_readlane_b32 %[pw1], %[vw1], %[u0]\n\t"
        "v_readlane_b32 %[pw2], %[vw2], %[u0]\n\t"
        "s_and_b32 %[sl], %[w0], 0x3ffff\n\t"
        "s_or_b32 %[sl], %[sl], %[t0]\n\t"
        "s_or_b32 %[gl], %[t2], %[t1]\n\t"
        "s_bfe_u32 %[rows], %[w0], 0x40012\n\t"
        "s_mov_b32 %[nsig], 0\n\t"
        "s_mov_b32 %[csgn], 0\n\t"
        "s_mov_b32 %[carry], 0\n\t"
        "s_mov_b32 %[gcarry], 0\n\t"
        "s_cmp_eq_u32 %[rows], 15\n\t"
        "s_cbranch_scc0 L%=zrows\n\t"
        "s_cmp_eq_u32 %[sl], 0\n\t"
        "s_cbranch_scc0 L%=zrows\n\t"
        TC_MQ_BODY("a", "strl", "L%=rl1")
        "L%=rl1:\n\t"
        "s_cmp_eq_u32 %[d], 0\n\t"
        "s_cbranch_scc1 L%=col\n\t"
        TC_MQ_BODY("b", "stuni", "L%=rl2")
        "L%=rl2:\n\t"
        "s_lshl_b32 %[r], %[d], 1\n\t"
        TC_MQ_BODY("c", "stuni", "L%=rl3")
        "L%=rl3:\n\t"
        "s_or_b32 %[r], %[r], %[d]\n\t"
        "s_lshl1_add_u32 %[r3], %[r], %[r]\n\t"
        TC_SIGN_UPDATE("e", "L%=rl4")
        "L%=rl4:\n\t"
        "s_lshl_b32 %[u0], 2, %[r]\n\t"
        "s_sub_u32 %[u0], %[u0], 1\n\t"
        "s_andn2_b32 %[rows], %[rows], %[u0]\n"
        "L%=zrows:\n\t"
        "s_cmp_eq_u32 %[rows], 0\n\t"
        "s_cbranch_scc1 L%=cend\n"
        TC_ZC_ROWS("q", "L%=cend")
        "L%=cend:\n\t"
        "s_cmp_eq_u32 %[nsig], 0\n\t"
        "s_cbranch_scc1 L%=col\n\t"
        "s_lshl_b32 %[u1], %[csgn], 8\n\t"
        "s_or_b32 %[u0], %[nsig], %[u1]\n\t"
        "s_mov_b32 m0, %[x]\n\t"
        "v_writelane_b32 %[res], %[u0], m0\n\t"
        "s_branch L%=col\n"
        // no lane read may still be in flight when the registers are handed back
        "L%=end:\n\t"
        "s_or_b32 %[u0], %[t2], %[t3]\n\t"
        "s_or_b32 %[u0], %[pw0], %[pw1]\n\t"
        "s_or_b32 %[u0], %[u0], %[pw2]"
        TC_STRIPE_OPERANDS, [strl] "+s"(strl), [stuni] "+s"(stuni)
        : [len] "s"(m.len), [wp] "s"(m.words), [vw0] "v"(vw0), [vw1] "v"(vw1), [vw2] "v"(vw2), [zcv] "v"(zcv),
          [scv] "v"(scv), [trm] "v"(trm), [trl] "v"(trl)
        : "scc", "m0");
}

// Zero-coding label (T.800 Table D.1) of a 9-bit window (centre ignored),
// computed per lane from the neighbour counts.
__device__ __forceinline__ uint32_t zc_lab_lane(uint32_t nb, int cls) {
    uint32_t h = (uint32_t)__popc(nb & 0x28u), v = (uint32_t)__popc(nb & 0x82u);
    const uint32_t dg = (uint32_t)__popc(nb & 0x145u);
    if (cls == 1) {
        const uint32_t t = h;
        h = v;
        v = t;
    }
    if (cls < 2) {
        if (h == 2) return 8;
        if (h == 1) return v ? 7 : (dg ? 6 : 5);
        if (v == 2) return 4;
        if (v == 1) return 3;
        return dg >= 2 ? 2 : dg;
    }
    const uint32_t hv = h + v;
    if (dg >= 3) return 8;
    if (dg == 2) return hv ? 7 : 6;
    if (dg == 1) return hv >= 2 ? 5 : (hv ? 4 : 3);
    return hv >= 2 ? 2 : hv;
}

struct BlockArgs {
    uint32_t data_off, data_len;
    uint64_t out_off;
    uint32_t out_pitch;
    int w, h, numbps, npasses, mode, cls;
    float delta;
};

__device__ __forceinline__ BlockArgs load_args(const tc_cblk_desc* __restrict__ dp) {
    const uint32_t* q = reinterpret_cast<const uint32_t*>(dp);  // 8 dwords, layout of tc_cblk_desc
    BlockArgs a;
    a.data_off = uni(q[0]);
    a.data_len = uni(q[1]);
    a.out_off = (uint64_t)uni(q[2]) | ((uint64_t)uni(q[3]) << 32);
    a.out_pitch = uni(q[4]);
    const uint32_t wh = uni(q[5]), pk = uni(q[6]);
    a.w = (int)(wh & 0xFFFF);
    a.h = (int)(wh >> 16);
    a.numbps = (int)(pk & 0xFF);
    a.npasses = (int)((pk >> 8) & 0xFF);
    a.cls = (int)((pk >> 16) & 0xFF);
    a.mode = (int)((pk >> 24) & 0xFF);
    a.delta = __uint_as_float(uni(q[7]));
    return a;
}

// bits 0..5 of v to bits 0,3,...,15
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4) | ((v & 8u) << 6) | ((v & 16u) << 8) | ((v & 32u) << 10);
}

// rows y0-1 .. y0+4 of a column mask as a 6-bit value (row y0-1 in bit 0)
__device__ __forceinline__ uint32_t win6(uint64_t m, int y0) {
    return y0 == 0 ? (uint32_t)(m << 1) & 63u : (uint32_t)(m >> (y0 - 1)) & 63u;
}

__device__ __forceinline__ uint32_t from_left(uint32_t v) {
    const uint32_t u = (uint32_t)__shfl_up((int)v, 1, 64);
    return threadIdx.x == 0 ? 0u : u;
}
__device__ __forceinline__ uint32_t from_right(uint32_t v) {
    const uint32_t u = (uint32_t)__shfl_down((int)v, 1, 64);
    return threadIdx.x == 63 ? 0u : u;
}

// Decoder state of one block.  Scalar members are wave-uniform; the lane
// arrays hold one entry per lane.
struct Blk {
    Mq m;
    uint32_t ctxv;            // lane l < 17: state of zero/sign/refinement context l
    // the context used last and its current state; the state goes back to
    // lane cl of ctxv when a decision uses another context
    uint32_t cl, cs;
    uint32_t st_rl, st_uni;
    uint32_t zcv, scv, trm, trl;
    uint64_t sig, sgn, pip, ref, ones;  // lane x: column x (bit y = row y)
    int w, h, cls;

    // Per-column words of a stripe for the significance and cleanup passes.
    // w0: significance window (18 bits) | candidate rows << 18 | significant
    // rows << 22; w1: zero-coding labels of rows 0..3 from the window at the
    // start of the stripe << 14 (4 bits each); w2: sign window (18 bits, the
    // layout of the significance window).
    __device__ __forceinline__ void column_words(int y0, uint32_t rowm, bool cleanup, uint32_t& w0, uint32_t& w1,
                                                 uint32_t& w2, uint64_t& cols) const {
        const uint32_t so = win6(sig, y0);
        const uint32_t sl = from_left(so), sr = from_right(so);
        const uint32_t slice = spread3(sl) | (spread3(so) << 1) | (spread3(sr) << 2);
        const uint32_t s4 = (so >> 1) & 15u;
        const uint32_t p4 = (uint32_t)(pip >> y0) & 15u;
        uint32_t cand;
        if (cleanup) {
            cand = ~s4 & ~p4 & rowm;
        } else {
            uint32_t nz = 0;
#pragma unroll
            for (int r = 0; r < 4; r++) nz |= (((slice >> (3 * r)) & 0x1EFu) != 0u ? 1u : 0u) << r;
            cand = ~s4 & nz & rowm;
        }
        if ((int)threadIdx.x >= w) cand = 0;  // columns outside the block
        w0 = slice | (cand << 18) | (s4 << 22);
        const uint32_t go = win6(sgn, y0);
        w2 = spread3(from_left(go)) | (spread3(go) << 1) | (spread3(from_right(go)) << 2);
        w1 = 0;
        cols = __ballot(cand != 0u);
        if (cols) {  // zero-coding labels of the four rows from the window as it is now
            uint32_t labs = 0;
#pragma unroll
            for (int r = 0; r < 4; r++) labs |= zc_lab_lane((slice >> (3 * r)) & 0x1FFu, cls) << (4 * r);
            w1 |= labs << 14;
        }
    }

    // significance propagation over one stripe (T.800 D.3.1)
    __device__ __forceinline__ void spp(int y0, int nr) {
        const uint32_t rowm = (1u << nr) - 1u;
        uint32_t vw0, vw1, vw2;
        uint64_t cols;
        column_words(y0, rowm, false, vw0, vw1, vw2, cols);
        uint32_t res = 0;  // lane x: coded rows | new significance << 4 | signs << 8
        if (cols) spp_stripe(m, cs, cl, ctxv, cols, vw0, vw1, vw2, res, rowm, (uint32_t)w, zcv, scv, trm, trl);
        const uint64_t nv = (uint64_t)(res & 15u) << y0, ns = (uint64_t)((res >> 4) & 15u) << y0;
        pip |= nv;
        sig |= ns;
        ones |= ns;
        sgn |= (uint64_t)((res >> 8) & 15u) << y0;
    }

    // magnitude refinement over one stripe (T.800 D.3.3)
    __device__ __forceinline__ void mrp(int y0, int nr) {
        const uint32_t rowm = (1u << nr) - 1u;
        const uint32_t so = win6(sig, y0);
        const uint32_t sl = from_left(so), sr = from_right(so);
        const uint32_t slice = spread3(sl) | (spread3(so) << 1) | (spread3(sr) << 2);
        const uint32_t s4 = (so >> 1) & 15u;
        const uint32_t p4 = (uint32_t)(pip >> y0) & 15u;
        const uint32_t r4 = (uint32_t)(ref >> y0) & 15u;
        uint32_t nz = 0;
#pragma unroll
        for (int r = 0; r < 4; r++) nz |= (((slice >> (3 * r)) & 0x1EFu) != 0u ? 1u : 0u) << r;
        const uint32_t cand = (int)threadIdx.x < w ? s4 & ~p4 & rowm : 0u;
        // the candidates' labels in scan order, 4 bits each (label - 14: 2 if
        // refined before, else 1 with a significant neighbour, else 0), 15 ends;
        // one label for all of them also as label | 4 << 24 | (count - 1) << 28
        uint32_t stream = 0, first = 0;
        bool same = true;
        int k = 0;
#pragma unroll
        for (int r = 0; r < 4; r++)
            if ((cand >> r) & 1u) {
                const uint32_t code = ((r4 >> r) & 1u) ? 2u : ((nz >> r) & 1u);
                if (k == 0) first = code;
                same = same && code == first;
                stream |= code << (4 * k);
                k++;
            }
        stream |= 15u << (4 * k);
        if (k > 0 && same) stream |= (first | 4u | ((uint32_t)(k - 1) << 4)) << 24;
        uint64_t cols = __ballot(cand != 0u);
        uint32_t res = 0;  // lane x: column x's decisions, first candidate highest
        if (cols) mq_mrp_stripe(m, cs, cl, ctxv, cols, stream, res, trm, trl);
        // back to rows: candidate j of n (scan order) holds bit n-1-j
        const int n = __popc(cand);
        uint32_t one = 0;
        int j = 0;
#pragma unroll
        for (int r = 0; r < 4; r++)
            if ((cand >> r) & 1u) {
                one |= ((res >> (n - 1 - j)) & 1u) << r;
                j++;
            }
        ref |= (uint64_t)cand << y0;
        ones |= (uint64_t)one << y0;
    }

    // cleanup with run-length mode over one stripe (T.800 D.3.4)
    __device__ __forceinline__ void cup(int y0, int nr) {
        const uint32_t rowm = (1u << nr) - 1u;
        uint32_t vw0, vw1, vw2;
        uint64_t cols;
        column_words(y0, rowm, true, vw0, vw1, vw2, cols);
        uint32_t res = 0;  // lane x: new significance | signs << 8
        if (cols) cup_stripe(m, cs, cl, ctxv, st_rl, st_uni, cols, vw0, vw1, vw2, res, zcv, scv, trm, trl);
        const uint64_t ns = (uint64_t)(res & 15u) << y0;
        sig |= ns;
        ones |= ns;
        sgn |= (uint64_t)((res >> 8) & 15u) << y0;
    }
};

template <typename M>
__device__ __forceinline__ void decode_block(const BlockArgs d, const uint32_t* __restrict__ arena_w,
                                             uint32_t arena_words, uint32_t* __restrict__ coef,
                                             M (*planes)[64], M* sgncols, uint32_t zcv,
                                             uint32_t scv, uint32_t trm, uint32_t trl, uint32_t ctx0) {
    const int lane = threadIdx.x;
    const int w = d.w, h = d.h;
    const int numbps = d.numbps, npasses = d.npasses;

    Blk B;
    B.w = w;
    B.h = h;
    B.cls = d.cls;
    B.ctxv = ctx0;
    B.cl = 0;
    B.cs = rl(ctx0, 0);
    B.st_rl = st_pack(3, 0);
    B.st_uni = st_pack(46, 0);
    B.zcv = zcv;
    B.scv = scv;
    B.trm = trm;
    B.trl = trl;
    B.sig = B.sgn = B.pip = B.ref = B.ones = 0;
    const uint32_t w0 = d.data_off >> 2;
    // bytes past the arena read as 0xFF, like bytes past the codeword
    const uint32_t avail = arena_words > w0 ? (arena_words - w0) * 4u : 0u;
    mq_init(B.m, arena_w + w0, d.data_len < avail ? d.data_len : avail);

    for (int k = 0; k < npasses; k++) {
        const int p = numbps - 1 - (k + 2) / 3;
        if (p < 0) break;
        const int type = k == 0 ? 2 : (k - 1) % 3;  // 0 sigprop, 1 magref, 2 cleanup
        for (int y0 = 0; y0 < h; y0 += 4) {
            const int nr = h - y0 < 4 ? h - y0 : 4;
            if (type == 0)
                B.spp(y0, nr);
            else if (type == 1)
                B.mrp(y0, nr);
            else
                B.cup(y0, nr);
        }
        // a plane ends after its cleanup pass, or when the passes run out
        if (type == 2 || k == npasses - 1) {
            planes[p][lane] = (M)B.ones;
            B.ones = 0;
            B.pip = 0;
        }
    }
    sgncols[lane] = (M)B.sgn;
    __syncthreads();

    // rebuild, dequantize and store the block's rows
    const int kl = npasses - 1;
    int plast = numbps - 1 - (kl + 2) / 3;
    if (plast < 0) plast = 0;
    const int tlast = kl == 0 ? 2 : (kl - 1) % 3;
    const uint32_t sigonly = 1u << plast;
    int span = 1;
    while (span < w) span <<= 1;
    const int rows_per = 64 / span;
    const int lx = lane % span, ly = lane / span;
    uint32_t* out = coef + d.out_off;
    if (lx < w) {
        const M sc = sgncols[lx];
        for (int y = ly; y < h; y += rows_per) {
            uint32_t mag = 0;
            for (int pp = plast; pp < numbps; pp++) mag |= (uint32_t)((planes[pp][lx] >> y) & 1u) << pp;
            uint32_t v = 0;
            if (mag) {
                const bool neg = ((sc >> y) & 1u) != 0;
                const int q = (tlast == 0 && mag != sigonly) ? plast + 1 : plast;
                if ((d.mode & TC_MODE_IRREVERSIBLE) == 0) {
                    const int32_t iv = (int32_t)(mag + (q > 0 ? (1u << (q - 1)) : 0u));
                    v = (uint32_t)(neg ? -iv : iv);
                } else {
                    const float f = ((float)mag + ldexpf(0.5f, q)) * d.delta;
                    v = __float_as_uint(neg ? -f : f);
                }
            }
            out[(uint32_t)y * d.out_pitch + lx] = v;
        }
    }
    __syncthreads();
}

// M: type of the per-plane column words kept in LDS (32 bits for blocks up to
// 32x32); NPL: bit-planes a block may carry (deeper blocks are left to the
// row-mask kernel); NL: the widest block.  Every lane writes its plane
// word (lanes past the block hold zeros), so no store is divergent.  LDS per
// work-group is (NPL+1)*64*sizeof(uint64_t): 8.3 KiB small, 10.5 KiB large.
template <typename M, int NPL, int NL>
__global__ __launch_bounds__(64) void t1_stripe_kernel(const tc_cblk_desc* __restrict__ descs,
                                                       const uint32_t* __restrict__ order, int n,
                                                       const uint32_t* __restrict__ arena_w, uint32_t arena_words,
                                                       uint32_t* __restrict__ coef) {
    __shared__ M planes[NPL][64];
    __shared__ M sgncols[64];
    const int lane = threadIdx.x;
    // lane tables: zero-coding labels for the 9-bit windows lane*8 .. lane*8+7
    // (4 bits each, one word per band class), sign entries for the sign
    // indices lane*4 .. lane*4+3 (8 bits each), MQ successor states
    uint32_t zc3[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        uint32_t v = 0;
        for (int b = 0; b < 8; b++) {
            const uint32_t i9 = (uint32_t)(lane * 8 + b);
            const uint32_t n8 = (i9 & 0xFu) | ((i9 >> 5) << 4);  // drop the centre bit
            v |= zc_label(c, n8) << (4 * b);
        }
        zc3[c] = v;
    }
    uint32_t scv = 0;
    for (int b = 0; b < 4; b++) {
        // key (TC_SIGN_UPDATE) -> the sign-coding table's index: significance
        // and sign of left, right, upper, lower neighbour, bits 0..7
        const uint32_t k = (uint32_t)(lane * 4 + b);
        const uint32_t i = ((k >> 3) & 1u) | (((k >> 2) & 1u) << 1) | (((k >> 5) & 1u) << 2) |
                           (((k >> 4) & 1u) << 3) | (((k >> 1) & 1u) << 4) | ((k & 1u) << 5) |
                           (((k >> 7) & 1u) << 6) | (((k >> 6) & 1u) << 7);
        scv |= sc_entry(i) << (8 * b);
    }
    const uint32_t trm = lane < 47 ? st_pack(kNmps[lane], 0) : 0u;
    const uint32_t trl =
        lane < 47 ? (st_pack(kNlps[lane], 0) | ((lane == 0 || lane == 6 || lane == 14) ? 1u : 0u)) : 0u;
    const uint32_t ctx0 = st_pack(lane == 0 ? 4 : 0, 0);
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        const uint32_t id = order ? uni(order[b]) : (uint32_t)b;
        const BlockArgs d = load_args(descs + id);
        if (d.numbps <= 0 || d.npasses <= 0 || d.numbps > NPL || d.w <= 0 || d.h <= 0 || d.w > NL ||
            d.h > (int)sizeof(uint64_t) * 8)
            continue;
        const uint32_t zcv = d.cls == 0 ? zc3[0] : (d.cls == 1 ? zc3[1] : zc3[2]);
        decode_block<M>(d, arena_w, arena_words, coef, planes, sgncols, zcv, scv, trm, trl, ctx0);
    }
}

}  // namespace

int t1_stripe_max_planes(bool small) { return small ? 32 : 20; }

hipError_t launch_t1_stripe(hipStream_t s, bool small, int grid, const tc_cblk_desc* descs, const uint32_t* order,
                            int n, const uint8_t* arena, uint32_t arena_bytes, uint32_t* coef) {
    if (n <= 0) return hipSuccess;
    if (grid > n) grid = n;
    if (grid < 1) grid = 1;
    if (small)
        hipLaunchKernelGGL((t1_stripe_kernel<uint32_t, 32, 32>), dim3(grid), dim3(64), 0, s, descs, order, n,
                           (const uint32_t*)arena, arena_bytes / 4, coef);
    else
        hipLaunchKernelGGL((t1_stripe_kernel<uint64_t, 20, 64>), dim3(grid), dim3(64), 0, s, descs, order, n,
                           (const uint32_t*)arena, arena_bytes / 4, coef);
    return hipGetLastError();
}

}  // namespace tc
