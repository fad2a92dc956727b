// Tier-1 EBCOT decoding with 64 code-blocks per wavefront advanced in
// lockstep: every step, each lane decodes exactly one MQ symbol of its own
// block, so one VALU instruction serves up to 64 blocks.
//
// Reference behaviour: g_decode, TC/coefficient_coder.cl:778-819 (pass order),
// the MQ decoder TC/coefficient_coder.cl:105-252, context formation
// :330-469 and the three pass functors :471-586, with the by-value no-op
// setters repaired and the pass count honoured (ITU-T T.800 Annex C, D).
//
// Why lockstep.  A lane that runs a block's nested pass / stripe / column /
// row loops on its own diverges from its neighbours at every branch, and a
// wave then issues the union of all lanes' paths.  Here the pass and stripe
// loops are wave-uniform (the host packs blocks with equal pass counts into a
// wave) and the work inside one stripe of one pass is flattened into a
// sequence of single-symbol steps.  A step is: pick the next coefficient
// (find-first-set over the stripe's candidate row masks) or continue a
// pending multi-symbol sequence (ZC 1 -> sign; run-length 1 -> two uniform
// symbols -> sign), form the context, decode one symbol, apply it.  All lanes
// execute the same instruction stream; only the lane masks differ.
//
// State.  A 64x64 block needs 4 bits per coefficient (significance, sign,
// coded-in-this-plane's-significance-pass, refined-once): 2 KiB, 128 KiB for
// 64 blocks, which LDS cannot hold for more than one wave per CU.  The full
// state therefore lives in a global scratch area laid out
// [wave][mask][row][lane] (each row load or store of the 64 lanes is one
// contiguous 512-byte access) and only the rows of the current stripe and
// its two halo rows are staged into LDS ([row][lane], conflict-free) while
// the stripe is decoded.  Context states (19 bytes per lane) and the MQ / zero-coding
// / sign-coding tables also live in LDS.
//
// Output.  Scattered read-modify-writes of single bits would saturate the
// memory system once a few waves share a SIMD, so a coefficient's word in the
// zero-filled plane is written once, by a plain store, when it turns
// significant (its top magnitude bit and sign).  The 1-bits of each
// refinement pass are collected per stripe as row masks and stored to a
// per-pass log in the scratch area ([wave][pass][row][lane], coalesced); after
// the decode the wave merges the logs into the words and dequantizes them in
// place (convert_lockstep_block).
//
// Code-block styles (descriptor mode bits): RESET re-initialises every context
// after each pass, CAUSAL drops the row below the stripe from the staged rows
// (so no context of the stripe sees it), SEGSYM decodes the four uniform
// symbols that follow each cleanup pass.
//
// Codewords are staged through a per-lane ring of 8 dwords in LDS, topped up
// (for every lane of the branch running short) before a lane reads past it,
// so no global load is in flight across symbol steps.  The MQ decoder keeps
// T.800's C register in the upper half of a 64-bit register and takes the
// codeword four bytes at a time into the lower half, so renormalisation is
// one shift, with no byte-in inside it (lane_fill states why this equals the
// byte-wise BYTEIN).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../tc_kernels.h"
#include "t1_common.hpp"

namespace tc {
namespace {

constexpr int kLockLabels = 19;  // 0-16 ZC/SC/MR, 17 run-length, 18 uniform
constexpr int kLaneCount = 64;

// Pending symbol of a multi-symbol sequence.
enum : uint32_t { PEND_NONE = 0, PEND_SC = 1, PEND_UNI1 = 2, PEND_UNI2 = 3 };

// Staged stripe rows, index into LdsLock::rows: significance rows y0-1..y0+4,
// then sign rows y0-1..y0+4.  A refinement pass stages no state: significance
// is fixed during it, so its context labels are formed as masks at stripe
// entry; the same rows then hold its decoded 1-bits (RO), its label-16
// candidates (RE) and label-15 candidates (RP), rows y0..y0+3 each.  The
// coded-in-this-pass rows are never read inside a pass and stay in registers.
constexpr int RS = 0, RX = 6, NROWS = 12;
constexpr int RO = 0, RE = 4, RP = 8;
// Masks of the global scratch state.
constexpr int GS = 0, GX = 1, GP = 2, GR = 3;

// A context state (T.800 C.2: index I and MPS) is one byte, st = I + 47 MPS,
// which is also its entry in tq: Qe(I) in bits 16-31, the two state bytes a
// renormalisation can lead to (MPS path NMPS(I) + 47 MPS in bits 0-6, LPS
// path NLPS(I) + 47 (MPS ^ SWITCH(I)) in bits 7-13) and MPS in bit 14, so
// the update is two field extracts.  The 94 entries fall in different LDS
// banks except entries 64 apart (two different states), where I | MPS << 6
// had put both MPS values of every index in one bank.  Each lane's 19 states
// sit in a 20-byte record (lane-major: the 64 lanes' accesses fall in 64
// different banks).  Bytes instead of packed 32-bit states take 3.6 KB off a
// wave's LDS, and the waves resident per compute unit are bound by LDS.
constexpr int kCtxStride = 20;
constexpr int kStates = 47;

__device__ __forceinline__ uint32_t tq_entry(int st) {
    const uint32_t mps = st >= kStates ? 1u : 0u;
    const int i = st - (int)mps * kStates;
    const uint32_t sw = (i == 0 || i == 6 || i == 14) ? 1u : 0u;
    return ((uint32_t)kQe[i] << 16) | ((uint32_t)kNmps[i] + mps * kStates) |
           (((uint32_t)kNlps[i] + (mps ^ sw) * kStates) << 7) | (mps << 14);
}
__device__ __forceinline__ uint8_t ctx_init(int c) { return (uint8_t)(c == 0 ? 4 : (c == 17 ? 3 : (c == 18 ? 46 : 0))); }

template <typename M>
struct LdsLock {
    uint32_t tq[2 * kStates];
    uint8_t ctx[kLaneCount][kCtxStride];
    M rows[NROWS][kLaneCount];
    uint32_t ring[9][kLaneCount];  // codeword dwords, slot = dword index mod 8; slot 8 repeats slot 0
    uint8_t zc[3][512];  // label by 3x3 significance window (centre ignored)
    uint8_t sc[256];     // sign label | xor<<4 by (sig,sign) of up,left,right,down
    int kmax, hmax;
};

// Per-lane MQ decoder reading its codeword from the LDS ring.
struct LaneMq {
    uint32_t A;      // MQ: T.800's A register in bits 31..16
    uint64_t C;      // MQ: T.800's C register in bits 63..32, codeword bytes taken ahead below
    int lsb;         // MQ: bit of C holding the lowest bit of the last byte taken
    int CT;          // raw segments: bits of `cur` still unread
    uint32_t bp, len, cur;  // cur: the byte at bp, the last one taken (0xFF past a marker or the end)
    uint32_t raw;    // 1: the current codeword segment is raw (BYPASS), C unused
    uint32_t base;   // arena byte offset of the codeword
    uint32_t rend;   // arena byte offset one past the ring's last byte
    const uint32_t* ring;  // this lane's column of the ring (stride kLaneCount)
};

// Codeword byte i (the ring must hold it); past the end the decoder sees 0xFF.
__device__ __forceinline__ uint32_t lane_byte(const LaneMq& m, uint32_t i) {
    const uint32_t a = m.base + i;
    const uint32_t b = (m.ring[((a >> 2) & 7) * kLaneCount] >> ((a & 3) * 8)) & 0xFFu;
    return i < m.len ? b : 0xFFu;
}

// Loads the 16-byte-aligned 32 bytes holding codeword byte bp (clamped to the
// arena's last 32 bytes; nwords is a multiple of 4) into the ring; at least
// 16 bytes past bp are then held.
__device__ __forceinline__ void lane_refill(LaneMq& m, uint32_t* ring_col, const uint32_t* __restrict__ words,
                                           uint32_t nwords) {
    uint32_t q0 = ((m.base + m.bp) >> 2) & ~3u;
    if (q0 > nwords - 8) q0 = nwords - 8;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4* src = reinterpret_cast<const u32x4*>(words + q0);
    u32x4 v[2];
#pragma unroll
    for (int j = 0; j < 2; j++) v[j] = __builtin_nontemporal_load(src + j);
    // q0 is a multiple of 4, so the first four dwords fill the ring half
    // starting at slot q0 & 4 and the next four the other half
#pragma unroll
    for (int j = 0; j < 2; j++) {
        uint32_t* const h = ring_col + (((q0 & 4u) ^ (4u * j)) * kLaneCount);
        h[0 * kLaneCount] = v[j].x;
        h[1 * kLaneCount] = v[j].y;
        h[2 * kLaneCount] = v[j].z;
        h[3 * kLaneCount] = v[j].w;
    }
    ring_col[8 * kLaneCount] = (q0 & 4u) ? v[1].x : v[0].x;  // slot 0 again
    m.rend = (q0 + 8) * 4;
}

// Takes codeword bytes into C below the last one, as many whole bytes as
// there is room for (lsb >= 8 on entry; at most four, lsb >= 0 after).
//
// T.800's BYTEIN (Figure C.19) adds the byte after B at 8 bits below it, or
// 7 after B = 0xFF; after 0xFF a byte above 0x8F is a marker, which feeds
// 0xFF at 8-bit spacing without advancing, as does the end of the codeword.
// Taking a byte earlier than BYTEIN would adds it into bits that are still
// zero, so the bits compared with Qe (63..48) are unchanged, with one
// exception: a byte of 0x80-0x8F after 0xFF overlaps the 0xFF's lowest bit
// and carries through it.  BYTEIN takes that byte when the 0xFF's lowest bit
// has been shifted past bit 48, so it is held back until then.  The common
// case (k = min(4, lsb / 8) bytes, no 0xFF among the last byte and the next
// k - 1, all k inside the codeword) takes them with one add; after a marker
// or the end, 8k one-bits.  The rest goes byte by byte.  Bytes up to lsb <= 48 are always
// held when Qe is compared, so every decision equals T.800's.
__device__ __forceinline__ void lane_fill(LaneMq& m, uint32_t* ring_col, const uint32_t* __restrict__ words,
                                         uint32_t nwords) {
    if (__ballot(m.rend - (m.base + m.bp) < 8u)) {
        if (m.rend - (m.base + m.bp) < 16u) lane_refill(m, ring_col, words, nwords);
    }
    const uint32_t a = m.base + m.bp + 1;
    const uint32_t* const rw = m.ring + ((a >> 2) & 7) * kLaneCount;  // slots s and s + 1 (slot 8 = slot 0)
    const uint32_t w0 = rw[0];
    const uint32_t w1 = rw[kLaneCount];
    const uint32_t le = __builtin_amdgcn_alignbyte(w1, w0, a & 3);  // bytes bp+1.. bp+4, the first lowest
    const uint32_t k = min(4u, (uint32_t)m.lsb >> 3);              // bytes there is room for
    const uint32_t nq = ~((le << 8) | m.cur);                       // the last byte and the next three, inverted
    // a zero byte among the low k of nq (borrows only run upwards, so the
    // test is exact for the low bytes)
    const uint32_t lowk = (uint32_t)((1ull << (8 * k)) - 1u);
    const bool noff = ((nq - 0x01010101u) & ~nq & 0x80808080u & lowk) == 0u;
    const bool inside = m.bp + k < m.len;
    const bool marker = m.cur == 0xFFu && (m.bp + 1 >= m.len || (le & 0xFFu) > 0x8Fu);
    if ((noff && inside) || marker) {
        // the first k bytes, placed below bit lsb
        const uint32_t word =
            (marker ? 0xFFFFFFFFu : __builtin_bswap32(le)) & (uint32_t)(0xFFFFFFFF00000000ull >> (8 * k));
        m.C += ((uint64_t)word << 32) >> (64 - m.lsb);
        m.lsb -= (int)(8 * k);
        m.bp += marker ? 0u : k;
        m.cur = marker ? 0xFFu : ((le >> (8 * k - 8)) & 0xFFu);
    } else {
        for (int i = 0; i < 4; i++) {
            if (m.lsb < 8) break;  // no room for another byte
            const uint32_t b = lane_byte(m, m.bp + 1);
            if (m.cur == 0xFFu) {
                if (b > 0x8Fu) {  // marker: 0xFF, no advance
                    m.lsb -= 8;
                    m.C += (uint64_t)0xFFu << m.lsb;
                    continue;
                }
                if (b >= 0x80u && m.lsb <= 48) break;  // carry byte: not before BYTEIN would take it
                m.lsb -= 7;
            } else {
                m.lsb -= 8;
            }
            m.C += (uint64_t)b << m.lsb;
            m.bp += 1;
            m.cur = b;
        }
    }
}

// INITDEC, T.800 Figure C.20 (the ring must hold the first bytes): the first
// byte goes to bits 55..62 (after INITDEC's shift by 7), the rest follow.
__device__ __forceinline__ void lane_init(LaneMq& m, uint32_t* ring_col, const uint32_t* __restrict__ words,
                                         uint32_t nwords) {
    m.bp = 0;
    m.cur = lane_byte(m, 0);
    m.C = (uint64_t)m.cur << 55;
    m.lsb = 55;
    lane_fill(m, ring_col, words, nwords);
    m.A = 0x8000u << 16;
}

// One raw bit (BYPASS segments, T.800 D.6): MSB first; the byte after 0xFF
// holds 7 bits below a stuffed 0; past the segment the bytes read 0xFF.  cur
// is the byte at bp and CT the bits of it still unread.
__device__ __forceinline__ uint32_t lane_raw(LaneMq& m, uint32_t* ring_col, const uint32_t* __restrict__ words,
                                            uint32_t nwords) {
    if (m.CT == 0) {
        if (m.rend - (m.base + m.bp) < 8u) lane_refill(m, ring_col, words, nwords);
        const uint32_t prev = m.cur;
        m.bp += 1;
        m.cur = lane_byte(m, m.bp);
        m.CT = prev == 0xFFu ? 7 : 8;
    }
    m.CT -= 1;
    return (m.cur >> m.CT) & 1u;
}

// Codeword segments of the BYPASS / TERMALL styles: pass k (0 = the first
// cleanup) ends one, and SPP/MRP passes from the fifth bit-plane on are raw.
__device__ __forceinline__ bool lock_seg_end(uint32_t style, int k) {
    if (style & TC_MODE_TERMALL) return true;
    return (style & TC_MODE_BYPASS) && (k == 9 || (k > 9 && (k - 1) % 3 != 0));
}
__device__ __forceinline__ bool lock_raw_pass(uint32_t style, int k) {
    return (style & TC_MODE_BYPASS) && k >= 10 && (k - 1) % 3 != 2;
}

// DECODE, T.800 Figures C.15-C.18, on a context state byte (I + 47 MPS),
// with the codeword taken ahead by lane_fill (lsb <= 48 on entry): RENORMD is
// one shift.
__device__ __forceinline__ uint32_t lane_decode(LaneMq& m, uint32_t& st, const uint32_t* tq) {
    // Nearly every step has some lane that renormalises, so the exchange is
    // computed for all lanes with selects rather than behind a branch.
    // A and Qe are held in bits 31..16, so C's top half compares with Qe
    // directly and the renormalisation shift is the leading zero count.
    const uint32_t t = tq[st];
    const uint32_t qe = t & 0xFFFF0000u;
    const uint32_t mps = (t >> 14) & 1u;
    uint32_t A = m.A - qe;
    const uint32_t chi = (uint32_t)(m.C >> 32);
    const bool lps = chi < qe;
    m.C = ((uint64_t)(lps ? chi : chi - qe) << 32) | (uint32_t)m.C;
    const bool renorm = lps || A < 0x80000000u;
    const bool take_mps = lps == (A < qe);  // LPS exchange: A < Qe gives MPS; MPS exchange: the reverse
    const uint32_t smps = t & 127u;
    const uint32_t slps = (t >> 7) & 127u;
    const uint32_t pick = take_mps ? ~0u : 0u;  // bitwise select keeps this off the branch path
    const uint32_t st2 = (smps & pick) | (slps & ~pick);
    const uint32_t d = (renorm && !take_mps) ? (mps ^ 1u) : mps;
    st = renorm ? st2 : st;
    A = lps ? qe : A;
    const int n = __builtin_clz(A);  // 0 unless renormalising (A > 0: Qe >= 1 and A - Qe > 0)
    m.A = A << n;
    m.C <<= n;
    m.lsb += n;
    return d;
}

// Bits x-1, x, x+1 of m (x-1 in bit 0; columns outside the row read 0),
// without a branch, so that the row loads feeding it issue together.
template <typename M>
__device__ __forceinline__ uint32_t col3(M m, int x) {
    const int sh = x > 0 ? x - 1 : 0;
    return ((uint32_t)(m >> sh) << (x == 0 ? 1 : 0)) & 7u;
}

// The positions of one stripe row that a significance pass coded: not
// significant at stripe entry, inside the block (valid), and with a
// significant neighbour when the scan reached them.  A neighbour earlier in
// the scan (in the row above, the one to the left, the row below's left one)
// counts with its state after the stripe (*_end), a later one (the row
// above's right one, the one to the right, the row below's centre and right)
// with its state at entry (*0).  Column x of a mask is bit x.
template <typename M>
__device__ __forceinline__ M sigprop_coded(M up_end, M up0, M row_end, M row0, M dn_end, M dn0, M valid) {
    return ~row0 & valid &
           ((up_end << 1) | up_end | (up0 >> 1) | (row_end << 1) | (row0 >> 1) | (dn_end << 1) | dn0 | (dn0 >> 1));
}

// OR into a lane's own LDS word (no return value: one LDS instruction).
__device__ __forceinline__ void lds_or(uint32_t* p, uint32_t v) { (void)atomicOr(p, v); }
__device__ __forceinline__ void lds_or(uint64_t* p, uint64_t v) {
    (void)atomicOr(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

// Bit x of four row masks as a 4-bit column nibble (row 0 in bit 0).
template <typename M>
__device__ __forceinline__ uint32_t nibble(M a0, M a1, M a2, M a3, int x) {
    return ((uint32_t)(a0 >> x) & 1u) | (((uint32_t)(a1 >> x) & 1u) << 1) | (((uint32_t)(a2 >> x) & 1u) << 2) |
           (((uint32_t)(a3 >> x) & 1u) << 3);
}

// Bits x and x+1 of four row masks as two column nibbles (row 0 in bit 0):
// n0 for column x, n1 for column x+1 (0 past the mask).
template <typename M>
__device__ __forceinline__ void nibble2(M a0, M a1, M a2, M a3, int x, uint32_t& n0, uint32_t& n1) {
    const uint32_t t0 = (uint32_t)(a0 >> x), t1 = (uint32_t)(a1 >> x), t2 = (uint32_t)(a2 >> x),
                   t3 = (uint32_t)(a3 >> x);
    n0 = (t0 & 1u) | ((t1 & 1u) << 1) | ((t2 & 1u) << 2) | ((t3 & 1u) << 3);
    n1 = ((t0 >> 1) & 1u) | (t1 & 2u) | ((t2 << 1) & 4u) | ((t3 << 2) & 8u);
}

// Decodes one stripe of one pass for every lane with work, one symbol per
// lane per step.  TYPE: 0 significance propagation (T.800 D.3.1), 1 magnitude
// refinement (D.3.3), 2 cleanup with run-length mode (D.3.4).
//
// c0..c3 hold the candidate columns of the stripe's rows.  A lane walks them
// in scan order with a column cursor: `cany` holds the columns still to visit;
// taking the lowest one loads its rows as a nibble `cur`, whose set bits are
// then popped one per step.  TYPE 0: a coefficient turning significant adds
// its later neighbours that are inside the block (rows vrows of the stripe,
// columns below wcols) and not significant: the row below in this column to
// `cur`, three rows of the next column to the nibble `nxt` that joins that
// column's rows when the walk reaches it (and the column to `cany`, so the
// walk's next column is that one).  c0..c3 are not changed; the positions
// coded in the pass are found from the stripe's rows afterwards
// (sigprop_coded).
// TYPE 1: the staged rows RE mark candidates refined before (label 16), RP
// the others with a significant neighbour (label 15; the rest take 14); each
// decoded 1-bit is ORed into the staged rows RO, returned as row masks in
// o0..o3.  TYPE 2: sany is the OR of the
// six staged significance rows (run-length test).  out points at row y0.
// RAW: some lane of the wave is in a raw (BYPASS) segment; without it the
// raw paths are not compiled into the symbol loop.
template <typename M, int TYPE, bool RAW>
__device__ __forceinline__ void stripe_symbols(LdsLock<M>& L, int lane, LaneMq& m, uint32_t* ring_col,
                                               const uint32_t* __restrict__ words, uint32_t nwords, bool work, M& c0,
                                               M& c1, M& c2, M& c3, M e0, M e1, M e2, M e3, M p0, M p1, M p2, M p3,
                                               M& o0, M& o1, M& o2, M& o3, M& sany, uint32_t cls,
                                               uint32_t* __restrict__ out, uint32_t pitch, uint32_t pbit,
                                               uint32_t vrows, int wcols) {
    auto lr = [&](int i) -> M& { return L.rows[i][lane]; };
    M cany = c0 | c1 | c2 | c3;
    uint32_t pend = PEND_NONE, cur = 0;
    uint32_t nxt = 0;  // TYPE 0: rows of the next column that joined
    int x = 0, r = 0;
    // (a do-while, tested at the bottom: with the test at the top the compiler
    // copied the lane state twice per step in the refinement loop)
    if (__ballot(work)) do {
        if (work) {
            // Codeword bytes: whenever one lane's register runs low (a
            // decision needs the bits down to 48), every lane tops up, so the
            // lanes' fills fall on the same steps and the wave runs the fill
            // far less often than once per lane.
            const bool raw = RAW && m.raw;
            if (__ballot(!raw && m.lsb > 48)) {
                if (!raw && m.lsb >= 8) lane_fill(m, ring_col, words, nwords);
            }
            bool rl = false;
            if (pend == PEND_NONE) {
                if (cur == 0) {
                    // next candidate column
                    x = (int)__builtin_ctzll((unsigned long long)cany);
                    cany &= cany - 1;
                    cur = nibble<M>(c0, c1, c2, c3, x);
                    if (TYPE == 0) {
                        cur |= nxt;
                        nxt = 0;
                    }
                    if (TYPE == 2) rl = cur == 15u && col3<M>(sany, x) == 0;
                }
                r = __builtin_ctz(cur);
                cur &= cur - 1;
            }
            const M bx = (M)1 << x;
            uint32_t lab, sxor = 0;
            M s1 = 0, x1 = 0;  // TYPE 0/2: significance and sign rows r+1 as read for the labels
            uint32_t w9 = 0;
            if (TYPE == 1) {
                lab = 14u + ((uint32_t)(lr(RP + r) >> x) & 1u) + (((uint32_t)(lr(RE + r) >> x) & 1u) << 1);
            } else {
                // Zero-coding and sign-coding labels from one read of the
                // staged rows, for every lane: nearly every step has lanes of
                // both kinds, and two branches would read the rows twice.
                s1 = lr(RS + r + 1);
                x1 = lr(RX + r + 1);
                w9 = col3<M>(lr(RS + r), x) | (col3<M>(s1, x) << 3) | (col3<M>(lr(RS + r + 2), x) << 6);
                const uint32_t xw = col3<M>(lr(RX + r), x) | (col3<M>(x1, x) << 3) | (col3<M>(lr(RX + r + 2), x) << 6);
                const uint32_t lz = L.zc[cls][w9];
                const uint32_t e = L.sc[(w9 & 0xAAu) | ((xw & 0xAAu) >> 1)];
                lab = pend == PEND_SC ? (e & 15u) : (pend == PEND_NONE ? (rl ? 17u : lz) : 18u);  // 17 run-length, 18 uniform
                sxor = e >> 4;
            }
            uint32_t st = L.ctx[lane][lab];
            const uint32_t bit = raw ? lane_raw(m, ring_col, words, nwords) : lane_decode(m, st, L.tq);
            L.ctx[lane][lab] = (uint8_t)st;
            uint32_t* const at = out + (uint32_t)r * pitch + x;
            if (TYPE == 1) {
                if (bit) lds_or(&lr(RO + r), bx);
            } else {
                // The decision's effect on the staged rows: a zero-coding 1
                // or a sign 1 sets its bit with one LDS OR.
                const bool zc = pend == PEND_NONE && !rl;
                const bool sig = zc && bit != 0u;
                const bool sgn = pend == PEND_SC;
                // a raw (BYPASS) segment carries the sign bit itself (T.800 D.6)
                const uint32_t sg = bit ^ (raw ? 0u : sxor);
                if (sig) lds_or(&lr(RS + r + 1), bx);
                if (sgn && sg) lds_or(&lr(RX + r + 1), bx);
                if (TYPE == 2) sany |= sig ? bx : (M)0;
                if (sgn) *at = pbit | (sg << 31);  // the word's first bits: the plane was zero here
                const uint32_t pend0 = pend;
                pend = sig ? PEND_SC : (sgn ? PEND_NONE : pend);
                if (TYPE == 0 && sig) {
                    // neighbours later in scan order may now be coded in this
                    // pass: the row below in this column (it joins the column's
                    // nibble) and rows r-1..r+1 of column x+1 (they join the
                    // next column's), where inside the block and not
                    // significant.  Those positions come later in the scan, so
                    // the window read for the labels holds their significance
                    // at stripe entry.
                    const uint32_t add0 = ((~w9 >> 7) & 1u) << (r + 1) & vrows & 15u;
                    const uint32_t n1 = ((w9 >> 2) & 1u) | ((w9 >> 4) & 2u) | ((w9 >> 6) & 4u);  // column x+1, rows r-1..r+1
                    const uint32_t add1 = x + 1 < wcols ? (((~n1 & 7u) << r) >> 1) & vrows & 15u : 0u;
                    cur |= add0;
                    nxt |= add1;
                    cany |= add1 ? (bx << 1) : (M)0;
                }
                if (TYPE == 2 && !zc && !sgn) {
                    if (rl) {
                        if (bit)
                            pend = PEND_UNI1;
                        else
                            cur = 0;  // the whole column stays insignificant
                    } else if (pend0 == PEND_UNI1) {
                        r = (int)(bit << 1);
                        pend = PEND_UNI2;
                    } else {
                        r |= (int)bit;
                        cur &= ~((2u << r) - 1u);  // rows above r were coded as insignificant
                        lds_or(&lr(RS + r + 1), bx);
                        sany |= bx;
                        pend = PEND_SC;
                    }
                }
            }
            work = pend != PEND_NONE || cur != 0 || cany != 0;
        }
    } while (__ballot(work));
    if (TYPE == 1) {
        o0 = lr(RO + 0);
        o1 = lr(RO + 1);
        o2 = lr(RO + 2);
        o3 = lr(RO + 3);
    }
}

// g: the wave's state rows [mask][row][lane]; lg: its refinement logs
// [pass][row][lane], one per refinement pass (pass k = 3j + 2 writes log j);
// a row holds one entry per block of the wave (stride bpw).
template <typename M, int MAXH>
__device__ void decode_lockstep(LdsLock<M>& L, const tc_cblk_desc& d, bool active, int lane,
                                const uint32_t* __restrict__ words, uint32_t nwords, uint32_t* __restrict__ coef,
                                M* __restrict__ g, M* __restrict__ lg, int bpw) {
    constexpr int NB = (int)sizeof(M) * 8;
    const int w = active ? d.w : 0, h = active ? d.h : 0;
    const int numbps = active ? d.numbps : 0, npasses = active ? d.npasses : 0;
    const uint32_t cls = active ? d.cls : 0;
    const uint32_t style = active ? (d.mode & TC_MODE_STYLE) : 0u;
    const M VM = (w >= NB) ? ~(M)0 : (((M)1 << w) - 1);
    uint32_t* const out = coef + (active ? d.out_off : 0);
    const uint32_t pitch = active ? d.out_pitch : 0;

    for (int c = 0; c < kLockLabels; c++) L.ctx[lane][c] = ctx_init(c);
    LaneMq m;
    uint32_t* const ring_col = &L.ring[0][lane];
    m.ring = ring_col;
    m.base = active ? d.data_off : 0;
    m.len = active ? d.data_len : 0;
    m.bp = 0;
    m.raw = 0;
    // BYPASS / TERMALL: the segment length table sits before the codeword
    const bool multi = (style & (TC_MODE_BYPASS | TC_MODE_TERMALL)) != 0;
    uint32_t seg = 0, nseg = 0, tab = 0;
    if (multi) {
        const uint32_t w0 = d.data_off >> 2;
        nseg = w0 > 0 ? words[w0 - 1] : 0u;
        if (nseg >= w0) nseg = 0;  // a table that cannot fit: no segments (bytes read as 0xFF)
        tab = w0 - 1 - nseg;
        m.len = nseg > 0 ? words[tab] : 0u;
    }
    if (active) {
        lane_refill(m, ring_col, words, nwords);
        lane_init(m, ring_col, words, nwords);
    }
    asm volatile("" : "+v"(m.len));  // the length's load is waited for here, not in the symbol loop

    auto gat = [&](int mask, int row) -> M& { return g[(mask * MAXH + row) * bpw + lane]; };
    auto lr = [&](int i) -> M& { return L.rows[i][lane]; };

    M carry_s = 0, carry_x = 0;  // rows y0-1 of the previous stripe
    const int kmax = L.kmax, hmax = L.hmax;
    for (int k = 0; k < kmax; k++) {
        const int type = k == 0 ? 2 : (k - 1) % 3;  // 0 sigprop, 1 magref, 2 cleanup (wave-uniform)
        const int p = numbps - 1 - (k + 2) / 3;
        const bool pass_on = active && k < npasses && p >= 0;
        const uint32_t pbit = pass_on ? (1u << p) : 0u;
        if (pass_on && multi && k > 0 && lock_seg_end(style, k - 1)) {
            // the next codeword segment: raw, or the MQ decoder re-initialised
            m.base += m.len;
            seg += 1;
            m.len = seg < nseg ? words[tab + seg] : 0u;
            // Consume the loaded length here: otherwise the wait for its
            // load lands on its first use, inside the symbol loop, where it
            // runs every step and waits for the coefficient stores in flight.
            asm volatile("" : "+v"(m.len));
            m.bp = 0;
            lane_refill(m, ring_col, words, nwords);
            if (lock_raw_pass(style, k)) {
                m.raw = 1;
                m.cur = lane_byte(m, 0);
                m.CT = 8;
            } else {
                m.raw = 0;
                lane_init(m, ring_col, words, nwords);
            }
        }
        // With 32-column masks the state rows of a stripe are loaded while
        // the stripe before it decodes: rows y0+4.. belong to the next
        // stripe, which this pass has not reached, so they hold the previous
        // pass's state.  (The 32x32 class's launch 21.1 -> 20.2 ms on the
        // mixed batch; with 64-column masks the held rows take the kernel
        // from 114 to 190 VGPRs and the 64x64 launch gained nothing, so it
        // loads each stripe's rows when the stripe starts.)
        constexpr bool kAhead = sizeof(M) == 4;
        M pS[5], pX[5], pP[4], pR[4];
        auto fetch = [&](int ys) {
#pragma unroll
            for (int r = 0; r < 5; r++) {
                // stripe-causal contexts take the row below the stripe as insignificant
                const bool ok = pass_on && k > 0 && ys + r < h && !(r == 4 && (style & TC_MODE_CAUSAL));
                pS[r] = ok ? gat(GS, ys + r) : (M)0;
                pX[r] = (ok && type != 1) ? gat(GX, ys + r) : (M)0;
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const bool ok = pass_on && k > 0 && ys + r < h;
                pP[r] = (ok && type != 0) ? gat(GP, ys + r) : (M)0;
                pR[r] = (ok && type == 1) ? gat(GR, ys + r) : (M)0;
            }
        };
        if (kAhead) fetch(0);
        for (int y0 = 0; y0 < hmax; y0 += 4) {
            const bool on = pass_on && y0 < h;
            if (!kAhead) fetch(y0);
            M c0 = 0, c1 = 0, c2 = 0, c3 = 0;  // candidate columns per stripe row
            M e0 = 0, e1 = 0, e2 = 0, e3 = 0;  // sigprop: significance at entry; magref: label-16 candidates
            M p0 = 0, p1 = 0, p2 = 0, p3 = 0;  // magref: label-15 candidates
            M sany = 0;                        // OR of the six significance rows
            M o0 = 0, o1 = 0, o2 = 0, o3 = 0;  // magref: 1-bits decoded in the stripe
            if (on) {
                M S[6], X[6], P[4], R[4];
                S[0] = y0 > 0 ? carry_s : (M)0;
                X[0] = (y0 > 0 && type != 1) ? carry_x : (M)0;
#pragma unroll
                for (int r = 0; r < 5; r++) {
                    S[r + 1] = pS[r];
                    X[r + 1] = pX[r];
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    P[r] = pP[r];
                    R[r] = pR[r];
                }
                if (type != 1) {
#pragma unroll
                    for (int i = 0; i < 6; i++) {
                        lr(RS + i) = S[i];
                        lr(RX + i) = X[i];
                        sany |= S[i];
                    }
                }
                M cr[4], er[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const M vr = (y0 + r < h) ? VM : (M)0;
                    if (type == 0) {
                        er[r] = S[r + 1];  // significance at entry, for sigprop_coded
                        cr[r] = ~S[r + 1] & vr & spread<M>(S[r] | S[r + 1] | S[r + 2]);
                    } else if (type == 1) {
                        cr[r] = S[r + 1] & ~P[r];
                        er[r] = cr[r] & R[r];
                    } else {
                        er[r] = 0;
                        cr[r] = ~S[r + 1] & ~P[r] & vr;
                    }
                }
                c0 = cr[0], c1 = cr[1], c2 = cr[2], c3 = cr[3];
                e0 = er[0], e1 = er[1], e2 = er[2], e3 = er[3];
                if (type == 1) {
                    // significance does not change in a refinement pass: every label is known now
                    M pr[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const M nb = spread<M>(S[r] | S[r + 2]) | (S[r + 1] << 1) | (S[r + 1] >> 1);
                        pr[r] = cr[r] & ~R[r] & nb;
                        // after the pass every coefficient significant before it has been refined
                        gat(GR, y0 + r) = cr[r];
                        lr(RO + r) = (M)0;
                        lr(RE + r) = er[r];
                        lr(RP + r) = pr[r];
                    }
                    p0 = pr[0], p1 = pr[1], p2 = pr[2], p3 = pr[3];
                    carry_s = S[4];
                }
            }

            if (kAhead && y0 + 4 < hmax) fetch(y0 + 4);
            const bool work = on && ((c0 | c1 | c2 | c3) != 0);
            const uint32_t vrows = (y0 < h ? 1u : 0u) | (y0 + 1 < h ? 2u : 0u) | (y0 + 2 < h ? 4u : 0u) | (y0 + 3 < h ? 8u : 0u);
            // raw segments occur in significance and refinement passes only
            const bool anyraw = type != 2 && __ballot(work && m.raw) != 0;
#define TC_STRIPE(T, R)                                                                                              \
    stripe_symbols<M, T, R>(L, lane, m, ring_col, words, nwords, work, c0, c1, c2, c3, e0, e1, e2, e3, p0, p1, p2, p3, \
                            o0, o1, o2, o3, sany, cls, out + (uint32_t)y0 * pitch, pitch, pbit, vrows, w)
            if (type == 0) {
                if (anyraw)
                    TC_STRIPE(0, true);
                else
                    TC_STRIPE(0, false);
                // c0..c3 become the positions coded in the pass; the rows
                // above and below the stripe have not changed
                const M E0 = lr(RS + 0), E1 = lr(RS + 1), E2 = lr(RS + 2), E3 = lr(RS + 3), E4 = lr(RS + 4),
                        E5 = lr(RS + 5);
                auto vrow = [&](int r) -> M { return (y0 + r < h) ? VM : (M)0; };
                c0 = sigprop_coded<M>(E0, E0, E1, e0, E2, e1, vrow(0));
                c1 = sigprop_coded<M>(E1, e0, E2, e1, E3, e2, vrow(1));
                c2 = sigprop_coded<M>(E2, e1, E3, e2, E4, e3, vrow(2));
                c3 = sigprop_coded<M>(E3, e2, E4, e3, E5, E5, vrow(3));
            } else if (type == 1) {
                if (anyraw)
                    TC_STRIPE(1, true);
                else
                    TC_STRIPE(1, false);
            } else {
                TC_STRIPE(2, false);
            }
#undef TC_STRIPE

            if (on) {
                if (type != 1) {
                    const M pr[4] = {c0, c1, c2, c3};  // sigprop: every position coded in the pass
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        gat(GS, y0 + i) = lr(RS + i + 1);
                        gat(GX, y0 + i) = lr(RX + i + 1);
                        gat(GP, y0 + i) = type == 0 ? pr[i] : (M)0;
                        // the first pass also clears the refined rows the scratch holds
                        if (k == 0) gat(GR, y0 + i) = (M)0;
                    }
                    carry_s = lr(RS + 4);
                    carry_x = lr(RX + 4);
                } else {
                    M* const lj = lg + (size_t)((k - 2) / 3) * MAXH * bpw;
                    lj[(y0 + 0) * bpw + lane] = o0;
                    lj[(y0 + 1) * bpw + lane] = o1;
                    lj[(y0 + 2) * bpw + lane] = o2;
                    lj[(y0 + 3) * bpw + lane] = o3;
                }
            }
        }
        if (pass_on && type == 2 && (style & TC_MODE_SEGSYM)) {
            // segmentation symbol: four uniform-context decisions after each
            // cleanup pass (decoded to keep the MQ state in step, not checked)
            for (int s4 = 0; s4 < 4; s4++) {
                if (m.lsb >= 32) lane_fill(m, ring_col, words, nwords);
                uint32_t st = L.ctx[lane][18];
                (void)lane_decode(m, st, L.tq);
                L.ctx[lane][18] = (uint8_t)st;
            }
        }
        if (pass_on && (style & TC_MODE_RESET))
            for (int c = 0; c < kLockLabels; c++) L.ctx[lane][c] = ctx_init(c);
    }
}

// Finishes block j of the wave: merges its refinement logs (log entry
// [j'][row][j] holds plane numbps-2-j') into the words its significance
// stores left, then rebuilds and dequantizes each coefficient as
// convert_block does (midpoint of the last coded plane).  The block's rows go
// in groups whose logs fit the wave's LDS stage area (cap entries): the
// group's log rows are loaded into LDS with one load per lane per entry
// batch, then the 64 lanes cover 64/span rows per step (span = w rounded up
// to a power of two), each lane with RU rows in flight, reading the log bits
// from LDS.  (Reading the logs from global memory per row group had cost one
// memory round trip per four rows: 0.4 ms of a 16-block wave.)
template <typename M, int MAXH>
__device__ void convert_lockstep_block(int lane, int j, const tc_cblk_desc& d, uint32_t* __restrict__ coef,
                                       const M* __restrict__ lg, int bpw, M* stage, int cap) {
    constexpr int RU = 16;
    const int numbps = d.numbps, np = d.npasses;
    if (numbps <= 0 || np <= 0) return;
    const int kl = np - 1;
    int plast = numbps - 1 - (kl + 2) / 3;
    if (plast < 0) plast = 0;
    const int tlast = kl == 0 ? 2 : (kl - 1) % 3;
    const int nlog = np / 3;  // refinement passes the block went through
    uint32_t* const out = coef + d.out_off;
    const int w = d.w, h = d.h;
    const uint32_t sigonly = 1u << plast;
    int span = 1;
    while (span < w) span <<= 1;
    const int rows_per = 64 / span;
    const int lx = lane % span, ly = lane / span;
    // rows per group: a multiple of rows_per whose logs fit the stage area
    int G = h;
    if (nlog > 0 && nlog * G > cap) G = cap / nlog;
    if (G > rows_per) G -= G % rows_per;
    if (G < 1) G = 1;
    for (int g0 = 0; g0 < h; g0 += G) {
        const int gn = min(G, h - g0);
        __syncthreads();  // the previous group's reads are done
        for (int i = lane; i < nlog * gn; i += 64) {
            const int t = i / gn, y = g0 + i % gn;
            stage[i] = lg[((size_t)t * MAXH + y) * bpw + j];
        }
        __syncthreads();
        if (lx >= w) continue;
        for (int y0 = g0 + ly; y0 < g0 + gn; y0 += RU * rows_per) {
            uint32_t v[RU], mag[RU];
            int yy[RU];
#pragma unroll
            for (int q = 0; q < RU; q++) {
                yy[q] = y0 + q * rows_per;
                const int yc = yy[q] < g0 + gn ? yy[q] : y0;  // rows past the group re-read row y0 (in bounds, unused)
                v[q] = out[(uint32_t)yc * d.out_pitch + lx];
            }
            int yo[RU];
#pragma unroll
            for (int q = 0; q < RU; q++) {
                yo[q] = (yy[q] < g0 + gn ? yy[q] : y0) - g0;
                mag[q] = v[q] & 0x7FFFFFFFu;
            }
            for (int t = 0; t < nlog; t++) {
                const M* const st = stage + t * gn;
                const int sh = numbps - 2 - t;
#pragma unroll
                for (int q = 0; q < RU; q++) mag[q] |= (uint32_t)((st[yo[q]] >> lx) & 1u) << sh;
            }
#pragma unroll
            for (int q = 0; q < RU; q++) {
                if (yy[q] >= g0 + gn || v[q] == 0) continue;
                const int qp = (tlast == 0 && mag[q] != sigonly) ? plast + 1 : plast;
                const bool neg = (v[q] >> 31) != 0;
                uint32_t res;
                if ((d.mode & TC_MODE_IRREVERSIBLE) == 0) {
                    const int32_t iv = (int32_t)(mag[q] + (qp > 0 ? (1u << (qp - 1)) : 0u));
                    res = (uint32_t)(neg ? -iv : iv);
                } else {
                    const float f = ((float)mag[q] + ldexpf(0.5f, qp)) * d.delta;
                    res = __float_as_uint(neg ? -f : f);
                }
                out[(uint32_t)yy[q] * d.out_pitch + lx] = res;
            }
        }
    }
}

template <typename M, int MAXH>
__global__ __launch_bounds__(64) void t1_lockstep_kernel(const tc_cblk_desc* __restrict__ descs,
                                                         const uint32_t* __restrict__ order, int n,
                                                         const uint8_t* __restrict__ arena, uint32_t arena_words,
                                                         uint32_t* __restrict__ coef, M* __restrict__ scratch,
                                                         int nlog_max, int bpw) {
    __shared__ LdsLock<M> L;
    const int lane = threadIdx.x;
    for (int i = lane; i < 2 * kStates; i += 64) L.tq[i] = tq_entry(i);
    for (int i = lane; i < 3 * 512; i += 64) {
        const uint32_t w9 = (uint32_t)(i & 511);
        L.zc[i >> 9][w9] = (uint8_t)zc_label(i >> 9, (w9 & 0xFu) | ((w9 >> 1) & 0xF0u));
    }
    for (int i = lane; i < 256; i += 64) {
        // index bits: sgn_u, sig_u, sgn_l, sig_l, sgn_r, sig_r, sgn_d, sig_d (bit 0 up)
        const uint32_t u = (uint32_t)i;
        const uint32_t e = ((u >> 3) & 1u) | (((u >> 2) & 1u) << 1) | (((u >> 5) & 1u) << 2) | (((u >> 4) & 1u) << 3) |
                           (((u >> 1) & 1u) << 4) | ((u & 1u) << 5) | (((u >> 7) & 1u) << 6) | (((u >> 6) & 1u) << 7);
        L.sc[i] = (uint8_t)sc_entry(e);
    }
    if (lane == 0) {
        L.kmax = 0;
        L.hmax = 0;
    }
    __syncthreads();

    // bpw (64, 32 or 16) blocks per wave: fewer than 64 leaves lanes idle but
    // makes more waves when a launch has few per CU
    const int base = blockIdx.x * bpw;
    const bool have = lane < bpw && base + lane < n;
    tc_cblk_desc d;
    if (have) d = descs[order ? order[base + lane] : (uint32_t)(base + lane)];
    const bool active = have && d.numbps > 0 && d.npasses > 0;
    if (active) {
        atomicMax(&L.kmax, (int)d.npasses);
        atomicMax(&L.hmax, (int)d.h);
    }
    __syncthreads();
    // lanes past bpw hold no block: their state rows are never touched
    M* const g = scratch + (size_t)blockIdx.x * (4 + nlog_max) * MAXH * bpw;
    M* const lg = g + 4 * MAXH * bpw;
    decode_lockstep<M, MAXH>(L, d, active, lane, reinterpret_cast<const uint32_t*>(arena), arena_words, coef, g, lg,
                             bpw);

    // make every lane's stores visible, and drop stale L1 lines, before the
    // cooperative read-back
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
    __syncthreads();
    for (int j = 0; j < bpw && base + j < n; j++) {
        const tc_cblk_desc dj = descs[order ? order[base + j] : (uint32_t)(base + j)];
        convert_lockstep_block<M, MAXH>(lane, j, dj, coef, lg, bpw, &L.rows[0][0], NROWS * kLaneCount);
    }
}

template <typename M, int MAXH>
static hipError_t launch_lockstep_variant(hipStream_t s, const tc_cblk_desc* descs, const uint32_t* order, int n,
                                          const uint8_t* arena, uint32_t arena_bytes, uint32_t* coef, void* scratch,
                                          int max_passes, int bpw) {
    if (n <= 0) return hipSuccess;
    const int grid = (n + bpw - 1) / bpw;
    hipLaunchKernelGGL((t1_lockstep_kernel<M, MAXH>), dim3(grid), dim3(kLaneCount), 0, s, descs, order, n, arena,
                       arena_bytes / 4, coef, static_cast<M*>(scratch), max_passes / 3, bpw);
    return hipGetLastError();
}

}  // namespace

// Scratch bytes the lockstep kernel needs for n blocks of one size class with
// at most max_passes coding passes each, bpw blocks per wave: per wave, four
// state masks and one refinement log per refinement pass, each MAXH rows of
// bpw entries.
size_t t1_lockstep_scratch_bytes(bool cls32, int n, int max_passes, int bpw) {
    const size_t waves = (size_t)((n + bpw - 1) / bpw);
    const size_t rows = (size_t)(4 + max_passes / 3);
    return cls32 ? waves * rows * 32 * (size_t)bpw * sizeof(uint32_t) : waves * rows * 64 * (size_t)bpw * sizeof(uint64_t);
}

hipError_t launch_t1_lockstep(hipStream_t s, bool cls32, int bpw, const tc_cblk_desc* descs, const uint32_t* order,
                              int n, const uint8_t* arena, uint32_t arena_bytes, uint32_t* coef, void* scratch,
                              int max_passes) {
    if (cls32)
        return launch_lockstep_variant<uint32_t, 32>(s, descs, order, n, arena, arena_bytes, coef, scratch, max_passes,
                                                     bpw);
    return launch_lockstep_variant<uint64_t, 64>(s, descs, order, n, arena, arena_bytes, coef, scratch, max_passes,
                                                 bpw);
}

}  // namespace tc
