// JPEG 2000 Part-1 encoder (host C++), used to produce the synthetic
// codestreams the benchmarks and tests decode: 12- and 16-bit images and large
// batches that the available third-party encoder cannot make quickly.
//
// Reversible path (RCT + 5/3) with every coding pass kept, one quality layer,
// LRCP order, one precinct per resolution, any tile size.  Irreversible path
// (ICT + 9/7, expounded quantization with step sizes scaled to the synthesis
// gains) with post-compression rate control: every coding pass gets a
// distortion-reduction estimate, passes are truncated on the lower convex hull
// of each block's rate-distortion curve at a common slope found by bisection
// so the codestream meets a target size.  Truncated codewords are cut at the
// shortest length that still decodes the kept passes, checked by re-decoding
// their MQ symbols.
//
// The encoder mirrors the decode-side definitions of ITU-T T.800 (Annex B
// geometry, B.10 packet headers and tag trees, C.2 MQ encoder, D tier-1
// contexts, E.1 quantization, F.4 forward wavelets, G.2/G.3 colour transforms).
// It is the forward counterpart of the reference's decoder (the reference's own
// forward kernels are TC/dwt_f53.cl, TC/dwt_f97.cl, TC/preprocess_rct.cl,
// TC/preprocess_ict.cl).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace tce {
namespace {

// ---------------------------------------------------------------- MQ encoder

const uint16_t QE[47] = {0x5601, 0x3401, 0x1801, 0x0AC1, 0x0521, 0x0221, 0x5601, 0x5401, 0x4801, 0x3801,
                         0x3001, 0x2401, 0x1C01, 0x1601, 0x5601, 0x5401, 0x5101, 0x4801, 0x3801, 0x3401,
                         0x3001, 0x2801, 0x2401, 0x2201, 0x1C01, 0x1801, 0x1601, 0x1401, 0x1201, 0x1101,
                         0x0AC1, 0x09C1, 0x08A1, 0x0521, 0x0441, 0x02A1, 0x0221, 0x0141, 0x0111, 0x0085,
                         0x0049, 0x0025, 0x0015, 0x0009, 0x0005, 0x0001, 0x5601};
const uint8_t NMPS[47] = {1,  2,  3,  4,  5,  38, 7,  8,  9,  10, 11, 12, 13, 29, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24,
                          25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 45, 46};
const uint8_t NLPS[47] = {1,  6,  9,  12, 29, 33, 6,  14, 14, 14, 17, 18, 20, 21, 14, 14, 15, 16, 17, 18, 19, 19, 20, 21,
                          22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 46};
inline bool SWITCH(int i) { return i == 0 || i == 6 || i == 14; }

enum { CTX_RUN = 17, CTX_UNI = 18, NCTX = 19 };

struct MQState {
    uint32_t A, C;
    int CT;
    size_t nbytes;  // bytes in buf (buf[0] is the byte before the codeword)
    uint8_t I[NCTX], MPS[NCTX];
};

struct MQEncoder {
    MQState s;
    std::vector<uint8_t> buf;
    void init() {
        s.A = 0x8000;
        s.C = 0;
        s.CT = 12;
        buf.assign(1, 0);
        reset_contexts();
    }
    // initial context states, T.800 Table D.7 (also the RESET style after each pass)
    void reset_contexts() {
        std::memset(s.I, 0, sizeof s.I);
        std::memset(s.MPS, 0, sizeof s.MPS);
        s.I[0] = 4;
        s.I[CTX_RUN] = 3;
        s.I[CTX_UNI] = 46;
    }
    void byteout() {
        uint8_t& B = buf.back();
        if (B == 0xFF) {
            buf.push_back((uint8_t)(s.C >> 20));
            s.C &= 0xFFFFF;
            s.CT = 7;
        } else if (s.C < 0x8000000) {
            buf.push_back((uint8_t)(s.C >> 19));
            s.C &= 0x7FFFF;
            s.CT = 8;
        } else {
            B++;
            if (B == 0xFF) {
                s.C &= 0x7FFFFFF;
                buf.push_back((uint8_t)(s.C >> 20));
                s.C &= 0xFFFFF;
                s.CT = 7;
            } else {
                buf.push_back((uint8_t)((s.C >> 19) & 0xFF));
                s.C &= 0x7FFFF;
                s.CT = 8;
            }
        }
    }
    void renorm() {
        do {
            s.A <<= 1;
            s.C <<= 1;
            if (--s.CT == 0) byteout();
        } while (!(s.A & 0x8000));
    }
    void encode(int d, int cx) {
        int i = s.I[cx];
        uint32_t qe = QE[i];
        s.A -= qe;
        if (d == s.MPS[cx]) {
            if (!(s.A & 0x8000)) {
                if (s.A < qe)
                    s.A = qe;
                else
                    s.C += qe;
                s.I[cx] = NMPS[i];
                renorm();
            } else {
                s.C += qe;
            }
        } else {
            if (s.A < qe)
                s.C += qe;
            else
                s.A = qe;
            if (SWITCH(i)) s.MPS[cx] ^= 1;
            s.I[cx] = NLPS[i];
            renorm();
        }
    }
    // FLUSH (T.800 C.2.9); returns the codeword (without the leading byte).
    std::vector<uint8_t> flush() {
        uint32_t tempc = s.C + s.A;
        s.C |= 0xFFFF;
        if (s.C >= tempc) s.C -= 0x8000;
        s.C <<= s.CT;
        byteout();
        s.C <<= s.CT;
        byteout();
        std::vector<uint8_t> out(buf.begin() + 1, buf.end());
        if (!out.empty() && out.back() == 0xFF) out.pop_back();
        return out;
    }
    size_t bytes_so_far() const { return buf.size() - 1; }
    // registers back to INITENC after a termination; the contexts keep their
    // states (T.800 D.4.1: only RESET resets them)
    void restart() {
        s.A = 0x8000;
        s.C = 0;
        s.CT = 12;
        buf.assign(1, 0);
    }
};

// Raw (bypass) coding of a segment, T.800 D.6: bits MSB first; after a 0xFF
// byte the next byte carries 7 bits below a stuffed 0.
struct RawEncoder {
    std::vector<uint8_t> out;
    uint32_t cur = 0;
    int left = 8;  // bits still free in cur
    bool any = false;
    void init() {
        out.clear();
        cur = 0;
        left = 8;
        any = false;
    }
    void put(int bit) {
        if (left == 0) {
            out.push_back((uint8_t)cur);
            left = cur == 0xFF ? 7 : 8;
            cur = 0;
        }
        left--;
        cur |= (uint32_t)(bit & 1) << left;
        any = true;
    }
    // pads the last byte with 0s; a final 0xFF gets a 0x00 after it so the
    // segment never ends on 0xFF
    std::vector<uint8_t> flush() {
        if (any) out.push_back((uint8_t)cur);
        if (!out.empty() && out.back() == 0xFF) out.push_back(0);
        return out;
    }
    size_t bytes_so_far() const { return out.size() + (any ? 1 : 0); }
};

// Codeword segments of the BYPASS (1) and TERMALL (4) styles (T.800 D.6,
// Table D.9): pass k (0 = first cleanup) ends a segment, and raw_pass: pass k
// is coded raw.
inline bool seg_end(int style, int k) {
    if (style & 4) return true;
    if (style & 1) return k == 9 || (k > 9 && (k - 1) % 3 != 0);
    return false;
}
inline bool raw_pass(int style, int k) { return (style & 1) && k >= 10 && (k - 1) % 3 != 2; }

// Recorded in place of a context label where the RESET style reset every
// context; never a real label.
constexpr uint8_t kResetMark = 255;

// MQ decoder used only to validate truncation points.
struct MQCheck {
    const uint8_t* d;
    size_t len, bp;
    uint32_t A, C, cur;
    int CT;
    uint8_t I[NCTX], MPS[NCTX];
    uint32_t byte(size_t i) const { return i < len ? d[i] : 0xFFu; }
    void bytein() {
        if (cur == 0xFF) {
            uint32_t b1 = byte(bp + 1);
            if (b1 > 0x8F) {
                C += 0xFF00;
                CT = 8;
            } else {
                bp++;
                cur = b1;
                C += b1 << 9;
                CT = 7;
            }
        } else {
            bp++;
            cur = byte(bp);
            C += cur << 8;
            CT = 8;
        }
    }
    void init(const uint8_t* data, size_t n) {
        d = data;
        len = n;
        bp = 0;
        cur = byte(0);
        C = cur << 16;
        bytein();
        C <<= 7;
        CT -= 7;
        A = 0x8000;
        std::memset(I, 0, sizeof I);
        std::memset(MPS, 0, sizeof MPS);
        I[0] = 4;
        I[CTX_RUN] = 3;
        I[CTX_UNI] = 46;
    }
    int decode(int cx) {
        if (cx == kResetMark) {  // RESET style: contexts back to their initial states
            std::memset(I, 0, sizeof I);
            std::memset(MPS, 0, sizeof MPS);
            I[0] = 4;
            I[CTX_RUN] = 3;
            I[CTX_UNI] = 46;
            return 0;
        }
        int i = I[cx];
        uint32_t qe = QE[i];
        int dd;
        A -= qe;
        if ((C >> 16) < qe) {
            if (A < qe) {
                dd = MPS[cx];
                I[cx] = NMPS[i];
            } else {
                dd = 1 - MPS[cx];
                if (SWITCH(i)) MPS[cx] ^= 1;
                I[cx] = NLPS[i];
            }
            A = qe;
        } else {
            C -= qe << 16;
            if (A & 0x8000) return MPS[cx];
            if (A < qe) {
                dd = 1 - MPS[cx];
                if (SWITCH(i)) MPS[cx] ^= 1;
                I[cx] = NLPS[i];
            } else {
                dd = MPS[cx];
                I[cx] = NMPS[i];
            }
        }
        do {
            if (CT == 0) bytein();
            A <<= 1;
            C <<= 1;
            CT--;
        } while (!(A & 0x8000));
        return dd;
    }
};

// ---------------------------------------------------------------- tier-1 encoder

int zc_label(int cls, int h, int v, int d) {  // T.800 Table D.1
    if (cls == 1) std::swap(h, v);
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

int sign_label(int H, int V, int* x) {  // Tables D.2/D.3
    *x = 0;
    if (H < 0 || (H == 0 && V < 0)) {
        H = -H;
        V = -V;
        *x = 1;
    }
    return H ? (V > 0 ? 13 : (V == 0 ? 12 : 11)) : (V ? 10 : 9);
}

struct Pass {
    size_t end_bytes;     // codeword bytes emitted when the pass ended
    double dist_reduction;
    size_t nsym;          // symbols coded up to the end of the pass
};

struct BlockResult {
    std::vector<uint8_t> data;  // full codeword
    int numbps = 0;
    int npasses = 0;            // passes included after truncation
    std::vector<Pass> passes;
    std::vector<uint8_t> sym_ctx, sym_bit;  // every coded symbol (for truncation checks)
    std::vector<uint32_t> trunc_len;       // decodable length of each pass prefix
    // BYPASS / TERMALL: length of each codeword segment and the pass count at
    // its end; truncation only at segment ends
    std::vector<uint32_t> seglen;
    std::vector<int> seg_passes;
};

// Encodes w x h signed coefficients (integer quantization indices) of a band
// with orientation orient.  weight scales the distortion estimate (square of
// the synthesis norm times the step size squared).  style: code-block style
// bits RESET (2), CAUSAL (8) and SEGSYM (32) are honoured (T.800 D.5-D.7);
// PTERM (16) needs nothing beyond this encoder's standard flush.
void t1_encode(const int32_t* coef, int w, int h, int pitch, int orient, double weight, bool record, BlockResult& R,
               int style = 0) {
    const bool causal = (style & 8) != 0;
    int cls = orient == 1 ? 1 : (orient == 3 ? 2 : 0);
    int W = w + 2;
    std::vector<uint8_t> sig((size_t)W * (h + 2)), vis((size_t)W * (h + 2)), ref((size_t)W * (h + 2)),
        sgn((size_t)W * (h + 2));
    std::vector<uint32_t> mag((size_t)w * h);
    uint32_t maxm = 0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int32_t v = coef[(size_t)y * pitch + x];
            uint32_t m = (uint32_t)(v < 0 ? -(int64_t)v : v);
            mag[(size_t)y * w + x] = m;
            sgn[(size_t)(y + 1) * W + x + 1] = v < 0;
            maxm = std::max(maxm, m);
        }
    int numbps = 0;
    while (numbps < 31 && (maxm >> numbps)) numbps++;
    R.numbps = numbps;
    R.passes.clear();
    R.sym_ctx.clear();
    R.sym_bit.clear();
    if (!numbps) {
        R.npasses = 0;
        R.data.clear();
        return;
    }
    MQEncoder mq;
    mq.init();
    RawEncoder raw;
    const bool term = (style & (1 | 4)) != 0;
    bool raw_now = false;
    R.seglen.clear();
    R.seg_passes.clear();
    auto enc = [&](int bit, int cx) {
        if (raw_now) {
            raw.put(bit);
            return;
        }
        mq.encode(bit, cx);
        if (record) {
            R.sym_ctx.push_back((uint8_t)cx);
            R.sym_bit.push_back((uint8_t)bit);
        }
    };
    auto AT = [&](int x, int y) { return (size_t)(y + 1) * W + x + 1; };
    // CAUSAL: the row below a stripe is never a neighbour of the stripe's last row
    auto below = [&](int y) { return !(causal && (y & 3) == 3); };
    auto counts = [&](int x, int y, int& hh, int& vv, int& dd) {
        size_t i = AT(x, y);
        const bool b = below(y);
        hh = sig[i - 1] + sig[i + 1];
        vv = sig[i - W] + (b ? sig[i + W] : 0);
        dd = sig[i - W - 1] + sig[i - W + 1] + (b ? sig[i + W - 1] + sig[i + W + 1] : 0);
        return hh + vv + dd;
    };
    auto contrib = [&](size_t i) { return sig[i] ? (sgn[i] ? -1 : 1) : 0; };
    double dist = 0;
    auto code_sign = [&](int x, int y) {
        size_t i = AT(x, y);
        int H = contrib(i - 1) + contrib(i + 1), V = contrib(i - W) + (below(y) ? contrib(i + W) : 0);
        H = H > 0 ? 1 : (H < 0 ? -1 : 0);
        V = V > 0 ? 1 : (V < 0 ? -1 : 0);
        int xr;
        int lab = sign_label(H, V, &xr);
        // a raw (BYPASS) segment carries the sign bit itself: no context, no XOR (T.800 D.6)
        enc(raw_now ? sgn[i] : (sgn[i] ^ xr), lab);
    };
    // distortion reduction of revealing bit p of magnitude m (midpoint reconstruction)
    auto dsig = [&](uint32_t m, int p) {
        double e0 = (double)m;                                   // error before: reconstruct 0
        double rec = std::ldexp(1.0, p) * 1.5;                   // first estimate 1.5 * 2^p
        double lo = (double)((m >> p) << p) + std::ldexp(0.5, p);  // midpoint of revealed interval
        (void)rec;
        double e1 = (double)m - lo;
        return (e0 * e0 - e1 * e1) * weight;
    };
    auto dref = [&](uint32_t m, int p) {
        double before = (double)((m >> (p + 1)) << (p + 1)) + std::ldexp(0.5, p + 1);
        double after = (double)((m >> p) << p) + std::ldexp(0.5, p);
        double e0 = (double)m - before, e1 = (double)m - after;
        return (e0 * e0 - e1 * e1) * weight;
    };
    int npass = 3 * numbps - 2;
    for (int k = 0; k < npass; k++) {
        int p = numbps - 1 - (k + 2) / 3;
        int type = k == 0 ? 2 : (k - 1) % 3;
        dist = 0;
        raw_now = raw_pass(style, k);
        if (type == 0) {
            for (int y0 = 0; y0 < h; y0 += 4)
                for (int x = 0; x < w; x++)
                    for (int y = y0; y < y0 + 4 && y < h; y++) {
                        size_t i = AT(x, y);
                        int hh, vv, dd;
                        if (sig[i] || !counts(x, y, hh, vv, dd)) continue;
                        uint32_t m = mag[(size_t)y * w + x];
                        int bit = (m >> p) & 1;
                        enc(bit, zc_label(cls, hh, vv, dd));
                        if (bit) {
                            code_sign(x, y);
                            sig[i] = 1;
                            dist += dsig(m, p);
                        }
                        vis[i] = 1;
                    }
        } else if (type == 1) {
            for (int y0 = 0; y0 < h; y0 += 4)
                for (int x = 0; x < w; x++)
                    for (int y = y0; y < y0 + 4 && y < h; y++) {
                        size_t i = AT(x, y);
                        if (!sig[i] || vis[i]) continue;
                        int hh, vv, dd;
                        int lab = ref[i] ? 16 : (counts(x, y, hh, vv, dd) ? 15 : 14);
                        uint32_t m = mag[(size_t)y * w + x];
                        enc((m >> p) & 1, lab);
                        ref[i] = 1;
                        dist += dref(m, p);
                    }
        } else {
            for (int y0 = 0; y0 < h; y0 += 4)
                for (int x = 0; x < w; x++) {
                    int y = y0;
                    if (y0 + 4 <= h) {
                        bool rl = true;
                        for (int r = 0; r < 4 && rl; r++) {
                            int hh, vv, dd;
                            size_t i = AT(x, y0 + r);
                            if (sig[i] || vis[i] || counts(x, y0 + r, hh, vv, dd)) rl = false;
                        }
                        if (rl) {
                            int first = -1;
                            for (int r = 0; r < 4; r++)
                                if ((mag[(size_t)(y0 + r) * w + x] >> p) & 1) {
                                    first = r;
                                    break;
                                }
                            if (first < 0) {
                                enc(0, CTX_RUN);
                                continue;
                            }
                            enc(1, CTX_RUN);
                            enc(first >> 1, CTX_UNI);
                            enc(first & 1, CTX_UNI);
                            y = y0 + first;
                            code_sign(x, y);
                            sig[AT(x, y)] = 1;
                            dist += dsig(mag[(size_t)y * w + x], p);
                            y++;
                        }
                    }
                    for (; y < y0 + 4 && y < h; y++) {
                        size_t i = AT(x, y);
                        if (sig[i] || vis[i]) continue;
                        int hh, vv, dd;
                        counts(x, y, hh, vv, dd);
                        uint32_t m = mag[(size_t)y * w + x];
                        int bit = (m >> p) & 1;
                        enc(bit, zc_label(cls, hh, vv, dd));
                        if (bit) {
                            code_sign(x, y);
                            sig[i] = 1;
                            dist += dsig(m, p);
                        }
                    }
                }
            std::fill(vis.begin(), vis.end(), 0);
            if (style & 32) {  // SEGSYM: 1010 in the uniform context
                enc(1, CTX_UNI);
                enc(0, CTX_UNI);
                enc(1, CTX_UNI);
                enc(0, CTX_UNI);
            }
        }
        if (style & 2) {
            mq.reset_contexts();
            if (record) {
                R.sym_ctx.push_back(kResetMark);
                R.sym_bit.push_back(0);
            }
        }
        if (term) {
            size_t done = R.data.size() + (raw_now ? raw.bytes_so_far() : mq.bytes_so_far());
            if (seg_end(style, k) || k == npass - 1) {
                std::vector<uint8_t> seg = raw_now ? raw.flush() : mq.flush();
                R.data.insert(R.data.end(), seg.begin(), seg.end());
                R.seglen.push_back((uint32_t)seg.size());
                R.seg_passes.push_back(k + 1);
                done = R.data.size();
                if (k + 1 < npass) {
                    if (raw_pass(style, k + 1))
                        raw.init();
                    else
                        mq.restart();
                }
            }
            R.passes.push_back(Pass{done, dist, R.sym_ctx.size()});
            continue;
        }
        R.passes.push_back(Pass{mq.bytes_so_far(), dist, R.sym_ctx.size()});
    }
    if (term) {
        R.npasses = npass;
        return;
    }
    R.data = mq.flush();
    R.npasses = npass;
}

// Shortest prefix of the codeword that decodes the first `npass` passes: start
// from bytes emitted + 2 and grow until the recorded symbols decode.
uint32_t truncation_length(const BlockResult& R, int npass) {
    if (!R.seglen.empty()) {  // segment styles: whole segments up to the pass count
        uint32_t L = 0;
        for (size_t s = 0; s < R.seglen.size() && R.seg_passes[s] <= npass; s++) L += R.seglen[s];
        return L;
    }
    if (npass >= (int)R.passes.size()) return (uint32_t)R.data.size();
    size_t nsym = R.passes[npass - 1].nsym;
    size_t L = std::min(R.data.size(), R.passes[npass - 1].end_bytes + 2);
    for (;; L++) {
        if (L >= R.data.size()) return (uint32_t)R.data.size();
        if (L > 0 && R.data[L - 1] == 0xFF) continue;  // never end a segment on 0xFF
        MQCheck m;
        m.init(R.data.data(), L);
        bool ok = true;
        for (size_t s = 0; s < nsym && ok; s++) ok = m.decode(R.sym_ctx[s]) == R.sym_bit[s];
        if (ok) return (uint32_t)L;
    }
}

// ---------------------------------------------------------------- transforms

inline int64_t floor_div(int64_t a, int64_t b) {
    int64_t q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}
inline int64_t ceil_div(int64_t a, int64_t b) { return -floor_div(-a, b); }

inline int mirror(int i, int i0, int i1) {
    int per = 2 * (i1 - i0 - 1);
    int m = (i - i0) % per;
    if (m < 0) m += per;
    if (m >= i1 - i0) m = per - m;
    return i0 + m;
}

// forward 1-D 5/3 over x[i0..i1) (absolute indexing through xs[i - i0 + 4])
void fwd53(int32_t* xs, int i0, int i1) {
    int n = i1 - i0;
    if (n == 1) {
        if (i0 & 1) xs[4] *= 2;
        return;
    }
    auto X = [&](int i) -> int32_t& { return xs[i - i0 + 4]; };
    for (int i = i0 - 4; i < i0; i++) X(i) = X(mirror(i, i0, i1));
    for (int i = i1; i < i1 + 4; i++) X(i) = X(mirror(i, i0, i1));
    for (int i = i0 - 3; i < i1 + 3; i++)
        if (i & 1) X(i) -= (int32_t)floor_div((int64_t)X(i - 1) + X(i + 1), 2);
    for (int i = i0; i < i1; i++)
        if (!(i & 1)) X(i) += (int32_t)floor_div((int64_t)X(i - 1) + X(i + 1) + 2, 4);
}

const double kA = -1.586134342059924, kB = -0.052980118572961, kG = 0.882911075530934, kD = 0.443506852043971,
             kK = 1.230174104914001;

void fwd97(double* xs, int i0, int i1) {
    int n = i1 - i0;
    if (n == 1) {
        if (i0 & 1) xs[6] *= 2;
        return;
    }
    auto X = [&](int i) -> double& { return xs[i - i0 + 6]; };
    for (int i = i0 - 6; i < i0; i++) X(i) = X(mirror(i, i0, i1));
    for (int i = i1; i < i1 + 6; i++) X(i) = X(mirror(i, i0, i1));
    for (int i = i0 - 5; i < i1 + 5; i++)
        if (i & 1) X(i) += kA * (X(i - 1) + X(i + 1));
    for (int i = i0 - 4; i < i1 + 4; i++)
        if (!(i & 1)) X(i) += kB * (X(i - 1) + X(i + 1));
    for (int i = i0 - 3; i < i1 + 3; i++)
        if (i & 1) X(i) += kG * (X(i - 1) + X(i + 1));
    for (int i = i0 - 2; i < i1 + 2; i++)
        if (!(i & 1)) X(i) += kD * (X(i - 1) + X(i + 1));
    for (int i = i0; i < i1; i++) X(i) = (i & 1) ? X(i) * kK : X(i) / kK;
}

// Multi-level forward transform of a tile-component into Mallat layout.
template <typename T, void (*F)(T*, int, int), int M>
void fdwt2(T* a, int w, int h, int x0, int y0, int levels) {
    std::vector<T> line((size_t)std::max(w, h) + 2 * M + 2), tmp((size_t)w * h);
    for (int n = 1; n <= levels; n++) {
        int64_t s = (int64_t)1 << (n - 1), s2 = (int64_t)1 << n;
        int u0 = (int)ceil_div(x0, s), u1 = (int)ceil_div(x0 + w, s), v0 = (int)ceil_div(y0, s), v1 = (int)ceil_div(y0 + h, s);
        int lw = (int)(ceil_div(x0 + w, s2) - ceil_div(x0, s2)), lh = (int)(ceil_div(y0 + h, s2) - ceil_div(y0, s2));
        int rw = u1 - u0, rh = v1 - v0;
        if (rw <= 0 || rh <= 0) continue;
        for (int x = 0; x < rw; x++) {  // vertical
            for (int y = 0; y < rh; y++) line[y + M] = a[(size_t)y * w + x];
            F(line.data(), v0, v1);
            for (int y = 0; y < rh; y++) tmp[(size_t)y * rw + x] = line[y + M];
        }
        for (int y = 0; y < rh; y++) {  // horizontal
            for (int x = 0; x < rw; x++) line[x + M] = tmp[(size_t)y * rw + x];
            F(line.data(), u0, u1);
            for (int x = 0; x < rw; x++) tmp[(size_t)y * rw + x] = line[x + M];
        }
        for (int y = 0; y < rh; y++)  // deinterleave into quadrants
            for (int x = 0; x < rw; x++) {
                int u = u0 + x, v = v0 + y;
                int cx = (u & 1) ? lw + (int)(floor_div(u, 2) - floor_div(u0, 2)) : (int)(u / 2 - ceil_div(u0, 2));
                int cy = (v & 1) ? lh + (int)(floor_div(v, 2) - floor_div(v0, 2)) : (int)(v / 2 - ceil_div(v0, 2));
                a[(size_t)cy * w + cx] = tmp[(size_t)y * rw + x];
            }
    }
}

// ---------------------------------------------------------------- tier-2 writer

struct BitWriter {
    std::vector<uint8_t>& out;
    uint32_t cur = 0;
    int n = 0, max = 8;
    explicit BitWriter(std::vector<uint8_t>& o) : out(o) {}
    void bit(int b) {
        cur = (cur << 1) | (uint32_t)(b & 1);
        if (++n == max) emit();
    }
    void bits(uint32_t v, int k) {
        for (int i = k - 1; i >= 0; i--) bit((v >> i) & 1);
    }
    void emit() {
        out.push_back((uint8_t)cur);
        max = cur == 0xFF ? 7 : 8;
        cur = 0;
        n = 0;
    }
    void align() {
        if (n) {
            cur <<= (max - n);
            emit();
        }
        if (max == 7) emit();  // a header ending on 0xFF is followed by a stuffed byte
    }
};

struct TagTree {
    struct Lvl {
        int w, h;
        size_t base;
    };
    std::vector<Lvl> lv;
    std::vector<int> val, low;
    std::vector<uint8_t> known;
    void init(int w, int h) {
        lv.clear();
        size_t total = 0;
        int lw = w, lh = h;
        for (;;) {
            lv.push_back({lw, lh, total});
            total += (size_t)lw * lh;
            if (lw == 1 && lh == 1) break;
            lw = (lw + 1) / 2;
            lh = (lh + 1) / 2;
        }
        val.assign(total, 0x3FFFFFFF);
        low.assign(total, 0);
        known.assign(total, 0);
    }
    void set(int x, int y, int v) {  // leaf value; parents hold the minimum
        for (size_t l = 0; l < lv.size(); l++) {
            size_t q = lv[l].base + (size_t)y * lv[l].w + x;
            val[q] = std::min(val[q], v);
            x >>= 1;
            y >>= 1;
        }
    }
    void encode(BitWriter& bw, int x, int y, int thr) {
        size_t path[40];
        int nl = (int)lv.size();
        for (int l = 0; l < nl; l++) {
            path[l] = lv[l].base + (size_t)y * lv[l].w + x;
            x >>= 1;
            y >>= 1;
        }
        int lo = 0;
        for (int l = nl - 1; l >= 0; l--) {
            size_t q = path[l];
            if (low[q] < lo) low[q] = lo;
            while (!known[q] && low[q] < thr) {
                if (low[q] >= val[q]) {
                    bw.bit(1);
                    known[q] = 1;
                } else {
                    bw.bit(0);
                    low[q]++;
                }
            }
            lo = low[q];
            if (!known[q]) break;
        }
    }
};

void put16(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}
void put32(std::vector<uint8_t>& o, uint32_t v) {
    put16(o, v >> 16);
    put16(o, v & 0xFFFF);
}

void parallel_for(int n, int threads, const std::function<void(int)>& f) {
    if (threads <= 1 || n <= 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<int> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < std::min(threads, n); t++)
        pool.emplace_back([&]() {
            for (;;) {
                int i = next.fetch_add(1);
                if (i >= n) break;
                f(i);
            }
        });
    for (auto& th : pool) th.join();
}

struct CB {
    int band_idx, x0, y0, x1, y1;
    BlockResult r;
    double weight = 1.0;
    int keep = 0;            // passes kept
    uint32_t keep_len = 0;   // bytes kept
};

struct BandG {
    int orient, x0, y0, x1, y1, ox, oy, ncx, ncy, first, xcb, ycb, nb;
    int eps, mu;
    double delta;
};

struct TCG {
    int x0, y0, x1, y1;
    std::vector<std::vector<BandG>> res;  // [r][band]
};

struct Params {
    int w, h, ncomp, prec, levels, xcb, ycb, tw, th, mct, irreversible, threads;
    double target_bytes;  // irreversible: codestream size target (0: keep every pass)
    double base_step;     // irreversible: step for a unit-gain band
    int cbsty = 0;        // code-block style (RESET 2, CAUSAL 8, PTERM 16, SEGSYM 32)
    int packed = 0;       // packet headers: 0 in the packets, 1 PPT, 2 PPM (A.7.4, A.7.5)
    int packed_max = 0;   // largest payload of one PPT/PPM segment (0: 65532)
    int dx[4] = {1, 1, 1, 1}, dy[4] = {1, 1, 1, 1};  // component sub-sampling (XRsiz, YRsiz)
};

// L2 norm of the synthesis basis function of a band at level nb (orientation o),
// measured by running the inverse of an impulse through the 1-D transforms.
double synthesis_norm(int nb, int orient, bool rev) {
    if (nb == 0) return 1.0;
    int N = 32 << nb;
    auto norm1 = [&](bool high) {
        // inverse of an impulse at the middle of the low or high band of level nb,
        // followed by low-band syntheses of the coarser-to-finer levels
        std::vector<double> a(N, 0.0);
        int lowlen = N >> nb;
        int idx = high ? lowlen + (lowlen >> 1) : (lowlen >> 1);
        a[idx] = 1.0;
        for (int n = nb; n >= 1; n--) {
            int len = N >> (n - 1), l = (len + 1) / 2;
            std::vector<double> x(len + 12);
            for (int i = 0; i < len; i++) x[i + 6] = (i & 1) ? a[l + i / 2] : a[i / 2];
            // inverse 1-D lifting (T.800 F.3.8.2 or F.3.8.1 without rounding)
            auto X = [&](int i) -> double& { return x[i + 6]; };
            for (int i = -6; i < 0; i++) X(i) = X(mirror(i, 0, len));
            for (int i = len; i < len + 6; i++) X(i) = X(mirror(i, 0, len));
            if (!rev) {
                for (int i = -5; i < len + 5; i++) X(i) = (i & 1) ? X(i) / kK : X(i) * kK;
                for (int i = -4; i < len + 4; i++)
                    if (!(i & 1)) X(i) -= kD * (X(i - 1) + X(i + 1));
                for (int i = -3; i < len + 3; i++)
                    if (i & 1) X(i) -= kG * (X(i - 1) + X(i + 1));
                for (int i = -2; i < len + 2; i++)
                    if (!(i & 1)) X(i) -= kB * (X(i - 1) + X(i + 1));
                for (int i = 0; i < len; i++)
                    if (i & 1) X(i) -= kA * (X(i - 1) + X(i + 1));
            } else {
                for (int i = -2; i < len + 2; i++)
                    if (!(i & 1)) X(i) -= (X(i - 1) + X(i + 1) + 2) / 4.0;
                for (int i = 0; i < len; i++)
                    if (i & 1) X(i) += (X(i - 1) + X(i + 1)) / 2.0;
            }
            for (int i = 0; i < len; i++) a[i] = X(i);
        }
        double s = 0;
        for (double v : a) s += v * v;
        return std::sqrt(s);
    };
    bool hx = orient == 1 || orient == 3, hy = orient == 2 || orient == 3;
    return norm1(hx) * norm1(hy);
}

}  // namespace

// Encodes interleaved samples (uint8 when prec <= 8, else uint16).
int encode(const void* pixels, const Params& P, std::vector<uint8_t>& out) {
    if (P.ncomp < 1 || P.ncomp > 4 || P.prec < 1 || P.prec > 16 || P.levels < 0 || P.levels > 10 || P.xcb < 2 ||
        P.ycb < 2 || P.xcb > 10 || P.ycb > 10 || P.xcb + P.ycb > 12)
        return -1;
    const bool irr = P.irreversible != 0;
    const int C = P.ncomp, W = P.w, H = P.h;
    const int tw = P.tw > 0 ? P.tw : W, th = P.th > 0 ? P.th : H;
    const int ntx = (W + tw - 1) / tw, nty = (H + th - 1) / th;
    bool subsampled = false;
    for (int c = 0; c < C; c++) {
        if (P.dx[c] < 1 || P.dy[c] < 1 || P.dx[c] > 255 || P.dy[c] > 255) return -1;
        subsampled |= P.dx[c] != 1 || P.dy[c] != 1;
    }
    if (subsampled && P.mct) return -1;  // the colour transform needs equal components
    const bool mct = P.mct && C >= 3;
    const int G = 2;  // guard bits
    // quantization parameters per (level, orientation), shared by every tile
    struct Q {
        int eps, mu;
        double delta, weight;
    };
    std::vector<Q> qp(3 * P.levels + 1);
    for (int i = 0; i < (int)qp.size(); i++) {
        int nb = i == 0 ? P.levels : P.levels - (i - 1) / 3;
        int orient = i == 0 ? 0 : (i - 1) % 3 + 1;
        int gain = orient == 0 ? 0 : (orient == 3 ? 2 : 1);
        double norm = synthesis_norm(nb, orient, !irr);
        if (!irr) {
            qp[i] = {P.prec + gain + (mct ? 1 : 0), 0, 1.0, norm * norm};
        } else {
            double step = P.base_step / norm;
            int Rb = P.prec + gain;
            int e = Rb - (int)std::floor(std::log2(step));
            double mant = step / std::ldexp(1.0, Rb - e) - 1.0;
            int mu = (int)std::lround(mant * 2048.0);
            if (mu >= 2048) {
                mu = 0;
                e -= 1;
            }
            if (e < 0) e = 0;
            if (e > 31) e = 31;
            double delta = std::ldexp(1.0 + mu / 2048.0, Rb - e);
            qp[i] = {e, mu, delta, norm * norm * delta * delta};
        }
    }
    // tile geometry and code-blocks
    struct TileWork {
        std::vector<TCG> comps;
        std::vector<std::vector<CB>> cbs;  // per component
        std::vector<std::vector<int32_t>> coef;
        std::vector<uint8_t> data;
    };
    std::vector<TileWork> tiles(ntx * nty);
    auto tx0_ref = [&](int px) { return px * tw; };
    auto ty0_ref = [&](int py) { return py * th; };
    auto tx1_ref = [&](int px) { return std::min(W, (px + 1) * tw); };
    auto ty1_ref = [&](int py) { return std::min(H, (py + 1) * th); };
    for (int t = 0; t < ntx * nty; t++) {
        TileWork& T = tiles[t];
        int px = t % ntx, py = t / ntx;
        T.comps.resize(C);
        T.cbs.resize(C);
        T.coef.resize(C);
        for (int c = 0; c < C; c++) {
            TCG& g = T.comps[c];
            // tile-component rectangle, T.800 B-12
            g.x0 = (int)ceil_div(tx0_ref(px), P.dx[c]);
            g.y0 = (int)ceil_div(ty0_ref(py), P.dy[c]);
            g.x1 = (int)ceil_div(tx1_ref(px), P.dx[c]);
            g.y1 = (int)ceil_div(ty1_ref(py), P.dy[c]);
            const int x0 = g.x0, y0 = g.y0, x1 = g.x1, y1 = g.y1;
            g.res.resize(P.levels + 1);
            for (int r = 0; r <= P.levels; r++) {
                int nb = r == 0 ? P.levels : P.levels - r + 1;
                int64_t rs = (int64_t)1 << (P.levels - r);
                int rx0 = (int)ceil_div(x0, rs), ry0 = (int)ceil_div(y0, rs);
                (void)rx0;
                (void)ry0;
                int lw = 0, lh = 0;
                if (r > 0) {
                    int64_t ls = (int64_t)1 << (P.levels - r + 1);
                    lw = (int)(ceil_div(x1, ls) - ceil_div(x0, ls));
                    lh = (int)(ceil_div(y1, ls) - ceil_div(y0, ls));
                }
                int nbands = r == 0 ? 1 : 3;
                for (int k = 0; k < nbands; k++) {
                    BandG B;
                    B.orient = r == 0 ? 0 : k + 1;
                    B.nb = nb;
                    int xo = (B.orient == 1 || B.orient == 3), yo = (B.orient == 2 || B.orient == 3);
                    if (nb == 0) {
                        B.x0 = x0;
                        B.y0 = y0;
                        B.x1 = x1;
                        B.y1 = y1;
                    } else {
                        int64_t bs = (int64_t)1 << nb, hs = (int64_t)1 << (nb - 1);
                        B.x0 = (int)ceil_div(x0 - hs * xo, bs);
                        B.y0 = (int)ceil_div(y0 - hs * yo, bs);
                        B.x1 = (int)ceil_div(x1 - hs * xo, bs);
                        B.y1 = (int)ceil_div(y1 - hs * yo, bs);
                    }
                    B.ox = xo ? lw : 0;
                    B.oy = yo ? lh : 0;
                    B.xcb = std::min(P.xcb, r == 0 ? 15 : 14);
                    B.ycb = std::min(P.ycb, r == 0 ? 15 : 14);
                    int qi = r == 0 ? 0 : 3 * (r - 1) + k + 1;
                    B.eps = qp[qi].eps;
                    B.mu = qp[qi].mu;
                    B.delta = qp[qi].delta;
                    B.first = (int)T.cbs[c].size();
                    if (B.x1 > B.x0 && B.y1 > B.y0) {
                        int64_t cw = (int64_t)1 << B.xcb, ch = (int64_t)1 << B.ycb;
                        int gx0 = (int)floor_div(B.x0, cw), gy0 = (int)floor_div(B.y0, ch);
                        B.ncx = (int)(ceil_div(B.x1, cw) - gx0);
                        B.ncy = (int)(ceil_div(B.y1, ch) - gy0);
                        for (int j = 0; j < B.ncy; j++)
                            for (int i = 0; i < B.ncx; i++) {
                                CB cb;
                                cb.band_idx = r * 3 + k;
                                cb.x0 = (int)std::max<int64_t>(B.x0, (gx0 + i) * cw);
                                cb.y0 = (int)std::max<int64_t>(B.y0, (gy0 + j) * ch);
                                cb.x1 = (int)std::min<int64_t>(B.x1, (gx0 + i + 1) * cw);
                                cb.y1 = (int)std::min<int64_t>(B.y1, (gy0 + j + 1) * ch);
                                cb.weight = qp[qi].weight;
                                T.cbs[c].push_back(cb);
                            }
                    } else {
                        B.ncx = B.ncy = 0;
                    }
                    g.res[r].push_back(B);
                }
            }
        }
    }
    // colour transform + DWT + quantization per tile-component
    const uint8_t* p8 = (const uint8_t*)pixels;
    const uint16_t* p16 = (const uint16_t*)pixels;
    // component c's sample at its own coordinates (x, y): the source pixel at
    // (x * XRsiz, y * YRsiz) on the reference grid
    auto sample = [&](int x, int y, int c) -> int32_t {
        size_t i = ((size_t)y * P.dy[c] * W + (size_t)x * P.dx[c]) * C + c;
        return P.prec <= 8 ? (int32_t)p8[i] : (int32_t)p16[i];
    };
    const int32_t shift = 1 << (P.prec - 1);
    parallel_for(ntx * nty * C, P.threads, [&](int job) {
        int t = job / C, c = job % C;
        TileWork& T = tiles[t];
        const TCG& g = T.comps[c];
        int w = g.x1 - g.x0, h = g.y1 - g.y0;
        std::vector<int32_t>& co = T.coef[c];
        co.assign((size_t)w * h, 0);
        if (!irr) {
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) {
                    int gx = g.x0 + x, gy = g.y0 + y;
                    int32_t v;
                    if (mct && c < 3) {
                        int32_t R = sample(gx, gy, 0) - shift, Gc = sample(gx, gy, 1) - shift, Bc = sample(gx, gy, 2) - shift;
                        v = c == 0 ? (int32_t)floor_div((int64_t)R + 2 * Gc + Bc, 4) : (c == 1 ? Bc - Gc : R - Gc);
                    } else {
                        v = sample(gx, gy, c) - shift;
                    }
                    co[(size_t)y * w + x] = v;
                }
            fdwt2<int32_t, fwd53, 4>(co.data(), w, h, g.x0, g.y0, P.levels);
        } else {
            std::vector<double> f((size_t)w * h);
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) {
                    int gx = g.x0 + x, gy = g.y0 + y;
                    double v;
                    if (mct && c < 3) {
                        double R = sample(gx, gy, 0) - shift, Gc = sample(gx, gy, 1) - shift, Bc = sample(gx, gy, 2) - shift;
                        v = c == 0 ? 0.299 * R + 0.587 * Gc + 0.114 * Bc
                                   : (c == 1 ? -0.16875 * R - 0.33126 * Gc + 0.5 * Bc : 0.5 * R - 0.41869 * Gc - 0.08131 * Bc);
                    } else {
                        v = sample(gx, gy, c) - shift;
                    }
                    f[(size_t)y * w + x] = v;
                }
            fdwt2<double, fwd97, 6>(f.data(), w, h, g.x0, g.y0, P.levels);
            // quantize per band (deadzone, T.800 E.1)
            for (int r = 0; r <= P.levels; r++)
                for (const BandG& B : g.res[r])
                    for (int y = 0; y < B.y1 - B.y0; y++)
                        for (int x = 0; x < B.x1 - B.x0; x++) {
                            size_t i = (size_t)(B.oy + y) * w + B.ox + x;
                            double v = f[i] / B.delta;
                            co[i] = (int32_t)(v < 0 ? -std::floor(-v) : std::floor(v));
                        }
        }
    });
    // tier-1 over every code-block
    struct Ref {
        int t, c, k;
    };
    std::vector<Ref> all;
    for (int t = 0; t < (int)tiles.size(); t++)
        for (int c = 0; c < C; c++)
            for (int k = 0; k < (int)tiles[t].cbs[c].size(); k++) all.push_back({t, c, k});
    const bool rate = irr && P.target_bytes > 0;
    parallel_for((int)all.size(), P.threads, [&](int i) {
        TileWork& T = tiles[all[i].t];
        CB& cb = T.cbs[all[i].c][all[i].k];
        const TCG& g = T.comps[all[i].c];
        const BandG& B = g.res[cb.band_idx / 3][cb.band_idx % 3];
        int w = g.x1 - g.x0;
        const int32_t* src = T.coef[all[i].c].data() + (size_t)(B.oy + cb.y0 - B.y0) * w + (B.ox + cb.x0 - B.x0);
        t1_encode(src, cb.x1 - cb.x0, cb.y1 - cb.y0, w, B.orient, cb.weight, rate, cb.r, P.cbsty);
        cb.keep = cb.r.npasses;
        cb.keep_len = (uint32_t)cb.r.data.size();
    });
    // rate control: common slope threshold over every block's convex hull
    if (rate) {
        std::vector<CB*> blocks;
        for (auto& T : tiles)
            for (auto& v : T.cbs)
                for (auto& cb : v)
                    if (cb.r.npasses) blocks.push_back(&cb);
        // hull slopes per block
        std::vector<std::vector<std::pair<int, double>>> hulls(blocks.size());
        for (size_t b = 0; b < blocks.size(); b++) {
            const BlockResult& R = blocks[b]->r;
            std::vector<std::pair<int, double>>& hl = hulls[b];  // (pass count, slope)
            double D = 0;
            std::vector<double> cumD(R.passes.size() + 1, 0), cumR(R.passes.size() + 1, 0);
            for (size_t k = 0; k < R.passes.size(); k++) {
                D += R.passes[k].dist_reduction;
                cumD[k + 1] = D;
                cumR[k + 1] = (double)R.passes[k].end_bytes + 1.0;
            }
            std::vector<int> pts{0};
            std::vector<uint8_t> cut(R.passes.size() + 1, R.seglen.empty() ? 1 : 0);
            for (int e : R.seg_passes) cut[e] = 1;
            for (int k = 1; k <= (int)R.passes.size(); k++) {
                if (!cut[k] || cumD[k] <= cumD[pts.back()]) continue;
                while (pts.size() >= 2) {
                    int a = pts[pts.size() - 2], bb = pts.back();
                    double s1 = (cumD[bb] - cumD[a]) / std::max(1e-9, cumR[bb] - cumR[a]);
                    double s2 = (cumD[k] - cumD[bb]) / std::max(1e-9, cumR[k] - cumR[bb]);
                    if (s2 >= s1)
                        pts.pop_back();
                    else
                        break;
                }
                pts.push_back(k);
            }
            for (size_t i = 1; i < pts.size(); i++) {
                double s = (cumD[pts[i]] - cumD[pts[i - 1]]) / std::max(1e-9, cumR[pts[i]] - cumR[pts[i - 1]]);
                hl.push_back({pts[i], s});
            }
        }
        auto size_at = [&](double lambda, std::vector<int>* keep) {
            double total = 0;
            for (size_t b = 0; b < blocks.size(); b++) {
                int k = 0;
                for (auto& pr : hulls[b])
                    if (pr.second >= lambda) k = pr.first;
                if (keep) (*keep)[b] = k;
                if (k) total += (double)blocks[b]->r.passes[k - 1].end_bytes + 3;
            }
            return total;
        };
        double budget = P.target_bytes * 0.97 - 200.0 - 24.0 * blocks.size() / 8.0;
        double lo = 1e-15, hi = 1e15;  // slope thresholds: lo keeps everything, hi nothing
        for (int it = 0; it < 120 && hi / lo > 1.0001; it++) {
            double mid = std::sqrt(lo * hi);
            if (size_at(mid, nullptr) > budget)
                lo = mid;
            else
                hi = mid;
        }
        std::vector<int> keep(blocks.size());
        size_at(hi, &keep);
        parallel_for((int)blocks.size(), P.threads, [&](int b) {
            CB& cb = *blocks[b];
            cb.keep = keep[b];
            cb.keep_len = cb.keep ? truncation_length(cb.r, cb.keep) : 0;
        });
    }
    // codestream: main header
    out.clear();
    put16(out, 0xFF4F);  // SOC
    put16(out, 0xFF51);  // SIZ
    put16(out, 38 + 3 * C);
    put16(out, 0);
    put32(out, (uint32_t)W);
    put32(out, (uint32_t)H);
    put32(out, 0);
    put32(out, 0);
    put32(out, (uint32_t)tw);
    put32(out, (uint32_t)th);
    put32(out, 0);
    put32(out, 0);
    put16(out, (uint32_t)C);
    for (int c = 0; c < C; c++) {
        out.push_back((uint8_t)(P.prec - 1));
        out.push_back((uint8_t)P.dx[c]);
        out.push_back((uint8_t)P.dy[c]);
    }
    put16(out, 0xFF52);  // COD
    put16(out, 12);
    out.push_back(0);                  // Scod: default precincts, no SOP/EPH
    out.push_back(0);                  // LRCP
    put16(out, 1);                     // one layer
    out.push_back(mct ? 1 : 0);
    out.push_back((uint8_t)P.levels);
    out.push_back((uint8_t)(P.xcb - 2));
    out.push_back((uint8_t)(P.ycb - 2));
    out.push_back((uint8_t)P.cbsty);   // code-block style
    out.push_back(irr ? 0 : 1);        // 9/7 or 5/3
    put16(out, 0xFF5C);  // QCD
    if (!irr) {
        put16(out, 3 + (uint32_t)qp.size());
        out.push_back((uint8_t)(G << 5));
        for (auto& q : qp) out.push_back((uint8_t)(q.eps << 3));
    } else {
        put16(out, 3 + 2 * (uint32_t)qp.size());
        out.push_back((uint8_t)((G << 5) | 2));
        for (auto& q : qp) put16(out, (uint32_t)((q.eps << 11) | q.mu));
    }
    // tiles: one tile-part each, packets in LRCP order with one layer; with
    // packed headers the packet headers go to PPT (tile-part header) or PPM
    // (main header) marker segments and the tile data holds the bodies only
    std::vector<uint8_t> parts;
    std::vector<std::vector<uint8_t>> packed_hdrs(tiles.size());
    const size_t seg_max = P.packed_max > 0 ? (size_t)std::min(P.packed_max, 65532) : 65532;
    for (int t = 0; t < (int)tiles.size(); t++) {
        TileWork& T = tiles[t];
        std::vector<uint8_t> body;
        std::vector<uint8_t>& phdr = packed_hdrs[t];
        std::vector<std::vector<TagTree>> incl(C), zbp(C);
        std::vector<std::vector<int>> lblock(C);
        for (int c = 0; c < C; c++) {
            incl[c].resize(3 * (P.levels + 1));
            zbp[c].resize(3 * (P.levels + 1));
            lblock[c].assign(T.cbs[c].size(), 3);
            for (int r = 0; r <= P.levels; r++)
                for (size_t k = 0; k < T.comps[c].res[r].size(); k++) {
                    const BandG& B = T.comps[c].res[r][k];
                    if (!B.ncx) continue;
                    TagTree& it = incl[c][r * 3 + k];
                    TagTree& zt = zbp[c][r * 3 + k];
                    it.init(B.ncx, B.ncy);
                    zt.init(B.ncx, B.ncy);
                    int Mb = G + B.eps - 1;
                    for (int j = 0; j < B.ncy; j++)
                        for (int i = 0; i < B.ncx; i++) {
                            CB& cb = T.cbs[c][B.first + j * B.ncx + i];
                            bool in = cb.keep > 0;
                            it.set(i, j, in ? 0 : 1);
                            zt.set(i, j, in ? Mb - cb.r.numbps : 0);
                            if (in && cb.r.numbps > Mb) return -2;  // guard bits exhausted
                        }
                }
        }
        for (int r = 0; r <= P.levels; r++)
            for (int c = 0; c < C; c++) {
                const TCG& g = T.comps[c];
                int64_t rs = (int64_t)1 << (P.levels - r);
                if (ceil_div(g.x1, rs) <= ceil_div(g.x0, rs) || ceil_div(g.y1, rs) <= ceil_div(g.y0, rs)) continue;
                std::vector<uint8_t> hdr;
                BitWriter bw(hdr);
                bool any = false;
                for (const BandG& B : g.res[r])
                    for (int i = 0; i < B.ncx * B.ncy; i++) any |= T.cbs[c][B.first + i].keep > 0;
                bw.bit(any ? 1 : 0);
                std::vector<const CB*> bodies;
                if (any) {
                    for (size_t k = 0; k < g.res[r].size(); k++) {
                        const BandG& B = g.res[r][k];
                        int Mb = G + B.eps - 1;
                        for (int j = 0; j < B.ncy; j++)
                            for (int i = 0; i < B.ncx; i++) {
                                int idx = B.first + j * B.ncx + i;
                                const CB& cb = T.cbs[c][idx];
                                incl[c][r * 3 + k].encode(bw, i, j, 1);
                                if (!cb.keep) continue;
                                zbp[c][r * 3 + k].encode(bw, i, j, Mb - cb.r.numbps + 1);
                                int np = cb.keep;
                                if (np == 1)
                                    bw.bit(0);
                                else if (np == 2)
                                    bw.bits(2, 2);
                                else if (np <= 5)
                                    bw.bits(0xC | (np - 3), 4);
                                else if (np <= 36)
                                    bw.bits((0xF << 5) | (np - 6), 9);
                                else
                                    bw.bits((0x1FFu << 7) | (uint32_t)(np - 37), 16);
                                // one length per codeword segment of the kept
                                // passes, in Lblock + floor(log2(passes)) bits
                                // each (B.10.7.1, B.10.7.2)
                                std::vector<std::pair<uint32_t, int>> segs;
                                if (cb.r.seglen.empty()) {
                                    segs.push_back({cb.keep_len, np});
                                } else {
                                    int prev = 0;
                                    for (size_t q = 0; q < cb.r.seglen.size() && cb.r.seg_passes[q] <= np; q++) {
                                        segs.push_back({cb.r.seglen[q], cb.r.seg_passes[q] - prev});
                                        prev = cb.r.seg_passes[q];
                                    }
                                }
                                auto flog2 = [](int v) {
                                    int l = 0;
                                    while ((1 << (l + 1)) <= v) l++;
                                    return l;
                                };
                                int& lb = lblock[c][idx];
                                int want = lb;
                                for (auto& sg : segs) {
                                    int need = 1;
                                    while (((uint32_t)1 << need) <= sg.first) need++;
                                    want = std::max(want, need - flog2(sg.second));
                                }
                                while (lb < want) {
                                    bw.bit(1);
                                    lb++;
                                }
                                bw.bit(0);
                                for (auto& sg : segs) bw.bits(sg.first, lb + flog2(sg.second));
                                bodies.push_back(&cb);
                            }
                    }
                }
                bw.align();
                std::vector<uint8_t>& dst = P.packed ? phdr : body;
                dst.insert(dst.end(), hdr.begin(), hdr.end());
                for (const CB* cb : bodies) body.insert(body.end(), cb->r.data.begin(), cb->r.data.begin() + cb->keep_len);
            }
        std::vector<uint8_t> tph;  // tile-part header segments
        if (P.packed == 1)
            for (size_t off = 0, z = 0; off < phdr.size() || z == 0; z++) {
                if (z > 255) return -4;  // Zppt is one byte
                size_t n = std::min(seg_max, phdr.size() - off);
                put16(tph, 0xFF61);  // PPT
                put16(tph, (uint32_t)(3 + n));
                tph.push_back((uint8_t)z);
                tph.insert(tph.end(), phdr.begin() + off, phdr.begin() + off + n);
                off += n;
            }
        put16(parts, 0xFF90);  // SOT
        put16(parts, 10);
        put16(parts, (uint32_t)t);
        put32(parts, (uint32_t)(12 + tph.size() + 2 + body.size()));
        parts.push_back(0);
        parts.push_back(1);
        parts.insert(parts.end(), tph.begin(), tph.end());
        put16(parts, 0xFF93);  // SOD
        parts.insert(parts.end(), body.begin(), body.end());
    }
    if (P.packed == 2) {
        // Nppm (4 bytes) + Ippm per tile-part in codestream order, split over
        // PPM segments wherever the payload limit falls
        std::vector<uint8_t> stream;
        for (auto& h : packed_hdrs) {
            put32(stream, (uint32_t)h.size());
            stream.insert(stream.end(), h.begin(), h.end());
        }
        for (size_t off = 0, z = 0; off < stream.size(); z++) {
            if (z > 255) return -4;  // Zppm is one byte
            size_t n = std::min(seg_max, stream.size() - off);
            put16(out, 0xFF60);  // PPM
            put16(out, (uint32_t)(3 + n));
            out.push_back((uint8_t)z);
            out.insert(out.end(), stream.begin() + off, stream.begin() + off + n);
            off += n;
        }
    }
    out.insert(out.end(), parts.begin(), parts.end());
    put16(out, 0xFFD9);  // EOC
    return 0;
}

}  // namespace tce

extern "C" {

// Returns 0 and a malloc'ed codestream in *out (free with tce_free).
int tce_encode(const void* pixels, int w, int h, int ncomp, int prec, int levels, int cbw_exp, int cbh_exp,
               int tile_w, int tile_h, int mct, int irreversible, double target_bytes, double base_step, int threads,
               int cbsty, int packed, int packed_max, const int* sub, uint8_t** out, size_t* out_len) {
    if (!pixels || !out || !out_len || w <= 0 || h <= 0 || (cbsty & ~63) || packed < 0 || packed > 2)
        return -1;
    tce::Params P{w, h, ncomp, prec, levels, cbw_exp, cbh_exp, tile_w, tile_h, mct, irreversible,
                  threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency()), target_bytes,
                  base_step > 0 ? base_step : 0.5, cbsty, packed, packed_max};
    if (sub)  // sub[2c], sub[2c+1]: XRsiz, YRsiz of component c (NULL: none)
        for (int c = 0; c < ncomp && c < 4; c++) {
            P.dx[c] = sub[2 * c];
            P.dy[c] = sub[2 * c + 1];
        }
    std::vector<uint8_t> cs;
    int rc = tce::encode(pixels, P, cs);
    if (rc) return rc;
    *out = (uint8_t*)std::malloc(cs.size());
    if (!*out) return -3;
    std::memcpy(*out, cs.data(), cs.size());
    *out_len = cs.size();
    return 0;
}

void tce_free(uint8_t* p) { std::free(p); }

// Tier-1 encode of one block of integer coefficients with every pass kept
// (test hook).  Returns the number of bit-planes; writes the codeword.
int tce_t1_encode_block(const int32_t* coef, int w, int h, int orient, uint8_t* out, int cap, int* len) {
    tce::BlockResult R;
    tce::t1_encode(coef, w, h, w, orient, 1.0, false, R);
    if ((int)R.data.size() > cap) return -1;
    std::memcpy(out, R.data.data(), R.data.size());
    *len = (int)R.data.size();
    return R.numbps;
}

int tce_t1_encode_truncated_style(const int32_t* coef, int w, int h, int orient, int keep, int style, uint8_t* out,
                                  int cap, int* numbps);

// Tier-1 encode and truncate after `keep` passes at the shortest decodable
// length (test hook).  Returns the length.
int tce_t1_encode_truncated(const int32_t* coef, int w, int h, int orient, int keep, uint8_t* out, int cap, int* numbps) {
    return tce_t1_encode_truncated_style(coef, w, h, orient, keep, 0, out, cap, numbps);
}

// Same with code-block style bits (RESET 2, CAUSAL 8, PTERM 16, SEGSYM 32).
int tce_t1_encode_truncated_style(const int32_t* coef, int w, int h, int orient, int keep, int style, uint8_t* out,
                                  int cap, int* numbps) {
    tce::BlockResult R;
    tce::t1_encode(coef, w, h, w, orient, 1.0, true, R, style);
    *numbps = R.numbps;
    if (!R.npasses) return 0;
    if (keep > R.npasses) keep = R.npasses;
    uint32_t L = tce::truncation_length(R, keep);
    if ((int)L > cap) return -1;
    std::memcpy(out, R.data.data(), L);
    return (int)L;
}

// Same for every style, BYPASS (1) and TERMALL (4) included: truncates after
// the last codeword segment that ends at or before `keep` passes, and returns
// the segment lengths (seglen[0..*nseg)) and the pass count kept.
int tce_t1_encode_segments(const int32_t* coef, int w, int h, int orient, int keep, int style, uint8_t* out, int cap,
                           int* numbps, uint32_t* seglen, int seg_cap, int* nseg, int* npasses) {
    tce::BlockResult R;
    tce::t1_encode(coef, w, h, w, orient, 1.0, true, R, style);
    *numbps = R.numbps;
    *nseg = 0;
    *npasses = 0;
    if (!R.npasses) return 0;
    if (keep > R.npasses) keep = R.npasses;
    if (R.seglen.empty()) {
        uint32_t L = tce::truncation_length(R, keep);
        if ((int)L > cap || seg_cap < 1) return -1;
        std::memcpy(out, R.data.data(), L);
        seglen[0] = L;
        *nseg = 1;
        *npasses = keep;
        return (int)L;
    }
    uint32_t L = 0;
    int n = 0, kept = 0;
    for (size_t s = 0; s < R.seglen.size() && R.seg_passes[s] <= keep; s++) {
        if (n >= seg_cap) return -1;
        seglen[n++] = R.seglen[s];
        L += R.seglen[s];
        kept = R.seg_passes[s];
    }
    if ((int)L > cap) return -1;
    std::memcpy(out, R.data.data(), L);
    *nseg = n;
    *npasses = kept;
    return (int)L;
}

}  // extern "C"
