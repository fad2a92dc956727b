/*
 * t800_oracle.c -- CPU restatement of the JPEG 2000 Part-1 decode path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker used by tests/, the smoke test
 * and the benchmark's CPU-baseline leg.  The product never links it.
 *
 * It is written for clarity, not speed: flags live in plain 2-D arrays with a
 * one-sample border, every pass walks every coefficient, the 9/7 lifting runs
 * in double precision.  Section references are to ITU-T T.800 (2002); the
 * reference file:line each part follows is given at each function.
 */
#include "t800_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXC 16
#define MAXLVL 32

/* ------------------------------------------------------------------------ */
/* MQ coder tables, T.800 Table C.2 (equal to TC/coefficient_coder.cl:36-60). */
static const uint16_t QE[47] = {
    0x5601, 0x3401, 0x1801, 0x0AC1, 0x0521, 0x0221, 0x5601, 0x5401, 0x4801, 0x3801,
    0x3001, 0x2401, 0x1C01, 0x1601, 0x5601, 0x5401, 0x5101, 0x4801, 0x3801, 0x3401,
    0x3001, 0x2801, 0x2401, 0x2201, 0x1C01, 0x1801, 0x1601, 0x1401, 0x1201, 0x1101,
    0x0AC1, 0x09C1, 0x08A1, 0x0521, 0x0441, 0x02A1, 0x0221, 0x0141, 0x0111, 0x0085,
    0x0049, 0x0025, 0x0015, 0x0009, 0x0005, 0x0001, 0x5601};
static const uint8_t NMPS[47] = {
    1,  2,  3,  4,  5,  38, 7,  8,  9,  10, 11, 12, 13, 29, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24,
    25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 45, 46};
static const uint8_t NLPS[47] = {
    1,  6,  9,  12, 29, 33, 6,  14, 14, 14, 17, 18, 20, 21, 14, 14, 15, 16, 17, 18, 19, 19, 20, 21,
    22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 46};
static const uint8_t SWITCH_[47] = {1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

/* Context labels used here: 0..8 zero coding, 9..13 sign, 14..16 refinement,
 * 17 run-length, 18 uniform (T.800 D.7; the reference numbers RUN=18, UNI=17,
 * TC/coefficient_coder.cl:65-66 -- only the labels differ). */
#define CX_RL 17
#define CX_UNI 18
#define NCTX 19

typedef struct {
    const uint8_t *d;
    int len, bp;
    uint32_t A, C;
    int CT;
    uint8_t I[NCTX], MPS[NCTX];
    long nsym;
    int raw; /* 1: the current segment is raw (BYPASS style, T.800 D.6) */
} mqd_t;

static int mq_byte(const mqd_t *m, int i) { return i < m->len ? m->d[i] : 0xFF; }

/* BYTEIN, T.800 Figure C.19 (the reference's look-ahead form is
 * TC/coefficient_coder.cl:121-136).  Past the end of the segment the decoder
 * sees 0xFF 0xFF, i.e. a marker, and keeps feeding 1-bits. */
static void mq_bytein(mqd_t *m) {
    if (mq_byte(m, m->bp) == 0xFF) {
        if (mq_byte(m, m->bp + 1) > 0x8F) {
            m->C += 0xFF00;
            m->CT = 8;
        } else {
            m->bp++;
            m->C += (uint32_t)mq_byte(m, m->bp) << 9;
            m->CT = 7;
        }
    } else {
        m->bp++;
        m->C += (uint32_t)mq_byte(m, m->bp) << 8;
        m->CT = 8;
    }
}

/* Context reset per T.800 Table D.7 (TC/coefficient_coder.cl:105-119). */
static void mq_reset_ctx(mqd_t *m) {
    memset(m->I, 0, sizeof m->I);
    memset(m->MPS, 0, sizeof m->MPS);
    m->I[CX_UNI] = 46;
    m->I[CX_RL] = 3;
    m->I[0] = 4;
}

/* INITDEC, T.800 Figure C.20 (TC/coefficient_coder.cl:202-221), on a new
 * codeword segment; the context states are left as they are. */
static void mq_start(mqd_t *m, const uint8_t *d, int len) {
    m->d = d;
    m->len = len;
    m->bp = 0;
    m->raw = 0;
    m->C = (uint32_t)mq_byte(m, 0) << 16;
    mq_bytein(m);
    m->C <<= 7;
    m->CT -= 7;
    m->A = 0x8000;
}

static void mq_init(mqd_t *m, const uint8_t *d, int len) {
    mq_start(m, d, len);
    m->nsym = 0;
    mq_reset_ctx(m);
}

/* A raw segment (T.800 D.6): bits MSB first; a byte after 0xFF holds 7 bits
 * below a stuffed 0; past the end the decoder sees 0xFF. */
static void raw_start(mqd_t *m, const uint8_t *d, int len) {
    m->d = d;
    m->len = len;
    m->bp = -1;
    m->raw = 1;
    m->C = 0;
    m->CT = 0;
}

static int raw_bit(mqd_t *m) {
    if (m->CT == 0) {
        int prev = m->bp >= 0 ? mq_byte(m, m->bp) : 0;
        m->bp++;
        m->C = (uint32_t)mq_byte(m, m->bp);
        m->CT = prev == 0xFF ? 7 : 8;
    }
    m->CT--;
    m->nsym++;
    return (int)((m->C >> m->CT) & 1);
}

/* RENORMD, T.800 Figure C.18 (TC/coefficient_coder.cl:138-150). */
static void mq_renorm(mqd_t *m) {
    do {
        if (m->CT == 0) mq_bytein(m);
        m->A <<= 1;
        m->C <<= 1;
        m->CT--;
    } while ((m->A & 0x8000) == 0);
}

/* DECODE with LPS/MPS exchange, T.800 Figures C.15-C.17
 * (TC/coefficient_coder.cl:152-252, with the MPS switch actually applied). */
static int mq_decode(mqd_t *m, int cx) {
    if (m->raw) return raw_bit(m);
    int I = m->I[cx];
    uint32_t qe = QE[I];
    int d;
    m->nsym++;
    m->A -= qe;
    if ((m->C >> 16) < qe) {
        if (m->A < qe) {
            d = m->MPS[cx];
            m->I[cx] = NMPS[I];
        } else {
            d = 1 - m->MPS[cx];
            if (SWITCH_[I]) m->MPS[cx] ^= 1;
            m->I[cx] = NLPS[I];
        }
        m->A = qe;
        mq_renorm(m);
    } else {
        m->C -= qe << 16;
        if ((m->A & 0x8000) == 0) {
            if (m->A < qe) {
                d = 1 - m->MPS[cx];
                if (SWITCH_[I]) m->MPS[cx] ^= 1;
                m->I[cx] = NLPS[I];
            } else {
                d = m->MPS[cx];
                m->I[cx] = NMPS[I];
            }
            mq_renorm(m);
        } else {
            d = m->MPS[cx];
        }
    }
    return d;
}

int or_mq_decode_symbols(const uint8_t *data, int len, const uint8_t *ctx, int n, uint8_t *out) {
    mqd_t m;
    if (!data && len) return -OR_ERR_ARG;
    mq_init(&m, data, len);
    for (int i = 0; i < n; i++) {
        if (ctx[i] >= NCTX) return -OR_ERR_ARG;
        out[i] = (uint8_t)mq_decode(&m, ctx[i]);
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* Context formation, T.800 D.3.1-D.3.3 (the reference's LUT forms are
 * TC/coefficient_coder.cl:330-469). */

/* Zero-coding label from horizontal (h), vertical (v), diagonal (d) counts of
 * significant neighbours.  cls: 0 = LL/LH, 1 = HL, 2 = HH (Table D.1). */
static int zc_label(int cls, int h, int v, int d) {
    if (cls == 1) {
        int t = h;
        h = v;
        v = t;
    }
    if (cls < 2) {
        if (h == 2) return 8;
        if (h == 1) return v >= 1 ? 7 : (d >= 1 ? 6 : 5);
        if (v == 2) return 4;
        if (v == 1) return 3;
        return d >= 2 ? 2 : d;
    }
    {
        int hv = h + v;
        if (d >= 3) return 8;
        if (d == 2) return hv >= 1 ? 7 : 6;
        if (d == 1) return hv >= 2 ? 5 : (hv == 1 ? 4 : 3);
        return hv >= 2 ? 2 : hv;
    }
}

/* Sign-coding label and XOR bit, Tables D.2/D.3. */
int or_sign_ctx(int h, int v, int *xorbit) {
    int x = 0, lab;
    if (h < 0) {
        h = -h;
        v = -v;
        x = 1;
    }
    if (h == 0 && v < 0) {
        v = -v;
        x = 1;
    }
    /* now h in {0,1}; (h=0,v=-1) folded to (0,1) with x=1 */
    if (h == 1)
        lab = (v == 1) ? 13 : (v == 0 ? 12 : 11);
    else
        lab = (v == 0) ? 9 : 10;
    if (xorbit) *xorbit = x;
    return lab;
}

void or_tables(uint8_t *zc, uint32_t *mq) {
    if (zc) {
        for (int cls = 0; cls < 3; cls++)
            for (int n = 0; n < 256; n++) {
                /* bit order tl,t,tr,l,r,bl,b,br */
                int tl = n & 1, t = (n >> 1) & 1, tr = (n >> 2) & 1, l = (n >> 3) & 1;
                int r = (n >> 4) & 1, bl = (n >> 5) & 1, b = (n >> 6) & 1, br = (n >> 7) & 1;
                zc[cls * 256 + n] = (uint8_t)zc_label(cls, l + r, t + b, tl + tr + bl + br);
            }
    }
    if (mq) {
        for (int i = 0; i < 47; i++) {
            mq[i] = QE[i];
            mq[47 + i] = NMPS[i];
            mq[94 + i] = NLPS[i];
            mq[141 + i] = SWITCH_[i];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Tier-1 decoding of one code-block (T.800 D.3-D.5).  Follows the pass order
 * of g_decode, TC/coefficient_coder.cl:778-819: a cleanup pass on the top
 * plane, then significance propagation, magnitude refinement and cleanup on
 * every lower plane, stripes of 4 rows, column-major inside a stripe
 * (TC/coefficient_coder.cl:681-776).  npasses bounds the passes decoded
 * (truncated code-blocks), which the reference ignored.  Code-block style
 * (COD/COC SPcod, Table A.19) adds, where set: RESET, every context back to
 * its initial state after each pass (D.6); CAUSAL, the row below a stripe
 * taken as insignificant when forming any context of the stripe's last row
 * (D.7); SEGSYM, four uniform-context symbols after each cleanup pass (D.5).
 * The reference decodes style 0 only. */

typedef struct {
    int w, h, W; /* W = w + 2 (bordered pitch) */
    uint8_t *sig, *vis, *ref, *sgn;
    int32_t *mag;
    int8_t *low;
    int causal;
} t1_t;

#define AT(t, x, y) (((y) + 1) * (t)->W + (x) + 1)

/* 1 when the row below (x, y) counts as a neighbour */
static int below(const t1_t *t, int y) { return !(t->causal && (y & 3) == 3); }

static int nb_counts(const t1_t *t, int x, int y, int *h, int *v, int *d) {
    const uint8_t *s = t->sig;
    int i = AT(t, x, y), W = t->W, b = below(t, y);
    *h = s[i - 1] + s[i + 1];
    *v = s[i - W] + (b ? s[i + W] : 0);
    *d = s[i - W - 1] + s[i - W + 1] + (b ? s[i + W - 1] + s[i + W + 1] : 0);
    return *h + *v + *d;
}

static int contrib(const t1_t *t, int i) { return t->sig[i] ? (t->sgn[i] ? -1 : 1) : 0; }

static void decode_sign(t1_t *t, mqd_t *m, int x, int y) {
    int i = AT(t, x, y), W = t->W, xr;
    int h = contrib(t, i - 1) + contrib(t, i + 1);
    int v = contrib(t, i - W) + (below(t, y) ? contrib(t, i + W) : 0);
    h = h > 0 ? 1 : (h < 0 ? -1 : 0);
    v = v > 0 ? 1 : (v < 0 ? -1 : 0);
    int lab = or_sign_ctx(h, v, &xr);
    /* in a raw (BYPASS) segment the sign is the bit itself, without the
     * sign-context XOR (T.800 D.6) */
    int bit = mq_decode(m, lab);
    t->sgn[i] = (uint8_t)(m->raw ? bit : bit ^ xr);
}

static void become_sig(t1_t *t, mqd_t *m, int x, int y, int p) {
    int i = AT(t, x, y);
    decode_sign(t, m, x, y);
    t->sig[i] = 1;
    t->mag[y * t->w + x] |= (int32_t)1 << p;
    t->low[y * t->w + x] = (int8_t)p;
}

static void pass_sigprop(t1_t *t, mqd_t *m, int cls, int p) {
    for (int y0 = 0; y0 < t->h; y0 += 4)
        for (int x = 0; x < t->w; x++)
            for (int y = y0; y < y0 + 4 && y < t->h; y++) {
                int i = AT(t, x, y), h, v, d;
                if (t->sig[i]) continue;
                if (nb_counts(t, x, y, &h, &v, &d) == 0) continue;
                if (mq_decode(m, zc_label(cls, h, v, d))) become_sig(t, m, x, y, p);
                t->vis[i] = 1;
            }
}

static void pass_magref(t1_t *t, mqd_t *m, int p) {
    for (int y0 = 0; y0 < t->h; y0 += 4)
        for (int x = 0; x < t->w; x++)
            for (int y = y0; y < y0 + 4 && y < t->h; y++) {
                int i = AT(t, x, y), h, v, d, lab;
                if (!t->sig[i] || t->vis[i]) continue;
                if (t->ref[i])
                    lab = 16;
                else
                    lab = nb_counts(t, x, y, &h, &v, &d) ? 15 : 14;
                if (mq_decode(m, lab)) t->mag[y * t->w + x] |= (int32_t)1 << p;
                t->low[y * t->w + x] = (int8_t)p;
                t->ref[i] = 1;
            }
}

static void pass_cleanup(t1_t *t, mqd_t *m, int cls, int p) {
    for (int y0 = 0; y0 < t->h; y0 += 4)
        for (int x = 0; x < t->w; x++) {
            int y = y0;
            if (y0 + 4 <= t->h) {
                int rl = 1;
                for (int k = 0; k < 4 && rl; k++) {
                    int i = AT(t, x, y0 + k), h, v, d;
                    if (t->sig[i] || t->vis[i] || nb_counts(t, x, y0 + k, &h, &v, &d)) rl = 0;
                }
                if (rl) {
                    if (!mq_decode(m, CX_RL)) continue;
                    int r = mq_decode(m, CX_UNI) << 1;
                    r |= mq_decode(m, CX_UNI);
                    y = y0 + r;
                    become_sig(t, m, x, y, p);
                    y++;
                }
            }
            for (; y < y0 + 4 && y < t->h; y++) {
                int i = AT(t, x, y), h, v, d;
                if (t->sig[i] || t->vis[i]) continue;
                nb_counts(t, x, y, &h, &v, &d);
                if (mq_decode(m, zc_label(cls, h, v, d))) become_sig(t, m, x, y, p);
            }
        }
    {
        int n = (t->w + 2) * (t->h + 2);
        memset(t->vis, 0, (size_t)n);
    }
}

int or_t1_decode(const uint8_t *data, int len, int w, int h, int orient, int numbps, int npasses,
                 int32_t *mag, int8_t *sign, int8_t *lowplane) {
    return or_t1_decode_style(data, len, w, h, orient, numbps, npasses, 0, mag, sign, lowplane);
}

int or_t1_decode_style(const uint8_t *data, int len, int w, int h, int orient, int numbps, int npasses, int style,
                       int32_t *mag, int8_t *sign, int8_t *lowplane) {
    return or_t1_decode_seg(data, len, &len, 1, w, h, orient, numbps, npasses, style, mag, sign, lowplane);
}

/* BYPASS / TERMALL codeword segments (T.800 D.6, Table D.9): pass k (0 = the
 * first cleanup) ends a segment; the SPP and MRP passes from the fifth
 * bit-plane on are raw under BYPASS. */
int or_seg_end(int style, int k) {
    if (style & OR_STYLE_TERMALL) return 1;
    if (style & OR_STYLE_BYPASS) return k == 9 || (k > 9 && (k - 1) % 3 != 0);
    return 0;
}

static int raw_pass(int style, int k) { return (style & OR_STYLE_BYPASS) && k >= 10 && (k - 1) % 3 != 2; }

int or_t1_decode_seg(const uint8_t *data, int len, const int *seglen, int nseg, int w, int h, int orient,
                     int numbps, int npasses, int style, int32_t *mag, int8_t *sign, int8_t *lowplane) {
    if (w <= 0 || h <= 0 || w > 1024 || h > 1024 || orient < 0 || orient > 3 || nseg < 1) return -OR_ERR_ARG;
    if (style & ~OR_STYLE_DECODABLE) return -OR_ERR_UNSUPPORTED;
    {
        long tot = 0;
        for (int s = 0; s < nseg; s++) tot += seglen[s] < 0 ? 0 : seglen[s];
        if (tot > len) return -OR_ERR_ARG;
    }
    int cls = orient == 1 ? 1 : (orient == 3 ? 2 : 0);
    t1_t t;
    int n = (w + 2) * (h + 2);
    t.w = w;
    t.h = h;
    t.W = w + 2;
    uint8_t *flags = (uint8_t *)calloc((size_t)n * 4, 1);
    if (!flags) return -OR_ERR_NOMEM;
    t.sig = flags;
    t.vis = flags + n;
    t.ref = flags + 2 * n;
    t.sgn = flags + 3 * n;
    t.mag = mag;
    t.low = lowplane;
    t.causal = (style & OR_STYLE_CAUSAL) != 0;
    for (int i = 0; i < w * h; i++) {
        mag[i] = 0;
        lowplane[i] = -1;
    }
    mqd_t m;
    int seg = 0, off = 0;
    mq_init(&m, data, seglen[0]);
    for (int k = 0; k < npasses; k++) {
        int p = numbps - 1 - (k + 2) / 3;
        if (p < 0) break;
        int type = k == 0 ? 2 : (k - 1) % 3; /* 0 sigprop, 1 magref, 2 cleanup */
        if (k > 0 && or_seg_end(style, k - 1)) {
            /* the next codeword segment: raw, or the MQ decoder re-initialised */
            off += seg < nseg ? seglen[seg] : 0;
            seg++;
            int sl = seg < nseg ? seglen[seg] : 0;
            if (raw_pass(style, k))
                raw_start(&m, data + off, sl);
            else
                mq_start(&m, data + off, sl);
        }
        if (type == 0)
            pass_sigprop(&t, &m, cls, p);
        else if (type == 1)
            pass_magref(&t, &m, p);
        else {
            pass_cleanup(&t, &m, cls, p);
            if (style & OR_STYLE_SEGSYM) {
                /* the segmentation symbol 1010 (D.5): anything else means the
                 * codeword (or the style signalled for it) is wrong */
                int seg = 0;
                for (int s = 0; s < 4; s++) seg = (seg << 1) | mq_decode(&m, CX_UNI);
                if (seg != 0xA) {
                    free(flags);
                    return -OR_ERR_CORRUPT;
                }
            }
        }
        if (style & OR_STYLE_RESET) mq_reset_ctx(&m);
    }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) sign[y * w + x] = (int8_t)t.sgn[AT(&t, x, y)];
    free(flags);
    return (int)m.nsym;
}

/* ------------------------------------------------------------------------ */
/* 1-D and 2-D wavelet transforms (T.800 F.3 / F.4), lifting form.  The
 * reference's primitives are TC/dwt_transform_buffer.cl:102-556 (truncating
 * '/' there is replaced by floor, T.800 F.3.8.1) with whole-sample symmetric
 * extension as in TC/dwt_io.cl:7-38 (mirror), here the general PSE of F.3.7
 * with origin parity. */

static int pse(int i, int i0, int i1) {
    int per = 2 * (i1 - i0 - 1);
    int m = (i - i0) % per;
    if (m < 0) m += per;
    if (m >= i1 - i0) m = per - m;
    return i0 + m;
}

static long floordiv(long a, long b) { /* b > 0 */
    long q = a / b;
    if ((a % b) != 0 && ((a < 0) != (b < 0))) q--;
    return q;
}

#define EXT 8

/* 1D_SR for 5/3 on x[i0..i1) addressed by absolute index (x points to index 0
 * of an array spanning i0-EXT..i1+EXT through an offset). */
static void sr53(int64_t *X, int i0, int i1) {
    if (i1 - i0 == 1) {
        if (i0 & 1) X[i0] /= 2;
        return;
    }
    for (int i = i0 - EXT; i < i0; i++) X[i] = X[pse(i, i0, i1)];
    for (int i = i1; i < i1 + EXT; i++) X[i] = X[pse(i, i0, i1)];
    /* STEP1: X(2n) = Y(2n) - floor((Y(2n-1)+Y(2n+1)+2)/4), i0/2 <= n < i1/2 + 1 */
    for (int i = i0 - 2; i < i1 + 2; i++)
        if (!(i & 1)) X[i] = X[i] - floordiv(X[i - 1] + X[i + 1] + 2, 4);
    /* STEP2: X(2n+1) = Y(2n+1) + floor((X(2n)+X(2n+2))/2), i0/2 <= n < i1/2 */
    for (int i = i0; i < i1; i++)
        if (i & 1) X[i] = X[i] + floordiv(X[i - 1] + X[i + 1], 2);
}

static void sd53(int64_t *X, int i0, int i1) {
    if (i1 - i0 == 1) {
        if (i0 & 1) X[i0] *= 2;
        return;
    }
    for (int i = i0 - EXT; i < i0; i++) X[i] = X[pse(i, i0, i1)];
    for (int i = i1; i < i1 + EXT; i++) X[i] = X[pse(i, i0, i1)];
    for (int i = i0 - 3; i < i1 + 3; i++)
        if (i & 1) X[i] = X[i] - floordiv(X[i - 1] + X[i + 1], 2);
    for (int i = i0; i < i1; i++)
        if (!(i & 1)) X[i] = X[i] + floordiv(X[i - 1] + X[i + 1] + 2, 4);
}

/* T.800 Table F.4 lifting parameters. */
static const double AL = -1.586134342059924;
static const double BE = -0.052980118572961;
static const double GA = 0.882911075530934;
static const double DE = 0.443506852043971;
static const double KK = 1.230174104914001;

static void sr97(double *X, int i0, int i1) {
    if (i1 - i0 == 1) {
        if (i0 & 1) X[i0] /= 2.0;
        return;
    }
    for (int i = i0 - EXT; i < i0; i++) X[i] = X[pse(i, i0, i1)];
    for (int i = i1; i < i1 + EXT; i++) X[i] = X[pse(i, i0, i1)];
    int a = i0 - 5, b = i1 + 5;
    for (int i = a; i < b; i++) X[i] = (i & 1) ? X[i] / KK : X[i] * KK; /* STEP1, STEP2 */
    for (int i = a + 1; i < b - 1; i++)
        if (!(i & 1)) X[i] -= DE * (X[i - 1] + X[i + 1]); /* STEP3 */
    for (int i = a + 2; i < b - 2; i++)
        if (i & 1) X[i] -= GA * (X[i - 1] + X[i + 1]); /* STEP4 */
    for (int i = a + 3; i < b - 3; i++)
        if (!(i & 1)) X[i] -= BE * (X[i - 1] + X[i + 1]); /* STEP5 */
    for (int i = i0; i < i1; i++)
        if (i & 1) X[i] -= AL * (X[i - 1] + X[i + 1]); /* STEP6 */
}

static void sd97(double *X, int i0, int i1) {
    if (i1 - i0 == 1) {
        if (i0 & 1) X[i0] *= 2.0;
        return;
    }
    for (int i = i0 - EXT; i < i0; i++) X[i] = X[pse(i, i0, i1)];
    for (int i = i1; i < i1 + EXT; i++) X[i] = X[pse(i, i0, i1)];
    int a = i0 - 5, b = i1 + 5;
    for (int i = a + 1; i < b - 1; i++)
        if (i & 1) X[i] += AL * (X[i - 1] + X[i + 1]);
    for (int i = a + 2; i < b - 2; i++)
        if (!(i & 1)) X[i] += BE * (X[i - 1] + X[i + 1]);
    for (int i = a + 3; i < b - 3; i++)
        if (i & 1) X[i] += GA * (X[i - 1] + X[i + 1]);
    for (int i = a + 4; i < b - 4; i++)
        if (!(i & 1)) X[i] += DE * (X[i - 1] + X[i + 1]);
    for (int i = i0; i < i1; i++) X[i] = (i & 1) ? X[i] * KK : X[i] / KK;
}

static long ceildiv(long a, long b) { return -floordiv(-a, b); }

/* Resolution rectangle at reduction 2^n of a tile-component (B.5). */
static void res_rect(int x0, int y0, int x1, int y1, int n, int *rx0, int *ry0, int *rx1, int *ry1) {
    long s = 1L << n;
    *rx0 = (int)ceildiv(x0, s);
    *ry0 = (int)ceildiv(y0, s);
    *rx1 = (int)ceildiv(x1, s);
    *ry1 = (int)ceildiv(y1, s);
}

/* 2D_SR on a Mallat-layout array: for level n from 'levels' down to 1 the
 * region [0,w_r) x [0,h_r) holds LL|HL over LH|HH; interleave, HOR_SR on
 * rows, then VER_SR on columns (F.3.2), result in place. */
#define DEF_2D(NAME, T, FN, IS_INV)                                                              \
    static int NAME(T *a, int w, int h, int x0, int y0, int levels) {                           \
        int maxd = (w > h ? w : h) + 2 * EXT + 4;                                               \
        T *line = (T *)malloc(sizeof(T) * (size_t)maxd);                                        \
        T *tmp = (T *)malloc(sizeof(T) * (size_t)w * (size_t)h + 1);                            \
        if (!line || !tmp) {                                                                    \
            free(line);                                                                         \
            free(tmp);                                                                          \
            return -OR_ERR_NOMEM;                                                               \
        }                                                                                       \
        for (int step = 0; step < levels; step++) {                                             \
            int n = IS_INV ? levels - step : step + 1; /* decomposition level handled */       \
            int u0, v0, u1, v1, l0x, l0y, l1x, l1y;                                             \
            res_rect(x0, y0, x0 + w, y0 + h, n - 1, &u0, &v0, &u1, &v1);                        \
            res_rect(x0, y0, x0 + w, y0 + h, n, &l0x, &l0y, &l1x, &l1y);                        \
            int rw = u1 - u0, rh = v1 - v0, lw = l1x - l0x, lh = l1y - l0y;                     \
            if (rw <= 0 || rh <= 0) continue;                                                   \
            if (IS_INV) {                                                                       \
                /* interleave quadrants into tmp (rw x rh) */                                   \
                for (int y = 0; y < rh; y++)                                                    \
                    for (int x = 0; x < rw; x++) {                                              \
                        int u = u0 + x, v = v0 + y;                                             \
                        int sx = (u & 1) ? lw + (int)(floordiv(u, 2) - floordiv(u0, 2))         \
                                         : (int)(u / 2 - ceildiv(u0, 2));                       \
                        int sy = (v & 1) ? lh + (int)(floordiv(v, 2) - floordiv(v0, 2))         \
                                         : (int)(v / 2 - ceildiv(v0, 2));                       \
                        tmp[y * rw + x] = a[sy * w + sx];                                       \
                    }                                                                           \
                for (int y = 0; y < rh; y++) { /* HOR_SR */                                     \
                    T *X = line + EXT + 2 - u0;                                                 \
                    for (int x = 0; x < rw; x++) X[u0 + x] = tmp[y * rw + x];                   \
                    FN(X, u0, u1);                                                              \
                    for (int x = 0; x < rw; x++) tmp[y * rw + x] = X[u0 + x];                   \
                }                                                                               \
                for (int x = 0; x < rw; x++) { /* VER_SR */                                     \
                    T *X = line + EXT + 2 - v0;                                                 \
                    for (int y = 0; y < rh; y++) X[v0 + y] = tmp[y * rw + x];                   \
                    FN(X, v0, v1);                                                              \
                    for (int y = 0; y < rh; y++) a[y * w + x] = X[v0 + y];                      \
                }                                                                               \
            } else {                                                                            \
                for (int y = 0; y < rh; y++)                                                    \
                    for (int x = 0; x < rw; x++) tmp[y * rw + x] = a[y * w + x];                \
                for (int x = 0; x < rw; x++) { /* VER_SD */                                     \
                    T *X = line + EXT + 2 - v0;                                                 \
                    for (int y = 0; y < rh; y++) X[v0 + y] = tmp[y * rw + x];                   \
                    FN(X, v0, v1);                                                              \
                    for (int y = 0; y < rh; y++) tmp[y * rw + x] = X[v0 + y];                   \
                }                                                                               \
                for (int y = 0; y < rh; y++) { /* HOR_SD */                                     \
                    T *X = line + EXT + 2 - u0;                                                 \
                    for (int x = 0; x < rw; x++) X[u0 + x] = tmp[y * rw + x];                   \
                    FN(X, u0, u1);                                                              \
                    for (int x = 0; x < rw; x++) tmp[y * rw + x] = X[u0 + x];                   \
                }                                                                               \
                for (int y = 0; y < rh; y++) /* deinterleave into quadrants */                  \
                    for (int x = 0; x < rw; x++) {                                              \
                        int u = u0 + x, v = v0 + y;                                             \
                        int sx = (u & 1) ? lw + (int)(floordiv(u, 2) - floordiv(u0, 2))         \
                                         : (int)(u / 2 - ceildiv(u0, 2));                       \
                        int sy = (v & 1) ? lh + (int)(floordiv(v, 2) - floordiv(v0, 2))         \
                                         : (int)(v / 2 - ceildiv(v0, 2));                       \
                        a[sy * w + sx] = tmp[y * rw + x];                                       \
                    }                                                                           \
            }                                                                                   \
        }                                                                                       \
        free(line);                                                                             \
        free(tmp);                                                                              \
        return OR_OK;                                                                           \
    }

DEF_2D(idwt53_i64, int64_t, sr53, 1)
DEF_2D(fdwt53_i64, int64_t, sd53, 0)
DEF_2D(idwt97_d, double, sr97, 1)
DEF_2D(fdwt97_d, double, sd97, 0)

static int run53(int32_t *a, int w, int h, int x0, int y0, int levels, int inv) {
    if (w <= 0 || h <= 0 || levels < 0 || levels > MAXLVL) return -OR_ERR_ARG;
    int64_t *t = (int64_t *)malloc(sizeof(int64_t) * (size_t)w * (size_t)h);
    if (!t) return -OR_ERR_NOMEM;
    for (long i = 0; i < (long)w * h; i++) t[i] = a[i];
    int rc = inv ? idwt53_i64(t, w, h, x0, y0, levels) : fdwt53_i64(t, w, h, x0, y0, levels);
    for (long i = 0; i < (long)w * h; i++) a[i] = (int32_t)t[i];
    free(t);
    return rc;
}

int or_idwt53(int32_t *a, int w, int h, int x0, int y0, int levels) { return run53(a, w, h, x0, y0, levels, 1); }
int or_fdwt53(int32_t *a, int w, int h, int x0, int y0, int levels) { return run53(a, w, h, x0, y0, levels, 0); }
int or_idwt97(double *a, int w, int h, int x0, int y0, int levels) {
    if (w <= 0 || h <= 0 || levels < 0 || levels > MAXLVL) return -OR_ERR_ARG;
    return idwt97_d(a, w, h, x0, y0, levels);
}
int or_fdwt97(double *a, int w, int h, int x0, int y0, int levels) {
    if (w <= 0 || h <= 0 || levels < 0 || levels > MAXLVL) return -OR_ERR_ARG;
    return fdwt97_d(a, w, h, x0, y0, levels);
}

/* ------------------------------------------------------------------------ */
/* Codestream model and tier-2 (T.800 Annex A, B).  Follows the subset of
 * TC/codestream.c:17-258 (main header), :260-298 (tile header), :336-470
 * (packets) and TC/codestream_image.c:14-277 (geometry), extended with
 * expounded quantization, COC/QCC, tile-part headers, COM/TLM/PLT skipping,
 * several layers and tile-parts, precincts (B.6), the five progression orders
 * (B.12.1) and empty packets. */

typedef struct {
    int nlev, xcb, ycb, cbsty, xform;
    int ppx[MAXLVL + 1], ppy[MAXLVL + 1];
} ccs_t; /* coding style of a component */

typedef struct {
    int style, guard, n;
    int eps[3 * MAXLVL + 1], mu[3 * MAXLVL + 1];
} qcs_t; /* quantization of a component */

typedef struct {
    int prog, nlayers, mct, sop, eph;
} gcs_t; /* COD global part */

/* One progression order change (POC, A.6.6): resolutions rs..re-1,
 * components cs..ce-1, layers below le, in order prog. */
typedef struct {
    int rs, cs, le, re, ce, prog;
} poc_t;
#define MAXPOC 64

typedef struct {
    int X0, Y0, X1, Y1, XT0, YT0, XT, YT, C;
    int prec[MAXC], sgnd[MAXC];
    int dx[MAXC], dy[MAXC]; /* component sub-sampling XRsiz, YRsiz */
    gcs_t g;
    ccs_t cc[MAXC];
    qcs_t qc[MAXC];
    int ntx, nty;
    int npoc, poc_tile; /* POC entries; poc_tile: taken from the tile's own headers */
    poc_t poc[MAXPOC];
    int nppm; /* PPM segments of the main header (A.7.4): index, offset, length */
    int ppm_z[256];
    size_t ppm_off[256], ppm_len[256];
} hdr_t;

typedef struct {
    const uint8_t *p;
    size_t len, pos;
    int err;
} rd_t;

static int rd8(rd_t *r) {
    if (r->pos + 1 > r->len) {
        r->err = OR_ERR_TRUNCATED;
        return 0;
    }
    return r->p[r->pos++];
}
static int rd16(rd_t *r) {
    int a = rd8(r);
    return (a << 8) | rd8(r);
}
static uint32_t rd32(rd_t *r) {
    uint32_t a = (uint32_t)rd16(r);
    return (a << 16) | (uint32_t)rd16(r);
}

static int parse_cs_params(rd_t *r, ccs_t *c, int user_prec) {
    c->nlev = rd8(r);
    c->xcb = rd8(r) + 2;
    c->ycb = rd8(r) + 2;
    c->cbsty = rd8(r);
    c->xform = rd8(r);
    if (c->nlev > MAXLVL || c->xcb > 10 || c->ycb > 10 || c->xcb + c->ycb > 12) return OR_ERR_UNSUPPORTED;
    for (int i = 0; i <= c->nlev; i++) {
        if (user_prec) {
            int b = rd8(r);
            c->ppx[i] = b & 15;
            c->ppy[i] = b >> 4;
        } else {
            c->ppx[i] = 15;
            c->ppy[i] = 15;
        }
    }
    return r->err;
}

static int parse_qc_params(rd_t *r, qcs_t *q, size_t end) {
    int s = rd8(r);
    q->style = s & 31;
    q->guard = s >> 5;
    q->n = 0;
    if (q->style == 0) {
        while (r->pos < end && q->n < 3 * MAXLVL + 1) {
            q->eps[q->n] = rd8(r) >> 3;
            q->mu[q->n] = 0;
            q->n++;
        }
    } else if (q->style == 1 || q->style == 2) {
        while (r->pos + 2 <= end && q->n < 3 * MAXLVL + 1) {
            int v = rd16(r);
            q->eps[q->n] = v >> 11;
            q->mu[q->n] = v & 0x7FF;
            q->n++;
        }
    } else
        return OR_ERR_UNSUPPORTED;
    return r->err;
}

/* Reads one marker segment of a (main or first tile-part) header that sets
 * coding parameters; cc_set/qc_set record which components a COC/QCC set. */
static int header_segment(rd_t *r, int mk, hdr_t *h, int *coc_set, int *qcc_set, int main_hdr) {
    int L = rd16(r);
    size_t start = r->pos, end = start + (size_t)L - 2;
    if (L < 2 || end > r->len) return OR_ERR_TRUNCATED;
    int rc = OR_OK;
    switch (mk) {
    case 0xFF52: { /* COD */
        int scod = rd8(r);
        h->g.prog = rd8(r);
        h->g.nlayers = rd16(r);
        h->g.mct = rd8(r);
        h->g.sop = (scod >> 1) & 1;
        h->g.eph = (scod >> 2) & 1;
        ccs_t c;
        rc = parse_cs_params(r, &c, scod & 1);
        if (rc) return rc;
        for (int i = 0; i < h->C; i++)
            if (!coc_set[i]) h->cc[i] = c;
        break;
    }
    case 0xFF53: { /* COC */
        int ci = h->C < 257 ? rd8(r) : rd16(r);
        int scoc = rd8(r);
        if (ci >= h->C) return OR_ERR_MARKER;
        rc = parse_cs_params(r, &h->cc[ci], scoc & 1);
        coc_set[ci] = 1;
        break;
    }
    case 0xFF5C: { /* QCD */
        qcs_t q;
        rc = parse_qc_params(r, &q, end);
        if (rc) return rc;
        for (int i = 0; i < h->C; i++)
            if (!qcc_set[i]) h->qc[i] = q;
        break;
    }
    case 0xFF5D: { /* QCC */
        int ci = h->C < 257 ? rd8(r) : rd16(r);
        if (ci >= h->C) return OR_ERR_MARKER;
        rc = parse_qc_params(r, &h->qc[ci], end);
        qcc_set[ci] = 1;
        break;
    }
    case 0xFF5F: { /* POC: a tile's own entries replace the main header's */
        int wide = h->C >= 257, n = (L - 2) / (wide ? 9 : 7);
        if (n < 1 || (L - 2) % (wide ? 9 : 7)) return OR_ERR_MARKER;
        if (!main_hdr && !h->poc_tile) {
            h->npoc = 0;
            h->poc_tile = 1;
        }
        for (int i = 0; i < n; i++) {
            if (h->npoc >= MAXPOC) return OR_ERR_UNSUPPORTED;
            poc_t *e = &h->poc[h->npoc++];
            e->rs = rd8(r);
            e->cs = wide ? rd16(r) : rd8(r);
            e->le = rd16(r);
            e->re = rd8(r);
            e->ce = wide ? rd16(r) : rd8(r);
            e->prog = rd8(r);
            if (e->ce == 0) e->ce = wide ? 16384 : 256;
            if (e->prog > 4) return OR_ERR_UNSUPPORTED;
        }
        break;
    }
    case 0xFF60: /* PPM: packed packet headers of every tile-part */
        if (!main_hdr || h->nppm >= 256 || L < 3) return OR_ERR_MARKER;
        h->ppm_z[h->nppm] = rd8(r);
        h->ppm_off[h->nppm] = r->pos;
        h->ppm_len[h->nppm] = end - r->pos;
        h->nppm++;
        break;
    case 0xFF64: /* COM */
    case 0xFF55: /* TLM */
    case 0xFF57: /* PLM */
    case 0xFF58: /* PLT */
    case 0xFF63: /* CRG */
        break;
    default:
        return OR_ERR_UNSUPPORTED; /* RGN, PPT outside a tile-part header, unknown */
    }
    if (rc) return rc;
    r->pos = end;
    return r->err;
}

static int parse_main(rd_t *r, hdr_t *h) {
    memset(h, 0, sizeof *h);
    if (rd16(r) != 0xFF4F) return OR_ERR_MARKER;
    if (rd16(r) != 0xFF51) return OR_ERR_MARKER;
    int L = rd16(r);
    size_t end = r->pos + (size_t)L - 2;
    rd16(r); /* Rsiz */
    h->X1 = (int)rd32(r);
    h->Y1 = (int)rd32(r);
    h->X0 = (int)rd32(r);
    h->Y0 = (int)rd32(r);
    h->XT = (int)rd32(r);
    h->YT = (int)rd32(r);
    h->XT0 = (int)rd32(r);
    h->YT0 = (int)rd32(r);
    h->C = rd16(r);
    if (r->err) return r->err;
    if (h->C <= 0 || h->C > MAXC) return OR_ERR_UNSUPPORTED;
    for (int i = 0; i < h->C; i++) {
        int s = rd8(r);
        h->prec[i] = (s & 0x7F) + 1;
        h->sgnd[i] = s >> 7;
        h->dx[i] = rd8(r);
        h->dy[i] = rd8(r);
        if (h->dx[i] < 1 || h->dy[i] < 1) return OR_ERR_MARKER;
        if (h->prec[i] > 16) return OR_ERR_UNSUPPORTED;
    }
    if (r->err) return r->err;
    r->pos = end;
    if (h->X1 <= h->X0 || h->Y1 <= h->Y0 || h->XT <= 0 || h->YT <= 0 || h->XT0 > h->X0 || h->YT0 > h->Y0 ||
        h->XT0 + h->XT <= h->X0 || h->YT0 + h->YT <= h->Y0)
        return OR_ERR_GEOMETRY;
    h->ntx = (int)ceildiv(h->X1 - h->XT0, h->XT);
    h->nty = (int)ceildiv(h->Y1 - h->YT0, h->YT);
    int coc_set[MAXC] = {0}, qcc_set[MAXC] = {0};
    int have_cod = 0, have_qcd = 0;
    for (;;) {
        if (r->pos + 2 > r->len) return OR_ERR_TRUNCATED;
        int mk = rd16(r);
        if (mk == 0xFF90) {
            r->pos -= 2;
            break;
        }
        if (mk == 0xFF52) have_cod = 1;
        if (mk == 0xFF5C) have_qcd = 1;
        int rc = header_segment(r, mk, h, coc_set, qcc_set, 1);
        if (rc) return rc;
    }
    if (!have_cod || !have_qcd) return OR_ERR_MARKER;
    return OR_OK;
}

/* ---- geometry of a tile ---- */

#define MAXSEG 96 /* codeword segments of one code-block (one per pass at most) */

typedef struct {
    int x0, y0, x1, y1;
    int included, zbp, npasses, lblock;
    uint8_t *data;
    int len, cap;
    int nseg, seglen[MAXSEG]; /* BYPASS / TERMALL: bytes of each codeword segment */
} cb_t;

/* codeword segment of pass k under the style */
static int seg_index(int style, int k) {
    int n = 0;
    for (int j = 0; j < k; j++) n += or_seg_end(style, j);
    return n;
}

typedef struct {
    int *val, *low, *known, *parent;
    int nleaf, nnode;
} tt_t;

/* the code-blocks of one band inside one precinct: columns i0..i0+w-1 and
 * rows j0..j0+h-1 of the band's code-block grid, with their tag trees (B.10.2) */
typedef struct {
    int i0, j0, w, h;
    tt_t incl, zbp;
} prc_t;

typedef struct {
    int orient, x0, y0, x1, y1, xcb, ycb;
    int ncx, ncy;
    cb_t *cb;
    prc_t *prc; /* one per precinct of the resolution */
    int eps, mu, mb;
} band_t;

typedef struct {
    int x0, y0, x1, y1, nb;
    int ppx, ppy;   /* precinct exponents PPx, PPy of the resolution */
    long px0, py0;  /* precinct index of the resolution's origin */
    int pw, ph;     /* precincts across and down (0 for an empty resolution) */
    int *nextl;     /* per precinct: the layer of its next packet */
    band_t b[3];
} res_t;

typedef struct {
    int x0, y0, x1, y1;   /* tile-component rectangle: ceil(tile / XRsiz) (B-12) */
    int rx0, ry0, rx1, ry1; /* the tile on the reference grid */
    int dx, dy;           /* XRsiz, YRsiz */
    int nres;
    res_t r[MAXLVL + 1];
    ccs_t cc;
    qcs_t qc;
} tc_t;

static int tt_init(tt_t *t, int w, int h) {
    int n = 0, ww = w, hh = h;
    int lw[40], lh[40], nl = 0;
    memset(t, 0, sizeof *t);
    if (w <= 0 || h <= 0) return OR_OK;
    do {
        lw[nl] = ww;
        lh[nl] = hh;
        n += ww * hh;
        nl++;
        ww = (ww + 1) / 2;
        hh = (hh + 1) / 2;
    } while (lw[nl - 1] * lh[nl - 1] > 1);
    t->nleaf = w * h;
    t->nnode = n;
    t->val = (int *)calloc((size_t)n, sizeof(int));
    t->low = (int *)calloc((size_t)n, sizeof(int));
    t->known = (int *)calloc((size_t)n, sizeof(int));
    t->parent = (int *)calloc((size_t)n, sizeof(int));
    if (!t->val || !t->low || !t->known || !t->parent) return OR_ERR_NOMEM;
    int base = 0;
    for (int l = 0; l < nl; l++) {
        int nbase = base + lw[l] * lh[l];
        for (int y = 0; y < lh[l]; y++)
            for (int x = 0; x < lw[l]; x++)
                t->parent[base + y * lw[l] + x] = (l + 1 < nl) ? nbase + (y / 2) * lw[l + 1] + (x / 2) : -1;
        base = nbase;
    }
    return OR_OK;
}

static void tt_free(tt_t *t) {
    free(t->val);
    free(t->low);
    free(t->known);
    free(t->parent);
}

typedef struct {
    const uint8_t *p;
    size_t len, pos;
    int buf, ct, err;
} bio_t;

/* Packet-header bit reader with bit-stuffing after 0xFF (T.800 B.10.1,
 * TC/io_buffered_stream.c:87-132). */
static int bio_bit(bio_t *b) {
    if (b->ct == 0) {
        int prev = b->buf;
        if (b->pos >= b->len) {
            b->err = OR_ERR_TRUNCATED;
            return 0;
        }
        b->buf = b->p[b->pos++];
        b->ct = (prev == 0xFF) ? 7 : 8;
    }
    b->ct--;
    return (b->buf >> b->ct) & 1;
}
static int bio_bits(bio_t *b, int n) {
    int v = 0;
    for (int i = 0; i < n; i++) v = (v << 1) | bio_bit(b);
    return v;
}
static void bio_align(bio_t *b) {
    /* a header ending on 0xFF owns the following (stuffed) byte too */
    if (b->buf == 0xFF) {
        if (b->pos < b->len)
            b->pos++;
        else
            b->err = OR_ERR_TRUNCATED;
    }
    b->ct = 0;
    b->buf = 0;
}

/* Tag-tree decoding (T.800 B.10.2; TC/codestream_tag_tree_encode.c:97-132):
 * returns 1 when the leaf value is below thr, reading only bits needed. */
static int tt_decode(tt_t *t, bio_t *b, int leaf, int thr) {
    int path[40], np = 0, n = leaf;
    while (n >= 0) {
        path[np++] = n;
        n = t->parent[n];
    }
    int low = 0;
    for (int k = np - 1; k >= 0; k--) {
        int q = path[k];
        if (t->low[q] < low) t->low[q] = low;
        while (!t->known[q] && t->low[q] < thr) {
            if (bio_bit(b))
                t->known[q] = 1;
            else
                t->low[q]++;
        }
        low = t->low[q];
        if (!t->known[q]) break;
    }
    return t->known[leaf] && t->low[leaf] < thr;
}

static int floorlog2(int v) {
    int l = 0;
    while (v > 1) {
        v >>= 1;
        l++;
    }
    return l;
}

static int bits_npasses(bio_t *b) { /* B.10.6 */
    if (!bio_bit(b)) return 1;
    if (!bio_bit(b)) return 2;
    int n = bio_bits(b, 2);
    if (n != 3) return 3 + n;
    n = bio_bits(b, 5);
    if (n != 31) return 6 + n;
    return 37 + bio_bits(b, 7);
}

static int cb_append(cb_t *c, const uint8_t *p, int n) {
    if (c->len + n > c->cap) {
        int nc = (c->len + n) * 2 + 64;
        uint8_t *d = (uint8_t *)realloc(c->data, (size_t)nc);
        if (!d) return OR_ERR_NOMEM;
        c->data = d;
        c->cap = nc;
    }
    memcpy(c->data + c->len, p, (size_t)n);
    c->len += n;
    return OR_OK;
}

static int build_tc(const hdr_t *h, int t, int c, tc_t *tc) {
    int p = t % h->ntx, q = t / h->ntx;
    int tx0 = h->XT0 + p * h->XT, ty0 = h->YT0 + q * h->YT;
    int tx1 = tx0 + h->XT, ty1 = ty0 + h->YT;
    if (tx0 < h->X0) tx0 = h->X0;
    if (ty0 < h->Y0) ty0 = h->Y0;
    if (tx1 > h->X1) tx1 = h->X1;
    if (ty1 > h->Y1) ty1 = h->Y1;
    memset(tc, 0, sizeof *tc);
    tc->rx0 = tx0;
    tc->ry0 = ty0;
    tc->rx1 = tx1;
    tc->ry1 = ty1;
    tc->dx = h->dx[c];
    tc->dy = h->dy[c];
    /* the tile-component rectangle (B-12): every later rectangle derives from it */
    tx0 = (int)ceildiv(tx0, tc->dx);
    ty0 = (int)ceildiv(ty0, tc->dy);
    tx1 = (int)ceildiv(tx1, tc->dx);
    ty1 = (int)ceildiv(ty1, tc->dy);
    tc->x0 = tx0;
    tc->y0 = ty0;
    tc->x1 = tx1;
    tc->y1 = ty1;
    tc->cc = h->cc[c];
    tc->qc = h->qc[c];
    int NL = tc->cc.nlev;
    tc->nres = NL + 1;
    if (tc->cc.cbsty & ~OR_STYLE_DECODABLE) return OR_ERR_UNSUPPORTED;
    if (tc->qc.style == 0 && tc->qc.n < 3 * NL + 1) return OR_ERR_MARKER;
    if (tc->qc.style == 2 && tc->qc.n < 3 * NL + 1) return OR_ERR_MARKER;
    if (tc->qc.style == 1 && tc->qc.n < 1) return OR_ERR_MARKER;
    for (int r = 0; r <= NL; r++) {
        res_t *R = &tc->r[r];
        res_rect(tx0, ty0, tx1, ty1, NL - r, &R->x0, &R->y0, &R->x1, &R->y1);
        /* precinct partition of the resolution, anchored at 0 (B.6, B-16) */
        R->ppx = tc->cc.ppx[r];
        R->ppy = tc->cc.ppy[r];
        if (r > 0 && (R->ppx == 0 || R->ppy == 0)) return OR_ERR_MARKER;
        if (R->x1 > R->x0 && R->y1 > R->y0) {
            R->px0 = floordiv(R->x0, 1L << R->ppx);
            R->py0 = floordiv(R->y0, 1L << R->ppy);
            long npx = ceildiv(R->x1, 1L << R->ppx) - R->px0;
            long npy = ceildiv(R->y1, 1L << R->ppy) - R->py0;
            if (npx * npy > (1L << 22)) return OR_ERR_UNSUPPORTED;
            R->pw = (int)npx;
            R->ph = (int)npy;
        }
        R->nextl = (int *)calloc((size_t)(R->pw * R->ph > 0 ? R->pw * R->ph : 1), sizeof(int));
        if (!R->nextl) return OR_ERR_NOMEM;
        R->nb = r == 0 ? 1 : 3;
        int nb = r == 0 ? NL : NL - r + 1;
        for (int k = 0; k < R->nb; k++) {
            band_t *B = &R->b[k];
            int o = r == 0 ? 0 : k + 1; /* 1 HL, 2 LH, 3 HH */
            int xob = (o == 1 || o == 3), yob = (o == 2 || o == 3);
            B->orient = o;
            if (nb == 0) {
                B->x0 = tx0; B->y0 = ty0; B->x1 = tx1; B->y1 = ty1;
            } else {
                long s = 1L << nb, hs = 1L << (nb - 1);
                B->x0 = (int)ceildiv(tx0 - hs * xob, s);
                B->y0 = (int)ceildiv(ty0 - hs * yob, s);
                B->x1 = (int)ceildiv(tx1 - hs * xob, s);
                B->y1 = (int)ceildiv(ty1 - hs * yob, s);
            }
            int xcb = tc->cc.xcb, ycb = tc->cc.ycb;
            int px = r == 0 ? tc->cc.ppx[r] : tc->cc.ppx[r] - 1;
            int py = r == 0 ? tc->cc.ppy[r] : tc->cc.ppy[r] - 1;
            if (xcb > px) xcb = px;
            if (ycb > py) ycb = py;
            B->xcb = xcb;
            B->ycb = ycb;
            if (B->x1 > B->x0 && B->y1 > B->y0) {
                B->ncx = (int)(ceildiv(B->x1, 1L << xcb) - floordiv(B->x0, 1L << xcb));
                B->ncy = (int)(ceildiv(B->y1, 1L << ycb) - floordiv(B->y0, 1L << ycb));
            } else {
                B->ncx = B->ncy = 0;
            }
            int ncb = B->ncx * B->ncy;
            B->cb = (cb_t *)calloc((size_t)(ncb ? ncb : 1), sizeof(cb_t));
            if (!B->cb) return OR_ERR_NOMEM;
            int gx0 = (int)floordiv(B->x0, 1L << xcb), gy0 = (int)floordiv(B->y0, 1L << ycb);
            for (int j = 0; j < B->ncy; j++)
                for (int i = 0; i < B->ncx; i++) {
                    cb_t *C = &B->cb[j * B->ncx + i];
                    int cx0 = (gx0 + i) << xcb, cy0 = (gy0 + j) << ycb;
                    C->x0 = cx0 > B->x0 ? cx0 : B->x0;
                    C->y0 = cy0 > B->y0 ? cy0 : B->y0;
                    C->x1 = (cx0 + (1 << xcb)) < B->x1 ? cx0 + (1 << xcb) : B->x1;
                    C->y1 = (cy0 + (1 << ycb)) < B->y1 ? cy0 + (1 << ycb) : B->y1;
                    C->lblock = 3;
                }
            /* the band's share of each precinct: its precinct partition is the
             * resolution's halved (r > 0), and code-blocks never straddle it */
            int np = R->pw * R->ph;
            B->prc = (prc_t *)calloc((size_t)(np ? np : 1), sizeof(prc_t));
            if (!B->prc) return OR_ERR_NOMEM;
            for (int q = 0; q < np; q++) {
                prc_t *P = &B->prc[q];
                if (B->ncx > 0 && B->ncy > 0) {
                    long X = R->px0 + q % R->pw, Y = R->py0 + q / R->pw;
                    long i0 = (X << (px - xcb)) - gx0, i1 = ((X + 1) << (px - xcb)) - gx0;
                    long j0 = (Y << (py - ycb)) - gy0, j1 = ((Y + 1) << (py - ycb)) - gy0;
                    if (i0 < 0) i0 = 0;
                    if (j0 < 0) j0 = 0;
                    if (i1 > B->ncx) i1 = B->ncx;
                    if (j1 > B->ncy) j1 = B->ncy;
                    if (i1 > i0 && j1 > j0) {
                        P->i0 = (int)i0;
                        P->j0 = (int)j0;
                        P->w = (int)(i1 - i0);
                        P->h = (int)(j1 - j0);
                    }
                }
                int rc = tt_init(&P->incl, P->w, P->h);
                if (!rc) rc = tt_init(&P->zbp, P->w, P->h);
                if (rc) return rc;
            }
            /* quantization exponent / mantissa, E.1 (derived: E-5) */
            int idx = r == 0 ? 0 : 3 * (r - 1) + k + 1;
            if (tc->qc.style == 1) {
                B->eps = tc->qc.eps[0] - NL + nb;
                if (r == 0) B->eps = tc->qc.eps[0];
                B->mu = tc->qc.mu[0];
            } else {
                B->eps = tc->qc.eps[idx];
                B->mu = tc->qc.mu[idx];
            }
            B->mb = tc->qc.guard + B->eps - 1; /* M_b, E-2 */
        }
    }
    return OR_OK;
}

static void free_tc(tc_t *tc) {
    for (int r = 0; r < tc->nres; r++) free(tc->r[r].nextl);
    for (int r = 0; r < tc->nres; r++)
        for (int k = 0; k < tc->r[r].nb; k++) {
            band_t *B = &tc->r[r].b[k];
            if (B->cb)
                for (int i = 0; i < B->ncx * B->ncy; i++) free(B->cb[i].data);
            free(B->cb);
            if (B->prc)
                for (int q = 0; q < tc->r[r].pw * tc->r[r].ph; q++) {
                    tt_free(&B->prc[q].incl);
                    tt_free(&B->prc[q].zbp);
                }
            free(B->prc);
        }
}

/* One packet of precinct q (B.9, B.10): header then body.  The header comes
 * from d at *pos, or from the packed headers hd at *hpos when hd is set. */
static int read_packet(const hdr_t *h, tc_t *tc, int r, int q, int layer, const uint8_t *d, size_t len,
                       size_t *pos, const uint8_t *hd, size_t hlen, size_t *hpos) {
    res_t *R = &tc->r[r];
    if (q >= R->pw * R->ph) return OR_OK; /* no precinct, no packet */
    if (h->g.sop && *pos + 6 <= len && d[*pos] == 0xFF && d[*pos + 1] == 0x91) *pos += 6;
    bio_t b;
    memset(&b, 0, sizeof b);
    b.p = hd ? hd : d;
    b.len = hd ? hlen : len;
    b.pos = hd ? *hpos : *pos;
    /* included segment lengths in header order */
    int present = bio_bit(&b);
    int nseg = 0, cap = 0;
    int *seglen = NULL, *segidx = NULL;
    cb_t **segcb = NULL;
    const int style = tc->cc.cbsty;
    const int multi = (style & (OR_STYLE_BYPASS | OR_STYLE_TERMALL)) != 0;
    if (present) {
        for (int k = 0; k < R->nb; k++) {
            band_t *B = &R->b[k];
            prc_t *P = &B->prc[q];
            for (int i = 0; i < P->w * P->h; i++) {
                cb_t *C = &B->cb[(P->j0 + i / P->w) * B->ncx + P->i0 + i % P->w];
                int inc;
                if (!C->included)
                    inc = tt_decode(&P->incl, &b, i, layer + 1);
                else
                    inc = bio_bit(&b);
                if (!inc) continue;
                if (!C->included) {
                    int z = 0;
                    while (!tt_decode(&P->zbp, &b, i, z + 1)) z++;
                    C->zbp = z;
                    C->included = 1;
                }
                int np = bits_npasses(&b);
                while (bio_bit(&b)) C->lblock++;
                /* one length per codeword segment the passes touch (B.10.7.2) */
                int k0 = C->npasses, k = k0;
                while (k < k0 + np && !b.err) {
                    int e = k;
                    if (multi)
                        while (e < k0 + np - 1 && !or_seg_end(style, e)) e++;
                    else
                        e = k0 + np - 1;
                    int L = bio_bits(&b, C->lblock + floorlog2(e - k + 1));
                    if (nseg == cap) {
                        cap = cap * 2 + 16;
                        seglen = (int *)realloc(seglen, sizeof(int) * (size_t)cap);
                        segidx = (int *)realloc(segidx, sizeof(int) * (size_t)cap);
                        segcb = (cb_t **)realloc(segcb, sizeof(cb_t *) * (size_t)cap);
                        if (!seglen || !segcb || !segidx) return OR_ERR_NOMEM;
                    }
                    seglen[nseg] = L;
                    segidx[nseg] = multi ? seg_index(style, k) : 0;
                    segcb[nseg] = C;
                    nseg++;
                    k = e + 1;
                }
                C->npasses += np;
                if (b.err) break;
            }
        }
    }
    bio_align(&b);
    if (b.err) {
        free(seglen);
        free(segidx);
        free(segcb);
        return b.err;
    }
    if (hd)
        *hpos = b.pos;
    else
        *pos = b.pos;
    if (h->g.eph) {
        /* EPH ends the header, wherever the header is */
        const uint8_t *e = hd ? hd : d;
        size_t elen = hd ? hlen : len, *ep = hd ? hpos : pos;
        if (*ep + 2 <= elen && e[*ep] == 0xFF && e[*ep + 1] == 0x92)
            *ep += 2;
        else {
            free(seglen);
            free(segidx);
            free(segcb);
            return OR_ERR_MARKER;
        }
    }
    for (int s = 0; s < nseg; s++) {
        if (*pos + (size_t)seglen[s] > len) {
            free(seglen);
            free(segidx);
            free(segcb);
            return OR_ERR_TRUNCATED;
        }
        int rc = cb_append(segcb[s], d + *pos, seglen[s]);
        if (rc) return rc;
        if (segidx[s] >= MAXSEG) return OR_ERR_UNSUPPORTED;
        segcb[s]->seglen[segidx[s]] += seglen[s];
        if (segidx[s] + 1 > segcb[s]->nseg) segcb[s]->nseg = segidx[s] + 1;
        *pos += (size_t)seglen[s];
    }
    free(seglen);
    free(segidx);
    free(segcb);
    return OR_OK;
}

/* Precinct of tile-component tc, resolution r at reference-grid position
 * (x, y) of a position-driven progression, or -1 when no precinct of that
 * resolution starts there (B.12.1.3-B.12.1.5: a precinct starts where the
 * position is a multiple of its size projected onto the reference grid, or
 * at the tile origin when the resolution's origin is not). */
static int prc_at(const tc_t *tc, int r, long x, long y) {
    const res_t *R = &tc->r[r];
    if (R->pw == 0 || R->ph == 0) return -1;
    int lv = tc->nres - 1 - r;
    long sx = 1L << (R->ppx + lv), sy = 1L << (R->ppy + lv);
    int okx = x % (tc->dx * sx) == 0 || (x == tc->rx0 && (((long)R->x0 << lv) % sx) != 0);
    int oky = y % (tc->dy * sy) == 0 || (y == tc->ry0 && (((long)R->y0 << lv) % sy) != 0);
    if (!okx || !oky) return -1;
    long i = floordiv(ceildiv(x, (long)tc->dx << lv), 1L << R->ppx) - R->px0;
    long j = floordiv(ceildiv(y, (long)tc->dy << lv), 1L << R->ppy) - R->py0;
    if (i < 0 || j < 0 || i >= R->pw || j >= R->ph) return -1;
    return (int)(j * R->pw + i);
}

#define OR_END 1000 /* internal: the tile's data ended at a packet boundary */

/* Reads packet (c, r, q, l) unless that precinct has no packets or its
 * packet of layer l came earlier (B.12.2: a packet appears once, and the
 * packets of a precinct come in layer order). */
typedef struct {
    const uint8_t *d, *hd; /* tile data; packed headers or NULL */
    size_t len, hlen, pos, hpos;
} pkt_src_t;

static int emit(const hdr_t *h, tc_t *tc, int r, int q, int l, pkt_src_t *S) {
    if (r >= tc->nres) return OR_OK;
    res_t *R = &tc->r[r];
    if (q < 0 || q >= R->pw * R->ph || R->nextl[q] != l) return OR_OK;
    if (S->hd ? S->hpos >= S->hlen : S->pos >= S->len) return OR_END;
    R->nextl[q]++;
    return read_packet(h, tc, r, q, l, S->d, S->len, &S->pos, S->hd, S->hlen, &S->hpos);
}

/* The packets of one progression volume (B.12.1) in order v->prog.  xs, ys:
 * position steps of the position-driven orders, no larger than any precinct
 * projected onto the reference grid. */
static int read_volume(const hdr_t *h, tc_t *tcs, const poc_t *v, long xs, long ys, pkt_src_t *S) {
    const long tx0 = tcs[0].rx0, ty0 = tcs[0].ry0, tx1 = tcs[0].rx1, ty1 = tcs[0].ry1;
    const int cs = v->cs, ce = v->ce < h->C ? v->ce : h->C, rs = v->rs, re = v->re;
    const int le = v->le < h->g.nlayers ? v->le : h->g.nlayers;
    int rc;
#define EM(c_, r_, q_, l_)                                          \
    do {                                                            \
        rc = emit(h, &tcs[c_], r_, q_, l_, S);                      \
        if (rc) return rc;                                          \
    } while (0)
#define NPRC(c_, r_) ((r_) < tcs[c_].nres ? tcs[c_].r[r_].pw * tcs[c_].r[r_].ph : 0)
    switch (v->prog) {
    case 0: /* LRCP */
        for (int l = 0; l < le; l++)
            for (int r = rs; r < re; r++)
                for (int c = cs; c < ce; c++)
                    for (int q = 0; q < NPRC(c, r); q++) EM(c, r, q, l);
        break;
    case 1: /* RLCP */
        for (int r = rs; r < re; r++)
            for (int l = 0; l < le; l++)
                for (int c = cs; c < ce; c++)
                    for (int q = 0; q < NPRC(c, r); q++) EM(c, r, q, l);
        break;
    case 2: /* RPCL */
        for (int r = rs; r < re; r++)
            for (long y = ty0; y < ty1; y += ys - y % ys)
                for (long x = tx0; x < tx1; x += xs - x % xs)
                    for (int c = cs; c < ce; c++) {
                        if (r >= tcs[c].nres) continue;
                        int q = prc_at(&tcs[c], r, x, y);
                        if (q >= 0)
                            for (int l = 0; l < le; l++) EM(c, r, q, l);
                    }
        break;
    case 3: /* PCRL */
        for (long y = ty0; y < ty1; y += ys - y % ys)
            for (long x = tx0; x < tx1; x += xs - x % xs)
                for (int c = cs; c < ce; c++)
                    for (int r = rs; r < re && r < tcs[c].nres; r++) {
                        int q = prc_at(&tcs[c], r, x, y);
                        if (q >= 0)
                            for (int l = 0; l < le; l++) EM(c, r, q, l);
                    }
        break;
    case 4: /* CPRL */
        for (int c = cs; c < ce; c++)
            for (long y = ty0; y < ty1; y += ys - y % ys)
                for (long x = tx0; x < tx1; x += xs - x % xs)
                    for (int r = rs; r < re && r < tcs[c].nres; r++) {
                        int q = prc_at(&tcs[c], r, x, y);
                        if (q >= 0)
                            for (int l = 0; l < le; l++) EM(c, r, q, l);
                    }
        break;
    default:
        return OR_ERR_UNSUPPORTED;
    }
#undef EM
#undef NPRC
    return OR_OK;
}

/* Runs every packet of a tile: the POC volumes in order, then the COD
 * progression over whatever they left (B.12). */
static int read_tile_packets(const hdr_t *h, tc_t *tcs, const uint8_t *d, size_t len, const uint8_t *hd,
                             size_t hlen) {
    pkt_src_t S = {d, hd, len, hlen, 0, 0};
    int maxr = 0;
    for (int c = 0; c < h->C; c++)
        if (tcs[c].nres > maxr) maxr = tcs[c].nres;
    /* position steps: the smallest precinct projected onto the reference grid */
    long xs = 1L << 40, ys = 1L << 40;
    for (int c = 0; c < h->C; c++)
        for (int r = 0; r < tcs[c].nres; r++) {
            int lv = tcs[c].nres - 1 - r;
            long a = (long)tcs[c].dx << (tcs[c].r[r].ppx + lv), b = (long)tcs[c].dy << (tcs[c].r[r].ppy + lv);
            if (a < xs) xs = a;
            if (b < ys) ys = b;
        }
    int rc = OR_OK;
    for (int i = 0; i < h->npoc && !rc; i++) rc = read_volume(h, tcs, &h->poc[i], xs, ys, &S);
    if (!rc) {
        poc_t all = {0, 0, h->g.nlayers, maxr, h->C, h->g.prog};
        rc = read_volume(h, tcs, &all, xs, ys, &S);
    }
    return rc == OR_END ? OR_OK : rc;
}

/* Collects the data of every tile-part of tile t into one buffer and applies
 * tile-part header overrides (A.4.2).  Packed packet headers (PPM of the main
 * header, PPT of the tile-part headers; A.7.4, A.7.5) are gathered per tile
 * in the same tile-part order. */
typedef struct {
    uint8_t *data, *hdr;
    size_t len, cap, hlen, hcap;
    hdr_t h; /* effective parameters for the tile */
    int seen, packed;
} tiledata_t;

static int buf_append(uint8_t **p, size_t *len, size_t *cap, const uint8_t *src, size_t n) {
    if (*len + n > *cap) {
        size_t nc = (*len + n) * 2 + 64;
        uint8_t *q = (uint8_t *)realloc(*p, nc);
        if (!q) return OR_ERR_NOMEM;
        *p = q;
        *cap = nc;
    }
    if (n) memcpy(*p + *len, src, n);
    *len += n;
    return OR_OK;
}

/* The payloads of n marker segments in order of their index byte z. */
static int concat_by_index(const uint8_t *cs, int n, const int *z, const size_t *off, const size_t *slen,
                           uint8_t **out, size_t *olen, size_t *ocap) {
    for (int want = 0, done = 0; done < n; want++) {
        if (want > 255) return OR_ERR_MARKER;
        for (int i = 0; i < n; i++)
            if (z[i] == want) {
                int rc = buf_append(out, olen, ocap, cs + off[i], slen[i]);
                if (rc) return rc;
                done++;
            }
    }
    return OR_OK;
}

static int gather_tiles(const uint8_t *cs, size_t len, const hdr_t *mh, tiledata_t *td, size_t start) {
    rd_t r = {cs, len, start, 0};
    int nt = mh->ntx * mh->nty;
    uint8_t *ppm = NULL;
    size_t ppm_len = 0, ppm_cap = 0, ppm_pos = 0;
    int rc = concat_by_index(cs, mh->nppm, mh->ppm_z, mh->ppm_off, mh->ppm_len, &ppm, &ppm_len, &ppm_cap);
    for (; !rc;) {
        if (r.pos + 2 > len) break; /* tolerate missing EOC */
        int mk = rd16(&r);
        if (mk == 0xFFD9) break;
        if (mk != 0xFF90) {
            rc = OR_ERR_MARKER;
            break;
        }
        size_t sot = r.pos - 2;
        int Lsot = rd16(&r);
        int I = rd16(&r);
        uint32_t Psot = rd32(&r);
        int TP = rd8(&r);
        rd8(&r);
        if (r.err) {
            rc = r.err;
            break;
        }
        size_t end = Psot ? sot + Psot : len - 2;
        if (Lsot != 10 || I >= nt) rc = OR_ERR_MARKER;
        else if (end > len || end < r.pos) rc = OR_ERR_TRUNCATED;
        if (rc) break;
        tiledata_t *T = &td[I];
        if (!T->seen) {
            T->h = *mh;
            T->seen = 1;
        }
        int coc_set[MAXC] = {0}, qcc_set[MAXC] = {0};
        int nppt = 0, ppt_z[256];
        size_t ppt_off[256], ppt_len[256];
        for (; !rc;) {
            int m = rd16(&r);
            if (r.err) {
                rc = r.err;
                break;
            }
            if (m == 0xFF93) break;
            if (TP != 0 && (m == 0xFF52 || m == 0xFF53 || m == 0xFF5C || m == 0xFF5D)) {
                rc = OR_ERR_MARKER;
                break;
            }
            if (m == 0xFF61) { /* PPT */
                int L = rd16(&r);
                size_t e = r.pos + (size_t)L - 2;
                if (L < 3 || e > len || nppt >= 256) {
                    rc = OR_ERR_MARKER;
                    break;
                }
                ppt_z[nppt] = rd8(&r);
                ppt_off[nppt] = r.pos;
                ppt_len[nppt] = e - r.pos;
                nppt++;
                r.pos = e;
                continue;
            }
            rc = header_segment(&r, m, &T->h, coc_set, qcc_set, 0);
        }
        if (rc) break;
        if (nppt && mh->nppm) rc = OR_ERR_MARKER; /* PPM and PPT exclude each other */
        if (!rc && nppt) {
            rc = concat_by_index(cs, nppt, ppt_z, ppt_off, ppt_len, &T->hdr, &T->hlen, &T->hcap);
            T->packed = 1;
        }
        if (!rc && mh->nppm) { /* Nppm, then Ippm, per tile-part in codestream order */
            if (ppm_pos + 4 > ppm_len) {
                rc = OR_ERR_TRUNCATED;
                break;
            }
            size_t n = ((size_t)ppm[ppm_pos] << 24) | ((size_t)ppm[ppm_pos + 1] << 16) |
                       ((size_t)ppm[ppm_pos + 2] << 8) | ppm[ppm_pos + 3];
            ppm_pos += 4;
            if (ppm_pos + n > ppm_len) {
                rc = OR_ERR_TRUNCATED;
                break;
            }
            rc = buf_append(&T->hdr, &T->hlen, &T->hcap, ppm + ppm_pos, n);
            ppm_pos += n;
            T->packed = 1;
        }
        if (!rc) rc = buf_append(&T->data, &T->len, &T->cap, cs + r.pos, end - r.pos);
        r.pos = end;
    }
    free(ppm);
    return rc;
}

static size_t find_codestream(const uint8_t *cs, size_t len, size_t *clen) {
    /* JP2 wrapper: signature box, then boxes; the codestream is in 'jp2c'. */
    if (len >= 12 && cs[0] == 0 && cs[1] == 0 && cs[2] == 0 && cs[3] == 12 && cs[4] == 'j' && cs[5] == 'P') {
        size_t pos = 0;
        while (pos + 8 <= len) {
            uint64_t L = ((uint32_t)cs[pos] << 24) | ((uint32_t)cs[pos + 1] << 16) | ((uint32_t)cs[pos + 2] << 8) | cs[pos + 3];
            size_t hdr = 8;
            if (L == 1) {
                if (pos + 16 > len) return (size_t)-1;
                L = 0;
                for (int i = 0; i < 8; i++) L = (L << 8) | cs[pos + 8 + i];
                hdr = 16;
            } else if (L == 0)
                L = len - pos;
            if (memcmp(cs + pos + 4, "jp2c", 4) == 0) {
                *clen = (size_t)L - hdr;
                if (pos + L > len) *clen = len - pos - hdr;
                return pos + hdr;
            }
            if (L < hdr) return (size_t)-1;
            pos += (size_t)L;
        }
        return (size_t)-1;
    }
    *clen = len;
    return 0;
}

int or_read_info(const uint8_t *cs0, size_t len0, or_info *info) {
    size_t len, off = find_codestream(cs0, len0, &len);
    if (off == (size_t)-1) return -OR_ERR_MARKER;
    const uint8_t *cs = cs0 + off;
    hdr_t h;
    rd_t r = {cs, len, 0, 0};
    int rc = parse_main(&r, &h);
    if (rc) return -rc;
    memset(info, 0, sizeof *info);
    info->x0 = h.X0;
    info->y0 = h.Y0;
    info->x1 = h.X1;
    info->y1 = h.Y1;
    info->tx0 = h.XT0;
    info->ty0 = h.YT0;
    info->tw = h.XT;
    info->th = h.YT;
    info->ntx = h.ntx;
    info->nty = h.nty;
    info->ncomp = h.C;
    for (int i = 0; i < h.C; i++) {
        info->prec[i] = h.prec[i];
        info->sgnd[i] = h.sgnd[i];
    }
    info->nlevels = h.cc[0].nlev;
    info->reversible = h.cc[0].xform == 1;
    info->mct = h.g.mct;
    info->cbw = 1 << h.cc[0].xcb;
    info->cbh = 1 << h.cc[0].ycb;
    info->nlayers = h.g.nlayers;
    info->progression = h.g.prog;
    info->qstyle = h.qc[0].style;
    info->guard_bits = h.qc[0].guard;
    for (int i = 0; i < h.C && i < 16; i++) {
        info->dx[i] = h.dx[i];
        info->dy[i] = h.dy[i];
    }
    return OR_OK;
}

static uint32_t fnv1a(const uint8_t *p, int n) {
    uint32_t hsh = 2166136261u;
    for (int i = 0; i < n; i++) {
        hsh ^= p[i];
        hsh *= 16777619u;
    }
    return hsh;
}

/* Shared front end: parse, gather tile data, build every tile-component and
 * read its packets; the callback consumes each decoded tile. */
typedef int (*tile_fn)(const hdr_t *h, int t, tc_t *tcs, void *user);

static int walk_tiles(const uint8_t *cs0, size_t len0, tile_fn fn, void *user) {
    size_t len, off = find_codestream(cs0, len0, &len);
    if (off == (size_t)-1) return OR_ERR_MARKER;
    const uint8_t *cs = cs0 + off;
    hdr_t h;
    rd_t r = {cs, len, 0, 0};
    int rc = parse_main(&r, &h);
    if (rc) return rc;
    int nt = h.ntx * h.nty;
    tiledata_t *td = (tiledata_t *)calloc((size_t)nt, sizeof(tiledata_t));
    if (!td) return OR_ERR_NOMEM;
    rc = gather_tiles(cs, len, &h, td, r.pos);
    for (int t = 0; t < nt && !rc; t++) {
        const hdr_t *th = td[t].seen ? &td[t].h : &h;
        tc_t *tcs = (tc_t *)calloc((size_t)h.C, sizeof(tc_t));
        if (!tcs) {
            rc = OR_ERR_NOMEM;
            break;
        }
        for (int c = 0; c < h.C && !rc; c++) rc = build_tc(th, t, c, &tcs[c]);
        if (!rc && td[t].seen)
            rc = read_tile_packets(th, tcs, td[t].data, td[t].len, td[t].packed ? td[t].hdr : NULL, td[t].hlen);
        if (!rc) rc = fn(th, t, tcs, user);
        for (int c = 0; c < h.C; c++) free_tc(&tcs[c]);
        free(tcs);
    }
    for (int t = 0; t < nt; t++) {
        free(td[t].data);
        free(td[t].hdr);
    }
    free(td);
    return rc;
}

typedef struct {
    or_cblk *out;
    int max, n;
    int32_t *nsym; /* when set: MQ symbols decoded per code-block */
} list_ctx;

static int list_tile(const hdr_t *h, int t, tc_t *tcs, void *user) {
    list_ctx *L = (list_ctx *)user;
    for (int c = 0; c < h->C; c++)
        for (int r = 0; r < tcs[c].nres; r++)
            for (int k = 0; k < tcs[c].r[r].nb; k++) {
                band_t *B = &tcs[c].r[r].b[k];
                for (int i = 0; i < B->ncx * B->ncy; i++) {
                    cb_t *C = &B->cb[i];
                    if (L->nsym && L->n < L->max) {
                        int w = C->x1 - C->x0, hh = C->y1 - C->y0, k = 0;
                        if (w > 0 && hh > 0 && C->npasses > 0) {
                            int32_t *mag = (int32_t *)malloc(sizeof(int32_t) * (size_t)w * hh);
                            int8_t *sg = (int8_t *)malloc((size_t)w * hh), *lp = (int8_t *)malloc((size_t)w * hh);
                            if (mag && sg && lp)
                                k = or_t1_decode_seg(C->data, C->len, C->nseg > 1 ? C->seglen : &C->len,
                                                     C->nseg > 1 ? C->nseg : 1, w, hh, B->orient, B->mb - C->zbp,
                                                     C->npasses, tcs[c].cc.cbsty, mag, sg, lp);
                            free(mag);
                            free(sg);
                            free(lp);
                        }
                        L->nsym[L->n] = k;
                    } else if (L->out && L->n < L->max) {
                        or_cblk *o = &L->out[L->n];
                        o->tile = t;
                        o->comp = c;
                        o->res = r;
                        o->band = B->orient;
                        o->x0 = C->x0;
                        o->y0 = C->y0;
                        o->x1 = C->x1;
                        o->y1 = C->y1;
                        o->npasses = C->npasses;
                        o->zero_bp = C->zbp;
                        o->mb = B->mb;
                        o->nbytes = C->len;
                        o->crc = fnv1a(C->data, C->len);
                    }
                    L->n++;
                }
            }
    return OR_OK;
}

int or_list_cblks(const uint8_t *cs, size_t len, or_cblk *out, int max) {
    list_ctx L = {out, max, 0, NULL};
    int rc = walk_tiles(cs, len, list_tile, &L);
    return rc ? -rc : L.n;
}

int or_t1_symbol_counts(const uint8_t *cs, size_t len, int32_t *nsym, int max) {
    list_ctx L = {NULL, max, 0, nsym};
    int rc = walk_tiles(cs, len, list_tile, &L);
    return rc ? -rc : L.n;
}

/* ---- full decode ---- */

typedef struct {
    int32_t *out;
    double *out_f;
    int W, H, X0, Y0;
    int reduce; /* finest DWT levels not reconstructed */
    int ix0, iy0, ix1, iy1; /* the image area at full resolution */
} dec_ctx;

/* Mallat position of a band of resolution r inside the tile-component array. */
static void band_origin(const tc_t *tc, int r, int orient, int *ox, int *oy) {
    *ox = 0;
    *oy = 0;
    if (r == 0) return;
    const res_t *lo = &tc->r[r - 1];
    if (orient == 1 || orient == 3) *ox = lo->x1 - lo->x0;
    if (orient == 2 || orient == 3) *oy = lo->y1 - lo->y0;
}

/* Decodes one tile into D->out.  With D->reduce = k > 0 only resolutions
 * 0..nres-1-k are decoded and inverse-transformed: the tile-component becomes
 * its resolution nres-1-k rectangle (coordinates ceil(x / 2^k), B.5), and the
 * inverse DWT runs nlev - k levels from that rectangle's origin. */
static int decode_tile(const hdr_t *h, int t, tc_t *tcs, void *user) {
    dec_ctx *D = (dec_ctx *)user;
    int C = h->C;
    if (C <= 0) return OR_ERR_UNSUPPORTED;
    for (int c = 0; c < C; c++)
        if (tcs[c].nres - 1 < D->reduce) return OR_ERR_UNSUPPORTED;
    int tws[MAXC], ths[MAXC]; /* each tile-component's (reduced) size */
    for (int c = 0; c < C; c++) {
        const res_t *top = &tcs[c].r[tcs[c].nres - 1 - D->reduce];
        tws[c] = top->x1 - top->x0;
        ths[c] = top->y1 - top->y0;
    }
    double **planes = (double **)calloc((size_t)C, sizeof(double *));
    if (!planes) return OR_ERR_NOMEM;
    int rc = OR_OK;
    for (int c = 0; c < C && !rc; c++) {
        tc_t *tc = &tcs[c];
        int rev = tc->cc.xform == 1;
        const int tw = tws[c], th = ths[c];
        const long n = (long)tw * th;
        int32_t *ia = rev ? (int32_t *)calloc((size_t)n, sizeof(int32_t)) : NULL;
        double *da = (double *)calloc((size_t)n, sizeof(double));
        planes[c] = da;
        if ((rev && !ia) || !da) {
            free(ia);
            rc = OR_ERR_NOMEM;
            break;
        }
        for (int r = 0; r < tc->nres - D->reduce && !rc; r++)
            for (int k = 0; k < tc->r[r].nb && !rc; k++) {
                band_t *B = &tc->r[r].b[k];
                int ox, oy;
                band_origin(tc, r, B->orient, &ox, &oy);
                /* Delta_b, E.1.1.1 with R_b = R_I + log2 gain_b */
                int gain = B->orient == 0 ? 0 : (B->orient == 3 ? 2 : 1);
                double delta = ldexp(1.0 + B->mu / 2048.0, h->prec[c] + gain - B->eps);
                for (int i = 0; i < B->ncx * B->ncy; i++) {
                    cb_t *CB = &B->cb[i];
                    int w = CB->x1 - CB->x0, hh = CB->y1 - CB->y0;
                    if (w <= 0 || hh <= 0) continue;
                    int32_t *mag = (int32_t *)malloc(sizeof(int32_t) * (size_t)w * hh);
                    int8_t *sg = (int8_t *)malloc((size_t)w * hh);
                    int8_t *lp = (int8_t *)malloc((size_t)w * hh);
                    if (!mag || !sg || !lp) {
                        free(mag);
                        free(sg);
                        free(lp);
                        rc = OR_ERR_NOMEM;
                        break;
                    }
                    int numbps = B->mb - CB->zbp;
                    int nr = or_t1_decode_seg(CB->data, CB->len, CB->nseg > 1 ? CB->seglen : &CB->len,
                                              CB->nseg > 1 ? CB->nseg : 1, w, hh, B->orient, numbps, CB->npasses,
                                              tc->cc.cbsty, mag, sg, lp);
                    if (nr < 0) rc = -nr;
                    for (int y = 0; y < hh && !rc; y++)
                        for (int x = 0; x < w; x++) {
                            int j = y * w + x;
                            int ax = ox + CB->x0 - B->x0 + x, ay = oy + CB->y0 - B->y0 + y;
                            long o = (long)ay * tw + ax;
                            if (lp[j] < 0) continue;
                            if (rev) {
                                int32_t v = mag[j] + (lp[j] >= 1 ? (1 << (lp[j] - 1)) : 0);
                                ia[o] = sg[j] ? -v : v;
                            } else {
                                double v = ((double)mag[j] + ldexp(0.5, lp[j])) * delta;
                                da[o] = sg[j] ? -v : v;
                            }
                        }
                    free(mag);
                    free(sg);
                    free(lp);
                }
            }
        if (!rc) {
            const res_t *top = &tc->r[tc->nres - 1 - D->reduce];
            if (rev) {
                rc = or_idwt53(ia, tw, th, top->x0, top->y0, tc->cc.nlev - D->reduce);
                for (long i = 0; i < n; i++) da[i] = ia[i];
            } else
                rc = or_idwt97(da, tw, th, top->x0, top->y0, tc->cc.nlev - D->reduce);
        }
        free(ia);
    }
    if (!rc) {
        /* inverse MCT (G.2 RCT / G.3 ICT), TC/Decoder.cpp:163-180 selection */
        if (h->g.mct && C >= 3) {
            int rev = tcs[0].cc.xform == 1;
            for (int c = 0; c < 3; c++)
                if (h->dx[c] != 1 || h->dy[c] != 1) rc = OR_ERR_UNSUPPORTED; /* MCT needs equal components */
            if (tcs[1].cc.xform != tcs[0].cc.xform || tcs[2].cc.xform != tcs[0].cc.xform) rc = OR_ERR_UNSUPPORTED;
            const long n = (long)tws[0] * ths[0];
            for (long i = 0; i < n && !rc; i++) {
                double y0 = planes[0][i], y1 = planes[1][i], y2 = planes[2][i];
                if (rev) {
                    long Y0 = (long)y0, Y1 = (long)y1, Y2 = (long)y2;
                    long G = Y0 - floordiv(Y1 + Y2, 4);
                    planes[0][i] = (double)(Y2 + G);
                    planes[1][i] = (double)G;
                    planes[2][i] = (double)(Y1 + G);
                } else {
                    planes[0][i] = y0 + 1.402 * y2;
                    planes[1][i] = y0 - 0.34413 * y1 - 0.71414 * y2;
                    planes[2][i] = y0 + 1.772 * y1;
                }
            }
        }
        /* DC level shift, rounding and clamping (G.1.2) */
        for (int c = 0; c < C && !rc; c++) {
            int rev = tcs[c].cc.xform == 1;
            double shift = h->sgnd[c] ? 0.0 : ldexp(1.0, h->prec[c] - 1);
            long lo = h->sgnd[c] ? -(1L << (h->prec[c] - 1)) : 0;
            long hi = h->sgnd[c] ? (1L << (h->prec[c] - 1)) - 1 : (1L << h->prec[c]) - 1;
            const int tw = tws[c], th = ths[c];
            for (int y = 0; y < th; y++)
                for (int x = 0; x < tw; x++) {
                    double v = planes[c][(long)y * tw + x] + shift;
                    long iv = rev ? (long)v : (long)rint(v);
                    if (iv < lo) iv = lo;
                    if (iv > hi) iv = hi;
                    const res_t *top = &tcs[c].r[tcs[c].nres - 1 - D->reduce];
                    /* component c's own area (ceil(area / XRsiz), B-12, then the
                     * reduction), packed at pitch cw from c * W * H */
                    long s2 = 1L << D->reduce;
                    long cx0 = ceildiv(ceildiv(D->ix0, h->dx[c]), s2), cy0 = ceildiv(ceildiv(D->iy0, h->dy[c]), s2);
                    long cw = ceildiv(ceildiv(D->ix1, h->dx[c]), s2) - cx0;
                    long o = (long)c * D->W * D->H + (long)(top->y0 - cy0 + y) * cw + (top->x0 - cx0 + x);
                    D->out[o] = (int32_t)iv;
                    if (D->out_f) D->out_f[o] = v;
                }
        }
    }
    for (int c = 0; c < C; c++) free(planes[c]);
    free(planes);
    (void)t;
    return rc;
}

int or_decode(const uint8_t *cs, size_t len, int32_t *out, double *out_f) {
    return or_decode_reduced(cs, len, 0, out, out_f);
}

static int ceil_shift(int v, int k) { return (int)(((long)v + (1L << k) - 1) >> k); }

int or_decode_reduced(const uint8_t *cs, size_t len, int reduce, int32_t *out, double *out_f) {
    or_info info;
    int rc = or_read_info(cs, len, &info);
    if (rc) return rc;
    if (reduce < 0 || reduce > 32) return -OR_ERR_UNSUPPORTED;
    int X0 = ceil_shift(info.x0, reduce), Y0 = ceil_shift(info.y0, reduce);
    int X1 = ceil_shift(info.x1, reduce), Y1 = ceil_shift(info.y1, reduce);
    dec_ctx D = {out, out_f, X1 - X0, Y1 - Y0, X0, Y0, reduce, info.x0, info.y0, info.x1, info.y1};
    rc = walk_tiles(cs, len, decode_tile, &D);
    return rc ? -rc : OR_OK;
}
