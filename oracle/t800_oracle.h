/*
 * t800_oracle.h -- CPU restatement of the JPEG 2000 Part-1 (ITU-T T.800) decode
 * path that the reference's OpenCL kernels are written to perform.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: tests/, the smoke test
 * and the CPU-baseline leg of the benchmark call it; the product path never
 * links, loads or calls it.
 *
 * What it follows (behaviour, not code):
 *   tier-2 subset .......... TC/codestream.c:233-258,336-470, TC/codestream_image.c:14-277,
 *                            TC/codestream_tag_tree_encode.c:97-132, TC/io_buffered_stream.c:87-132
 *   tier-1 (MQ + 3 passes) . TC/coefficient_coder.cl:36-60,105-252,330-586,778-819
 *                            (with the by-value no-op setters repaired, T.800 Annex C/D)
 *   dequantization ......... TC/Quantizer.cpp:37-65 (M_b, shift), T.800 E.1
 *   inverse 5/3, 9/7 ....... TC/dwt_r53.cl, TC/dwt_r97.cl, TC/dwt_transform_buffer.cl:102-556,
 *                            TC/dwt_io.cl:7-38 (floor instead of truncating '/', FP64 9/7)
 *   colour / DC shift ...... TC/preprocess_rct_inverse.cl:18-36, TC/preprocess_dc_level_shift_inverse.cl:16-19,
 *                            TC/Preprocessor.cpp:52-55, TC/Decoder.cpp:163-180 (ICT per T.800 G.3)
 */
#ifndef T800_ORACLE_H
#define T800_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_OK = 0,
    OR_ERR_TRUNCATED = 1,
    OR_ERR_MARKER = 2,
    OR_ERR_UNSUPPORTED = 3,
    OR_ERR_GEOMETRY = 4,
    OR_ERR_NOMEM = 5,
    OR_ERR_ARG = 6,
    OR_ERR_CORRUPT = 7, /* a codeword check failed (segmentation symbol) */
};

typedef struct {
    int32_t x0, y0, x1, y1;      /* image area on the reference grid */
    int32_t tx0, ty0, tw, th;    /* tile grid origin and nominal tile size */
    int32_t ntx, nty;            /* tiles across / down */
    int32_t ncomp;
    int32_t prec[16];            /* bit depth per component */
    int32_t sgnd[16];            /* 1 = signed samples */
    int32_t nlevels;             /* decomposition levels (component 0) */
    int32_t reversible;          /* 1 = 5/3, 0 = 9/7 (component 0) */
    int32_t mct;                 /* multiple component transform flag */
    int32_t cbw, cbh;            /* nominal code-block width / height (component 0) */
    int32_t nlayers;
    int32_t progression;
    int32_t qstyle;              /* quantization style of component 0 */
    int32_t guard_bits;
    int32_t dx[16], dy[16];      /* component sub-sampling XRsiz, YRsiz */
} or_info;

typedef struct {
    int32_t tile, comp, res, band;   /* band: 0 LL, 1 HL, 2 LH, 3 HH */
    int32_t x0, y0, x1, y1;          /* code-block rectangle, subband coordinates */
    int32_t npasses;                 /* coding passes included over all layers */
    int32_t zero_bp;                 /* missing most-significant bit-planes */
    int32_t mb;                      /* magnitude bit-planes of the subband, M_b */
    int32_t nbytes;                  /* codeword bytes over all layers */
    uint32_t crc;                    /* FNV-1a over the codeword bytes */
} or_cblk;

/* Reads SIZ/COD/QCD of the main header. */
int or_read_info(const uint8_t *cs, size_t len, or_info *info);

/* Tier-2 table: every code-block in codestream order of tiles, then component,
 * resolution, band, raster.  Returns the count (or -error); writes at most max. */
int or_list_cblks(const uint8_t *cs, size_t len, or_cblk *out, int max);

/* MQ symbols decoded per code-block (tier-2 order); returns the count. */
int or_t1_symbol_counts(const uint8_t *cs, size_t len, int32_t *nsym, int max);

/* Full decode.  out: planar int32 [ncomp][y1-y0][x1-x0] after colour transform,
 * DC shift, rounding and clamping.  out_f (optional): the same values before
 * rounding/clamping (FP64 for the 9/7 path, exact integers for 5/3). */
int or_decode(const uint8_t *cs, size_t len, int32_t *out, double *out_f);
/* Same with the `reduce` finest DWT levels not reconstructed: out holds the
 * image at that resolution, ceil(X / 2^reduce) wide (T.800 B.5), after the
 * colour transform and level shift. */
int or_decode_reduced(const uint8_t *cs, size_t len, int reduce, int32_t *out, double *out_f);

/* Tier-1 decode of one code-block (code-block style 0).  orient: 0 LL,
 * 1 HL, 2 LH, 3 HH.  numbps = M_b - zero_bp bit-planes.  Outputs per
 * coefficient (row-major w*h): magnitude bits, sign (0/1) and the lowest
 * bit-plane coded for it (-1 when it never became significant).  Returns the
 * number of MQ symbols decoded, or -error. */
/* Code-block style bits (T.800 Table A.19); the oracle decodes all six. */
#define OR_STYLE_BYPASS 1
#define OR_STYLE_RESET 2
#define OR_STYLE_TERMALL 4
#define OR_STYLE_CAUSAL 8
#define OR_STYLE_PTERM 16
#define OR_STYLE_SEGSYM 32
#define OR_STYLE_DECODABLE 63
/* or_t1_decode with code-block style bits */
int or_t1_decode_style(const uint8_t *data, int len, int w, int h, int orient, int numbps, int npasses, int style,
                       int32_t *mag, int8_t *sign, int8_t *lowplane);
int or_t1_decode(const uint8_t *data, int len, int w, int h, int orient,
                 int numbps, int npasses, int32_t *mag, int8_t *sign, int8_t *lowplane);
/* With the codeword split into nseg codeword segments of seglen[] bytes (in
 * order, summing to at most len): needed for BYPASS and TERMALL, whose passes
 * start new segments (or_seg_end(style, k): pass k ends one). */
int or_t1_decode_seg(const uint8_t *data, int len, const int *seglen, int nseg, int w, int h, int orient,
                     int numbps, int npasses, int style, int32_t *mag, int8_t *sign, int8_t *lowplane);
int or_seg_end(int style, int k);

/* Inverse 2-D DWT of one tile-component held in Mallat layout (subbands in
 * quadrants, LL top-left), in place.  (x0,y0) is the tile-component origin
 * on the canvas (it fixes the parity of every level), w,h its size. */
int or_idwt53(int32_t *a, int w, int h, int x0, int y0, int levels);
int or_idwt97(double *a, int w, int h, int x0, int y0, int levels);

/* Forward transforms (same layout conventions), used to build test inputs. */
int or_fdwt53(int32_t *a, int w, int h, int x0, int y0, int levels);
int or_fdwt97(double *a, int w, int h, int x0, int y0, int levels);

/* MQ arithmetic decoder over a symbol stream: decodes n symbols with the
 * given context labels (0..18, initial states per T.800 Table D.7). */
int or_mq_decode_symbols(const uint8_t *data, int len, const uint8_t *ctx, int n, uint8_t *out);

/* T.800 constant tables, for pinning against the reference's literals.
 * zc: [3][256] zero-coding labels indexed by the 8-neighbour significance
 * (bit order tl,t,tr,l,r,bl,b,br = bits 0..7), class 0 = LL/LH, 1 = HL, 2 = HH.
 * mq: [4][47] Qe, NMPS, NLPS, SWITCH. */
void or_tables(uint8_t *zc, uint32_t *mq);

/* Sign context (label 9..13) and XOR bit for horizontal/vertical sign
 * contributions h, v in {-1,0,1}. */
int or_sign_ctx(int h, int v, int *xorbit);

#ifdef __cplusplus
}
#endif
#endif
