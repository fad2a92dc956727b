/*
 * tc_decode.h -- C ABI of the JPEG 2000 decode path on MI355X.
 *
 * Decoder entry points replace Decoder::decode (TC/Decoder.h:14-33,
 * TC/Decoder.cpp:94-216) for callers that bind the path through an FFI; the
 * C++ drop-in classes (include/tc/Decoder.h and the stage classes) sit on the
 * same implementation.  Three forms are offered:
 *   - per image:  tc_decoder_decode (host codestream in, host image out);
 *   - resident:   tc_decoder_prepare once, tc_decoder_run any number of
 *                 times on the device-resident data, tc_decoder_fetch;
 *   - batched:    tc_decoder_prepare with n codestreams, one launch per
 *                 stage for the whole batch, optionally replayed from a
 *                 captured graph.
 * Output images are interleaved (pixel-major), 1 byte per sample when every
 * component has at most 8 bits, else 2 bytes (little-endian); signed
 * components are stored as two's complement of that width.
 */
#ifndef TC_DECODE_H
#define TC_DECODE_H

#include "tc_device.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tc_decoder_s *tc_decoder;

typedef struct {
    int32_t x0, y0, width, height;
    int32_t ncomp;
    int32_t max_prec;
    int32_t bytes_per_sample; /* of the decoded output: 1 or 2 */
    int32_t reversible;       /* component 0: 1 = 5/3, 0 = 9/7 */
    int32_t nlevels, ntx, nty, cbw, cbh, nlayers, mct, progression;
    int32_t prec[4], sgnd[4]; /* first four components */
    /* 1 when a component is sub-sampled (XRsiz or YRsiz > 1): the output then
     * holds each component as its own plane of comp_width x comp_height
     * samples, one plane after the other (at most four components); else the
     * components are interleaved per pixel */
    int32_t planar;
    int32_t comp_width[4], comp_height[4];
    /* 1 when the input was a JP2 file (T.800 Annex I), whose image header the
     * decoder checked against the codestream; colourspace is its first usable
     * colour specification: the enumerated colour space (16 sRGB, 17
     * greyscale, 18 sYCC), -1 for an ICC profile, 0 when there is none (and
     * for a raw codestream).  The decoder returns the decoded components as
     * they are: no colour-space conversion is applied. */
    int32_t jp2, colourspace;
} tc_image_info;

/* tier-2 view of one code-block, for checking the parser */
typedef struct {
    int32_t tile, comp, res, band; /* band: 0 LL, 1 HL, 2 LH, 3 HH */
    int32_t x0, y0, x1, y1;
    int32_t npasses, zero_bp, mb, nbytes;
    uint32_t crc; /* FNV-1a of the codeword bytes */
} tc_cblk_info;

typedef struct {
    float fill_ms, t1_ms, dwt_ms, final_ms, total_ms; /* of the last timed run */
    int64_t cblks, codeword_bytes, samples, pixels;
    int64_t dwt_bytes, final_bytes; /* algorithmic HBM bytes of the DWT levels / final level */
    int32_t bpw_small, bpw_large;   /* blocks per wave used (0: class absent) */
    int32_t images;
    int64_t device_bytes;
    int64_t t1_samples; /* coefficients of the code-blocks tier-1 decodes (blocks with coding passes) */
    /* The tier-1 launches of the plan (at most four: per block class, the
     * one-wave-per-block launch of its heaviest blocks, then its main launch),
     * each with its block class (0: blocks up to 32x32, 1: larger), packing
     * (blocks per wave as in tc_decoder_set_blocks_per_wave), block count,
     * algorithmic bytes (codeword bytes + 4 B per coefficient of its blocks)
     * and its time on the stream it ran on in the last timed run. */
    int32_t t1_launches;
    int32_t t1_launch_class[4], t1_launch_bpw[4];
    int64_t t1_launch_blocks[4], t1_launch_bytes[4];
    float t1_launch_ms[4];
    int32_t heavy_small, heavy_large; /* heavy-block thresholds the plan applied (0: none) */
} tc_decoder_stats;

/* Main-header information (SIZ/COD/QCD) of a raw codestream or JP2 file.
 * A JP2 file (signature box first) is read as the reference's box reader
 * does (TC/boxes.c:142-307): the File Type box must name 'jp2 ', the JP2
 * Header box must start with an Image Header box and precede the codestream
 * box.  Beyond the reference, the image header (size, component count, bit
 * depths, compression type 7) must agree with SIZ (else TC_ERR_GEOMETRY), and
 * files with palette (pclr) or component-mapping (cmap) boxes fail with
 * TC_ERR_UNSUPPORTED.  Every decode entry point applies the same checks. */
tc_status tc_image_probe(const uint8_t *cs, size_t len, tc_image_info *info);

/* Same for the image decoded with its `reduce` finest DWT levels discarded
 * (tc_decoder_set_reduce): x0, y0, width, height and nlevels are those of the
 * resolution `reduce` levels below full, ceil(x / 2^reduce).  Fails when the
 * first tile-component has fewer than `reduce` levels. */
tc_status tc_image_probe_reduced(const uint8_t *cs, size_t len, int reduce, tc_image_info *info);

/* Parses the whole codestream and lists its code-blocks (tile, component,
 * resolution, band, raster order).  *count receives the total; at most max are
 * written to out (which may be NULL). */
tc_status tc_image_list_cblks(const uint8_t *cs, size_t len, tc_cblk_info *out, int max, int *count);

/* Codeword bytes of code-block `index` (order of tc_image_list_cblks), its
 * segments concatenated; *nbytes receives the size (out may be NULL). */
tc_status tc_image_cblk_data(const uint8_t *cs, size_t len, int index, uint8_t *out, size_t cap, size_t *nbytes);

tc_status tc_decoder_create(tc_device dev, tc_decoder *out);
tc_status tc_decoder_destroy(tc_decoder dec);

/* Code-block-to-wave packing for blocks up to 32x32 (small_blocks = 1) or up to
 * 64x64 (small_blocks = 0): 0 automatic, 1 one wavefront per block
 * (tc_launch_t1_decode_stripe / _wave), TC_BPW_LOCKSTEP 64 blocks per wavefront
 * in lockstep, TC_BPW_LOCKSTEP32 32 and TC_BPW_LOCKSTEP16 16
 * (tc_launch_t1_decode_lockstep). */
#define TC_BPW_LOCKSTEP (-64)
#define TC_BPW_LOCKSTEP32 (-32)
#define TC_BPW_LOCKSTEP16 (-16)
tc_status tc_decoder_set_blocks_per_wave(tc_decoder dec, int small_blocks, int bpw);

/* Resolution-reduced decode (applies from the next prepare or decode): the
 * `reduce` finest DWT levels of every tile-component are not reconstructed and
 * their code-blocks are not decoded; the output is the image at that
 * resolution (size from tc_image_probe_reduced), after the colour transform
 * and level shift.  0 decodes at full resolution.  Decoding fails for an image
 * with a tile-component of fewer levels. */
tc_status tc_decoder_set_reduce(tc_decoder dec, int reduce);

/* Region decode (applies from the next prepare or decode; no counterpart in
 * the reference, the "decode area" of JPEG 2000 decoders): only the w x h
 * window at (x, y) from the image origin is output, clipped to the image (of
 * the reduced image when a reduce is set), and only the code-blocks that can
 * influence it are decoded.  The window's samples equal those of the full
 * decode.  w or h 0 decodes the whole image.  The output of every image of a
 * batch is its window, row-major and interleaved like the full image. */
tc_status tc_decoder_set_region(tc_decoder dec, int x, int y, int w, int h);

/* Zero-copy host output (applies from the next prepare or decode; no
 * counterpart in the reference, whose output is mapped back with
 * clEnqueueMapBuffer, TC/Decoder.cpp:81-91): enable 1 makes the decoder keep
 * its output images in pinned host memory that the final-level kernel writes
 * directly over the bus.  tc_decoder_output then returns host addresses,
 * readable once tc_decoder_sync returns; tc_decoder_fetch copies from them. */
tc_status tc_decoder_set_host_output(tc_decoder dec, int enable);

/* Kernel for blocks decoded one per wavefront: 0 = stripe-column form
 * (tc_launch_t1_decode_stripe, default), 1 = row-mask form
 * (tc_launch_t1_decode_wave). */
tc_status tc_decoder_set_t1_kernel(tc_decoder dec, int kind);

/* Launch configuration of a decoder (applies from the next prepare). */
typedef struct {
    int32_t bpw_small, bpw_large; /* per block class: 0 auto, 1, TC_BPW_LOCKSTEP, _LOCKSTEP32 or _LOCKSTEP16 */
    int32_t t1_kernel;            /* one-wave-per-block kernel: 0 stripe, 1 row-mask */
    int32_t dwt_shape;            /* inverse-DWT tile shape, see tc_idwt_tile_shape */
    float t1_ms, dwt_ms;          /* stage times the autotuner measured for this choice */
    /* per block class packed in lockstep: blocks of more estimated MQ symbols
     * than this decode one wavefront per block in a launch of their own, beside
     * the lockstep launch (0: no split, -1: the default) */
    int32_t heavy_small, heavy_large;
    int32_t heavy_waves;          /* waves per compute unit of that launch (0: 4) */
} tc_tuning;

tc_status tc_decoder_get_tuning(tc_decoder dec, tc_tuning *out);
tc_status tc_decoder_set_tuning(tc_decoder dec, const tc_tuning *in);

/* Autotuner: times the prepared batch under candidate code-block-to-wave
 * packings (per block class, in up to two rounds, since the two classes'
 * launches share the tier-1 stage), heavy-block thresholds and inverse-DWT
 * tile shapes, keeps the fastest (a candidate must beat the incumbent by more
 * than 1 %, the run-to-run spread), and re-prepares with it.  *best (may be
 * NULL) receives the choice. */
tc_status tc_decoder_autotune(tc_decoder dec, int reps, tc_tuning *best);

/* Per-image decode: parse, upload, decode, download (blocking). */
tc_status tc_decoder_decode(tc_decoder dec, const uint8_t *cs, size_t len, void *out, size_t out_cap,
                            tc_image_info *info);

/* Resident / batched form. */
tc_status tc_decoder_prepare(tc_decoder dec, int n, const uint8_t *const *cs, const size_t *lens);
tc_status tc_decoder_run(tc_decoder dec, int timed);
tc_status tc_decoder_sync(tc_decoder dec);
tc_status tc_decoder_fetch(tc_decoder dec, int i, void *out, size_t out_cap);
tc_status tc_decoder_output(tc_decoder dec, int i, const void **device_ptr, size_t *bytes);
tc_status tc_decoder_capture_graph(tc_decoder dec);
tc_status tc_decoder_launch_graph(tc_decoder dec);
tc_status tc_decoder_get_stats(tc_decoder dec, tc_decoder_stats *stats);
tc_stream tc_decoder_stream(tc_decoder dec);

/* Per-image decode that allocates no memory on the device or pinned host
 * memory.  No counterpart in the reference: it is for callers that cannot
 * allocate (serving loops that must not call hipMalloc, harnesses that forbid
 * it).  Everything the kernels need beyond the codestream and the output --
 * the uploaded codewords and descriptors, the coefficient planes, the
 * inverse-DWT work buffers and the tier-1 state -- lives in one static device
 * array of TC_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from its
 * first use on that GPU until the process ends.  The decoded image goes to
 * out_device, device memory of out_cap bytes that the caller owns, in
 * tc_decoder_decode's layout (info, when not NULL, receives the image's
 * description).  The call is blocking: the stream has finished the decode when
 * it returns.  An image whose working memory exceeds the array fails with
 * TC_ERR_CAPACITY and launches nothing.  Every call on one GPU shares that
 * GPU's array and leaves its contents behind (each call clears what must
 * start at zero); calls that overlapped on one GPU would race over it, so the
 * library serialises them per GPU with a lock held for the whole call.
 * Host-side parsing and planning use ordinary heap memory. */
#define TC_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
tc_status tc_decode_noalloc(tc_stream stream, const uint8_t *cs, size_t len, void *out_device, size_t out_cap,
                            tc_image_info *info);

/* Decodes by running the drop-in stage classes one after the other, as
 * TC/Decoder.cpp:153-180 does (CoefficientCoder::decode_tile,
 * Quantizer::dequantize_tile, DWT::iwt per tile, then Preprocessor), with each
 * tile-component mapped to the host and packed into the same interleaved layout
 * as tc_decoder_decode.  Slower than the fused path; for checking the stages. */
tc_status tc_decode_staged(tc_device dev, const uint8_t *cs, size_t len, void *out, size_t out_cap);

#ifdef __cplusplus
}
#endif
#endif
