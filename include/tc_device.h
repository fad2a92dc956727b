/*
 * tc_device.h -- C ABI of the MI355X (gfx950) device layer.
 *
 * Replaces the reference's OpenCL seam: DeviceKernel (TC/DeviceKernel.h:10-31,
 * TC/DeviceKernel.cpp:11-197), DeviceQueue (TC/DeviceQueue.h:6-16,
 * TC/DeviceQueue.cpp:14-35), the typedefs of TC/platform.h:9-11 and the raw
 * cl* calls the stage classes make (TC/CoefficientCoder.cpp:185-205,
 * TC/Quantizer.cpp:72-115, TC/DWT.cpp:71-84, TC/DWTKernel.cpp:54-74,135-147,183,
 * TC/Decoder.cpp:84,136,201-206).  Plain pointers and sizes only; no exception
 * crosses this boundary: every call returns a tc_status and sets a
 * thread-local message readable through tc_last_error().
 *
 * Kernels are compiled ahead of time for gfx950 and launched through one typed
 * entry point each (no runtime JIT, no generic setArg).
 */
#ifndef TC_DEVICE_H
#define TC_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* tDeviceRC / DeviceSuccess, TC/platform.h:10-11 */
typedef int32_t tc_status;
#define TC_SUCCESS 0
#define TC_ERR_INVALID_ARG 1
#define TC_ERR_TRUNCATED 2
#define TC_ERR_BAD_MARKER 3
#define TC_ERR_UNSUPPORTED 4
#define TC_ERR_GEOMETRY 5
#define TC_ERR_OUT_OF_MEMORY 6
#define TC_ERR_DEVICE 7
#define TC_ERR_CAPACITY 8
#define TC_ERR_BUSY 9
#define TC_ERR_NOT_FOUND 10

typedef struct tc_device_s *tc_device; /* ocl_args_d_t{context,device}, TC/ocl_util.h:17-25 */
typedef struct tc_stream_s *tc_stream; /* DeviceQueue / cl_command_queue */
typedef struct tc_buf_s *tc_buf;       /* tDeviceMem = cl_mem, TC/platform.h:9 */
typedef struct tc_event_s *tc_event;
typedef struct tc_graph_s *tc_graph;

typedef struct {
    char name[64];
    char arch[32];
    int32_t compute_units;
    int32_t wavefront_size;
    int32_t xcds;
    uint64_t lds_bytes_per_workgroup; /* CL_DEVICE_LOCAL_MEM_SIZE, TC/DeviceKernel.cpp:54-74 */
    uint64_t global_mem_bytes;
    uint32_t mem_base_align; /* requiredOpenCLAlignment, TC/basic.cpp:209 */
    int32_t clock_khz;
    int32_t memory_clock_khz;
    int32_t memory_bus_bits;
} tc_device_props;

const char *tc_last_error(void);
const char *tc_status_string(tc_status s);

/* Device (InitOpenCL, TC/ocl_util.cpp:425-465) */
tc_status tc_device_open(int ordinal, tc_device *out);
tc_status tc_device_close(tc_device dev);
tc_status tc_device_get_props(tc_device dev, tc_device_props *props);
tc_status tc_device_count(int *count);

/* Queue (DeviceQueue::finish / flush, TC/DeviceQueue.cpp:14-35) */
tc_status tc_stream_create(tc_device dev, tc_stream *out);
tc_status tc_stream_destroy(tc_stream s);
tc_status tc_stream_sync(tc_stream s);  /* clFinish */
tc_status tc_stream_flush(tc_stream s); /* clFlush: work is already submitted on HIP */
tc_device tc_stream_device(tc_stream s);

/* Buffers (clCreateBuffer / clReleaseMemObject / clEnqueueFillBuffer /
 * clEnqueueCopyBufferRect / clEnqueueMapBuffer) */
tc_status tc_buf_alloc(tc_device dev, size_t bytes, tc_buf *out);
tc_status tc_buf_free(tc_buf b);
void *tc_buf_device_ptr(tc_buf b);
tc_device tc_buf_device(tc_buf b);
size_t tc_buf_size(tc_buf b);
tc_status tc_buf_fill32(tc_stream s, tc_buf b, size_t byte_off, uint32_t value, size_t count);
tc_status tc_buf_copy_h2d(tc_stream s, tc_buf dst, size_t dst_off, const void *src, size_t bytes);
tc_status tc_buf_copy_d2h(tc_stream s, void *dst, tc_buf src, size_t src_off, size_t bytes);
tc_status tc_buf_copy_d2d(tc_stream s, tc_buf dst, size_t dst_off, tc_buf src, size_t src_off, size_t bytes);
/* origins and region in bytes / rows / slices, as clEnqueueCopyBufferRect */
tc_status tc_buf_copy_rect(tc_stream s, tc_buf dst, tc_buf src, const size_t src_origin[3], const size_t dst_origin[3],
                           const size_t region[3], size_t src_row_pitch, size_t src_slice_pitch,
                           size_t dst_row_pitch, size_t dst_slice_pitch);
/* blocking map for reading (clEnqueueMapBuffer with CL_TRUE, TC/Decoder.cpp:81-91) */
tc_status tc_buf_map_read(tc_stream s, tc_buf b, size_t off, size_t bytes, void **host_ptr);
tc_status tc_buf_unmap(tc_buf b, void *host_ptr);
tc_status tc_host_alloc_pinned(size_t bytes, void **out);
tc_status tc_host_free_pinned(void *p);

/* Events (profiling, eventExecutionTime TC/basic.cpp:269-281) */
tc_status tc_event_create(tc_event *out);
tc_status tc_event_destroy(tc_event e);
tc_status tc_event_record(tc_event e, tc_stream s);
tc_status tc_event_sync(tc_event e);
/* Work queued on s after this call waits for the work before e (cross-stream
 * dependency; also recorded as a graph edge while s is being captured). */
tc_status tc_stream_wait_event(tc_stream s, tc_event e);
/* TC_SUCCESS when the work before the event has finished, TC_ERR_BUSY if not */
tc_status tc_event_query(tc_event e);
tc_status tc_event_elapsed_ms(tc_event start, tc_event stop, float *ms);

/* Graph capture of the launches issued on a stream (batched form). */
tc_status tc_graph_begin(tc_stream s);
tc_status tc_graph_end(tc_stream s, tc_graph *out);
tc_status tc_graph_launch(tc_graph g, tc_stream s);
tc_status tc_graph_destroy(tc_graph g);

/* ---- typed kernel launches (DeviceKernel::enqueue, TC/DeviceKernel.cpp:180-197) ---- */

/* Tier-1 decode of n code-blocks (g_decode, TC/coefficient_coder.cl:778) with
 * one wavefront per code-block (latency-bound launches with few blocks): grid
 * work-groups loop over the n blocks in order.  descs: device array of
 * tc_cblk_desc (thousandthchicken_amd/csrc/tc_kernels.h), order: device array
 * of n indices into descs (packing order) or NULL, arena: codewords, coef:
 * zero-filled coefficient planes (int32 or float).  Codewords must
 * start at 4-byte aligned offsets; arena_bytes bounds every codeword read.
 * small_blocks: every block is at most 32 wide.  Writes every coefficient of each
 * decoded block (no zero-fill needed for them). */
tc_status tc_launch_t1_decode_wave(tc_stream s, int small_blocks, int grid, const void *descs, const uint32_t *order,
                                   int n, const uint8_t *arena, size_t arena_bytes, void *coef);

/* Same contract as tc_launch_t1_decode_wave (one wavefront per code-block),
 * with block state held column-major across the lanes and the scan restricted
 * to the columns holding candidates.  small_blocks: every block is at most
 * 32x32.  Blocks with more bit-planes than TC_T1_STRIPE_MAX_PLANES(small) are
 * skipped (their coefficients keep what the plane held); decode them with
 * tc_launch_t1_decode_wave. */
#define TC_T1_STRIPE_MAX_PLANES(small) ((small) ? 32 : 20)
tc_status tc_launch_t1_decode_stripe(tc_stream s, int small_blocks, int grid, const void *descs, const uint32_t *order,
                                     int n, const uint8_t *arena, size_t arena_bytes, void *coef);

/* Same work as tc_launch_t1_decode_wave with blocks_per_wave (64, 32 or 16)
 * code-blocks per wavefront decoded in lockstep, one MQ symbol per lane per
 * step (throughput form for large batches; fewer than 64 per wave makes more
 * waves for launches with few blocks per CU).  Blocks that share a wavefront
 * should have equal pass counts (the packing order sorts them so); any mix
 * decodes correctly.  max_passes: the largest npasses of the n blocks (at
 * most 255).  scratch: device memory of at least
 * tc_t1_lockstep_scratch_bytes(small_blocks, blocks_per_wave, n, max_passes) bytes,
 * 8-byte aligned, holding per-block significance / sign / pass state and the
 * refinement 1-bits of each pass (no initialisation needed).  Codewords may
 * start at any byte offset; arena is 16-byte aligned and arena_bytes (a
 * multiple of 16, at least 64) readable bytes cover every codeword.  This is
 * the only tier-1 launch that decodes the code-block style bits of each
 * descriptor's mode (RESET, CAUSAL, SEGSYM, BYPASS, TERMALL; T.800 Table A.19);
 * the others ignore them.  A BYPASS or TERMALL block's codeword starts 4-byte
 * aligned, after its codeword-segment lengths and their count (one 4-byte word
 * each, the count last). */
size_t tc_t1_lockstep_scratch_bytes(int small_blocks, int blocks_per_wave, int n, int max_passes);
tc_status tc_launch_t1_decode_lockstep(tc_stream s, int small_blocks, int blocks_per_wave, const void *descs,
                                       const uint32_t *order, int n, int max_passes, const uint8_t *arena,
                                       size_t arena_bytes, void *scratch, size_t scratch_bytes, void *coef);

/* Output tile shapes of the inverse DWT launches: shape 0..*count-1 covers
 * *w x *h output samples per work-group (the workgroup-window choice of
 * DWTKernel::enqueue, TC/DWTKernel.cpp:89-125, made by the autotuner here). */
tc_status tc_idwt_tile_shape(int shape, int *w, int *h, int *count);

/* One inverse DWT level over a job list (run, TC/dwt_r53.cl:183 / TC/dwt_r97.cl:229).
 * jobs: device array of tc_dwt_job built for tile shape `shape`; wg_job:
 * work-group -> job map of nwg entries. */
tc_status tc_launch_idwt_level(tc_stream s, int irreversible, int shape, const void *jobs, const uint32_t *wg_job,
                               int nwg, void *base);

/* Final level fused with inverse RCT/ICT, DC shift, rounding, clamping and packing
 * (tcr_kernel / tci_kernel / idc_level_shift_kernel, TC/preprocess_*_inverse.cl).
 * ncomp: 1 or 3; out_bytes: 1 (8-bit) or 2 (>8-bit). */
tc_status tc_launch_idwt_final(tc_stream s, int irreversible, int shape, int ncomp, int out_bytes, const void *jobs,
                               const uint32_t *wg_job, int nwg, const void *base, void *out);

/* One forward DWT level over a job list (encoder side; run, TC/dwt_f53.cl:184 /
 * TC/dwt_f97.cl:220): each tc_fdwt_job (thousandthchicken_amd/csrc/tc_kernels.h)
 * splits a resolution raster into its four subbands, int32 5/3 or float 9/7.
 * Jobs are built for tile shape `shape` (tc_idwt_tile_shape); wg_job maps each of
 * the nwg work-groups to its job. */
tc_status tc_launch_fdwt_level(tc_stream s, int irreversible, int shape, const void *jobs, const uint32_t *wg_job,
                               int nwg, void *base);

/* First forward level fused with the DC level shift and the forward RCT (5/3) or
 * ICT (9/7) (fdc_level_shift_kernel TC/preprocess_dc_level_shift.cl:16,
 * rct_kernel TC/preprocess_rct.cl:18, ict_kernel TC/preprocess_ict.cl:20): reads
 * the interleaved image (in_bytes 1: 8-bit, 2: 16-bit; nc 1 or 3) and writes the
 * level-1 subbands of every component per tc_fdwt_first_job. */
tc_status tc_launch_fdwt_first(tc_stream s, int irreversible, int shape, int nc, int in_bytes, const void *jobs,
                               const uint32_t *wg_job, int nwg, const void *img, void *base);

/* Standalone inverse colour transform / DC shift over int32 planes in place, for the
 * Preprocessor drop-in (color_decoder_lossless/lossy, idc_level_shifting,
 * TC/Preprocessor.cpp:102-157).  kind: 0 DC shift only, 1 inverse RCT, 2 inverse ICT,
 * 3 DC shift of float samples.  For kinds 2 and 3 the planes hold float samples on
 * entry and rounded int32 samples on return. */
tc_status tc_launch_colour(tc_stream s, int kind, void *c0, void *c1, void *c2, size_t count, int32_t shift,
                           int32_t minv, int32_t maxv);

/* fill of 32-bit words (clEnqueueFillBuffer), usable under graph capture */
tc_status tc_launch_fill32(tc_stream s, void *ptr, uint32_t value, uint64_t count);

#ifdef __cplusplus
}
#endif
#endif
