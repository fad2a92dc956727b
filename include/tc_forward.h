/*
 * tc_forward.h -- C ABI of the GPU forward transform (encoder front end).
 *
 * The encoder-side counterpart of the decode path: the DC level shift and the
 * forward RCT/ICT (fdc_level_shift_kernel TC/preprocess_dc_level_shift.cl:16,
 * rct_kernel TC/preprocess_rct.cl:18, ict_kernel TC/preprocess_ict.cl:20,
 * driven by Preprocessor::dc_level_shifting / color_trans_gpu,
 * TC/Preprocessor.cpp:44-157) fused into the first level of the forward 5/3 or
 * 9/7 DWT (DWTForward53::dwt TC/DWTForward53.cpp:20-43, DWTForward97::dwt
 * TC/DWTForward97.cpp:20-43), then the remaining levels, one launch per level.
 *
 * Output: one plane of w x h elements per component (int32 for 5/3, float for
 * 9/7), each in the Mallat layout of a single tile at origin (0, 0): level n's
 * LL | HL over LH | HH in the top-left ceil(w/2^(n-1)) x ceil(h/2^(n-1))
 * region, pitch w -- the layout the reference's forward kernels leave after
 * their LL copy-back (TC/DWTKernel.cpp:127-154), with T.800 F.4 arithmetic.
 */
#ifndef TC_FORWARD_H
#define TC_FORWARD_H

#include "tc_device.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tc_forward_s *tc_forward;

tc_status tc_forward_create(tc_device dev, tc_forward *out);
tc_status tc_forward_destroy(tc_forward f);

/* Plans the transform of a w x h image of ncomp (1 or 3) interleaved components
 * with `bits` bits each (1..16: 8-bit samples up to 8 bits, 16-bit samples
 * above; unsigned, shifted by 2^(bits-1)), `levels` DWT levels (0..32),
 * reversible (5/3 + RCT) or irreversible (9/7 + ICT); mct: apply the colour
 * transform (ncomp 3 only).  Allocates the device image, planes and work
 * buffers and uploads the launch plan. */
tc_status tc_forward_prepare(tc_forward f, int w, int h, int ncomp, int bits, int levels, int irreversible, int mct);
/* Tile shape of the launches (tc_idwt_tile_shape); takes effect at the next prepare. */
tc_status tc_forward_set_shape(tc_forward f, int shape);
/* Copies the host image (h rows of w * ncomp samples) into the device image. */
tc_status tc_forward_upload(tc_forward f, const void *pixels);
/* Enqueues the transform of the device image on the forward stream. */
tc_status tc_forward_run(tc_forward f);
/* Same, timed with events on the forward stream: the fused first level alone
 * and the whole transform, in milliseconds (waits for the transform to end). */
tc_status tc_forward_run_timed(tc_forward f, float *first_ms, float *total_ms);
tc_status tc_forward_sync(tc_forward f);
/* Copies the planes (ncomp * w * h elements of 4 bytes) to the host. */
tc_status tc_forward_download(tc_forward f, void *planes);
tc_stream tc_forward_stream(tc_forward f);
/* Device addresses of the image and of plane 0 (planes follow at w * h elements). */
void *tc_forward_image_ptr(tc_forward f);
void *tc_forward_planes_ptr(tc_forward f);

/* Runs one of the DWT drop-in classes (include/tc/DWTForward53.h,
 * DWTForward97.h, DWTReverse53.h, DWTReverse97.h) over a host plane of w x h
 * int32 (5/3) or float (9/7) elements: DWTKernel::run(T *in, ...) then
 * mapOutputBufferToHost (TC/DWTKernel.cpp:45-78,180-191).  fwd: raster to
 * Mallat layout, else back.  For checking the classes. */
tc_status tc_dwt_plane(tc_device dev, int fwd, int irreversible, const void *in, void *out, int w, int h, int levels);

#ifdef __cplusplus
}
#endif
#endif
