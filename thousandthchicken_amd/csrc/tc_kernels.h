// Device-side work descriptors shared by the HIP kernels and the host planner.
// Plain structs, no HIP types: the host fills arrays of them and uploads them.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// One code-block for tier-1 decoding (the reference's per-block descriptor is
// CodeBlockAdditionalInfo, TC/coefficientcoder_common.h:5-24; here with 16-bit
// dimensions, the pass count, the codeword location and the dequantization
// step folded in).  32 bytes.
typedef struct tc_cblk_desc {
    uint32_t data_off;   // byte offset of the codeword in the codeword arena
    uint32_t data_len;   // codeword bytes
    uint64_t out_off;    // element offset of coefficient (0,0) of the block
    uint32_t out_pitch;  // elements per row of the coefficient plane
    uint16_t w, h;       // block size (<= 64 x 64)
    uint8_t numbps;      // bit-planes to decode: M_b minus missing MSBs
    uint8_t npasses;     // coding passes present (<= 3*numbps-2)
    uint8_t cls;         // context class: 0 = LL/LH, 1 = HL, 2 = HH
    uint8_t mode;        // TC_MODE_* bits: irreversible (float out, else int32) and the code-block style
    float delta;         // quantization step Delta_b (irreversible)
} tc_cblk_desc;

// tc_cblk_desc::mode bits.  The style bits (T.800 Table A.19) are decoded by
// the lockstep tier-1 kernel only; the planner sends every class holding such
// a block there.
#define TC_MODE_IRREVERSIBLE 1
#define TC_MODE_RESET 2    // every context reset after each pass (D.6)
#define TC_MODE_CAUSAL 4   // stripe-causal contexts (D.7)
#define TC_MODE_SEGSYM 8   // segmentation symbol after each cleanup pass (D.5)
// BYPASS: SPP/MRP passes from the fifth bit-plane on are raw; TERMALL: every
// pass ends a codeword segment (D.6, Table D.9).  With either, the block's
// codeword is its segments in order, and the 4-byte words before data_off
// (which is then 4-byte aligned) hold the segment lengths and their count:
// word data_off/4 - 1 = nseg, words data_off/4 - 1 - nseg + s = length of s.
#define TC_MODE_BYPASS 16
#define TC_MODE_TERMALL 32
#define TC_MODE_STYLE (TC_MODE_RESET | TC_MODE_CAUSAL | TC_MODE_SEGSYM | TC_MODE_BYPASS | TC_MODE_TERMALL)

// One level of a 2-D inverse DWT over one tile-component: rebuilds the
// resolution rectangle [u0,u1) x [v0,v1) from its LL (previous level's output
// or the coefficient plane) and the HL/LH/HH quadrants of the coefficient plane.
// Offsets are in elements from the kernel's base pointer.
typedef struct tc_dwt_job {
    uint64_t ll_off, hl_off, lh_off, hh_off, out_off;
    uint32_t ll_pitch, band_pitch, out_pitch;
    int32_t u0, u1, v0, v1;
    int32_t tiles_x;     // work-group tiles across
    int32_t wg_base;     // first work-group index of this job in the launch
    int32_t pad;
} tc_dwt_job;

// Final level fused with the inverse component transform, DC level shift,
// rounding, clamping and packing into the interleaved output image.
// Components 0..nc-1 of the job share the rectangle.  levels == 0 means the
// tile-component has no DWT (the LL band is the image).
typedef struct tc_final_job {
    uint64_t ll_off[3], hl_off[3], lh_off[3], hh_off[3];
    uint32_t ll_pitch[3], band_pitch[3];
    int32_t shift[3], minv[3], maxv[3];
    int32_t u0, u1, v0, v1;
    uint64_t out_off;       // output element index of (u0,v0), first component
    uint32_t out_pitch;     // output elements per image row
    uint32_t out_cstride;   // output elements between adjacent pixels
    int32_t tiles_x, wg_base;
    int32_t levels;         // 0: copy through
    int32_t mct;            // 1: apply RCT (5/3) / ICT (9/7) to components 0..2
    // output window inside [u0,u1) x [v0,v1): work-group tiles start at
    // (wx0, wy0), only samples inside are written, and out_off addresses
    // (wx0, wy0); the whole rectangle unless a region decode narrows it
    int32_t wx0, wx1, wy0, wy1;
} tc_final_job;

// One level of a 2-D forward DWT over one tile-component (encoder side): splits
// the resolution rectangle [u0,u1) x [v0,v1), a raster at in_off with in_pitch
// elements per row, into its LL/HL/LH/HH subbands.  Offsets are in elements
// from the kernel's base pointer.
typedef struct tc_fdwt_job {
    uint64_t in_off, ll_off, hl_off, lh_off, hh_off;
    uint32_t in_pitch, ll_pitch, band_pitch;
    int32_t u0, u1, v0, v1;
    int32_t tiles_x;     // work-group tiles across
    int32_t wg_base;     // first work-group index of this job in the launch
    int32_t pad;
} tc_fdwt_job;

// First forward level fused with the DC level shift and the forward RCT (5/3)
// or ICT (9/7): reads the interleaved 8/16-bit image and writes the level-1
// subbands of components 0..nc-1.  levels == 0 writes the shifted (and
// colour-transformed) samples to ll_off instead.
typedef struct tc_fdwt_first_job {
    uint64_t ll_off[3], hl_off[3], lh_off[3], hh_off[3];
    uint32_t ll_pitch[3], band_pitch[3];
    int32_t shift[3];
    int32_t u0, u1, v0, v1;
    uint64_t in_off;        // image element index of (u0,v0), first component
    uint32_t in_pitch;      // image elements per row
    uint32_t in_cstride;    // image elements between adjacent pixels
    int32_t tiles_x, wg_base;
    int32_t levels;         // 0: no DWT, the samples go to ll_off
    int32_t mct;            // 1: forward RCT (5/3) / ICT (9/7) on components 0..2
} tc_fdwt_first_job;

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
static_assert(sizeof(tc_cblk_desc) == 32, "tc_cblk_desc is read as 8 dwords by the kernels");
#endif
