// Elementwise kernels over 32-bit planes.
//
// Standalone inverse colour transforms and DC level shift over planes in
// place, for callers that run the reference's Preprocessor stage on its own
// (TC/Preprocessor.cpp:44-157; kernels TC/preprocess_rct_inverse.cl:18-36,
// TC/preprocess_ict_inverse.cl:22-41 with T.800 G.3 constants,
// TC/preprocess_dc_level_shift_inverse.cl:16-19).  The decode pipeline itself
// fuses this work into the final inverse-DWT level.
//
// The 32-bit fill clears the coefficient planes before tier-1 ORs its bits
// into them (the reference's clEnqueueFillBuffer, TC/CoefficientCoder.cpp:196):
// 16-byte stores when the start is 16-byte aligned, so one launch writes at
// close to the HBM store rate.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tc {
namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__global__ void colour_kernel(int kind, int32_t* __restrict__ c0, int32_t* __restrict__ c1,
                              int32_t* __restrict__ c2, size_t n, int shift, int lo, int hi) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (kind == 0) {
            c0[i] = clampi(c0[i] + shift, lo, hi);
        } else if (kind == 1) {
            int y0 = c0[i], y1 = c1[i], y2 = c2[i];
            int g = y0 - ((y1 + y2) >> 2);
            c0[i] = clampi(y2 + g + shift, lo, hi);
            c1[i] = clampi(g + shift, lo, hi);
            c2[i] = clampi(y1 + g + shift, lo, hi);
        } else if (kind == 3) {
            c0[i] = clampi((int)rintf(__int_as_float(c0[i])) + shift, lo, hi);
        } else {
            float y0 = __int_as_float(c0[i]), y1 = __int_as_float(c1[i]), y2 = __int_as_float(c2[i]);
            float r = fmaf(1.402f, y2, y0);
            float g = fmaf(-0.71414f, y2, fmaf(-0.34413f, y1, y0));
            float b = fmaf(1.772f, y1, y0);
            c0[i] = clampi((int)rintf(r) + shift, lo, hi);
            c1[i] = clampi((int)rintf(g) + shift, lo, hi);
            c2[i] = clampi((int)rintf(b) + shift, lo, hi);
        }
    }
}

__global__ void fill32_kernel(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void fill128_kernel(uint4* __restrict__ p, uint32_t v, uint64_t n4) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint4 q = make_uint4(v, v, v, v);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) p[i] = q;
}

}  // namespace

hipError_t launch_fill32(hipStream_t s, uint32_t* p, uint32_t v, uint64_t n) {
    if (!n) return hipSuccess;
    uint64_t head = 0;
    if (((uintptr_t)p & 15) == 0) {
        // body of whole 16-byte words, then the (< 4 element) tail below
        const uint64_t n4 = n / 4;
        if (n4) {
            uint64_t blocks = (n4 + 255) / 256;
            if (blocks > 8192) blocks = 8192;
            hipLaunchKernelGGL(fill128_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (uint4*)p, v, n4);
        }
        head = n4 * 4;
    }
    const uint64_t rest = n - head;
    if (rest) {
        uint64_t blocks = (rest + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(fill32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p + head, v, rest);
    }
    return hipGetLastError();
}

hipError_t launch_colour(hipStream_t s, int kind, void* c0, void* c1, void* c2, size_t n, int shift, int lo,
                         int hi) {
    if (!n) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(colour_kernel, dim3((unsigned)blocks), dim3(256), 0, s, kind, (int32_t*)c0, (int32_t*)c1,
                       (int32_t*)c2, n, shift, lo, hi);
    return hipGetLastError();
}

}  // namespace tc
