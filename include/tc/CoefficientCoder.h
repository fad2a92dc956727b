// Drop-in of TC/CoefficientCoder.h:20-61: tier-1 decoding of a tile's code-blocks.
// decode_tile writes each tile-component's coefficient plane (Mallat layout,
// int32 for 5/3, float for 9/7) with the dequantization already applied, so the
// separate pass of Quantizer::dequantize_tile has nothing left to do.
#pragma once
#include <vector>

#include "DeviceKernel.h"
#include "codestream_image_types.h"

#define LL_LH_SUBBAND 0
#define HL_SUBBAND 1
#define HH_SUBBAND 2

class CoefficientCoder : public DeviceKernel {
  public:
    explicit CoefficientCoder(KernelInitInfoBase initInfo);
    virtual ~CoefficientCoder(void);
    void decode_tile(type_tile* tile);           // TC/CoefficientCoder.cpp:62-88
    // packing of the tile's blocks (0 automatic, 1 one wave per block, or
    // TC_BPW_LOCKSTEP / _LOCKSTEP32 / _LOCKSTEP16 from tc_decode.h)
    void set_blocks_per_wave(int small_blocks, int bpw) { (small_blocks ? bpw32_ : bpw64_) = bpw; }
    float last_kernel_ms() const { return last_ms_; }

  private:
    tc_device device() const;
    tc_buf d_upload_ = nullptr;
    tc_buf d_scratch_ = nullptr;  // lockstep tier-1 state (code-block styles)
    size_t upload_cap_ = 0;
    int bpw32_ = 0, bpw64_ = 0;
    float last_ms_ = 0.f;
};
