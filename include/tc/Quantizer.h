// Drop-in of TC/Quantizer.h:10-28.  Inverse quantization (T.800 E.1: the
// reversible shift and Delta_b * (q + r)) happens in the tier-1 kernel's store,
// so dequantize_tile only checks that the tile went through decode_tile.
#pragma once
#include "DeviceKernel.h"
#include "codestream_image_types.h"

class Quantizer {
  public:
    explicit Quantizer(KernelInitInfoBase initInfo) : initInfo(initInfo) {}
    virtual ~Quantizer(void) {}
    void dequantize_tile(type_tile* tile);       // TC/Quantizer.cpp:181-216
  private:
    KernelInitInfoBase initInfo;
};
