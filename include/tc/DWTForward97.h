// Drop-in of TC/DWTForward97.h: irreversible 9/7 forward DWT in FP32 (T.800
// F.4.8.2, K scaling on both axes), of one plane at origin (0, 0) from a raster
// into the Mallat layout.
#pragma once
#include "DWTKernel.h"

class DWTForward97 : public DWTKernel<float> {
  public:
    explicit DWTForward97(KernelInitInfoBase initInfo);
    virtual ~DWTForward97(void) {}
  protected:
    bool irreversible() const override { return true; }
    bool forward() const override { return true; }
};
