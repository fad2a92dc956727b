// Drop-in of TC/DWTReverse97.h: irreversible 9/7, FP32 lifting.
#pragma once
#include "DWTKernel.h"

class DWTReverse97 : public DWTKernel<float> {
  public:
    explicit DWTReverse97(KernelInitInfoBase initInfo);
    virtual ~DWTReverse97(void) {}
  protected:
    bool irreversible() const override { return true; }
};
