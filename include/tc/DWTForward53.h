// Drop-in of TC/DWTForward53.h: reversible 5/3 forward DWT, integer lifting
// with floor (T.800 F.4.8.1), of one plane at origin (0, 0) from a raster into
// the Mallat layout (run(in, out, sizeX, sizeY, levels) of TC/DWTKernel.h:10-11).
#pragma once
#include "DWTKernel.h"

class DWTForward53 : public DWTKernel<int> {
  public:
    explicit DWTForward53(KernelInitInfoBase initInfo);
    virtual ~DWTForward53(void) {}
  protected:
    bool irreversible() const override { return false; }
    bool forward() const override { return true; }
};
