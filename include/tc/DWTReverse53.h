// Drop-in of TC/DWTReverse53.h: reversible 5/3, integer lifting with floor.
#pragma once
#include "DWTKernel.h"

class DWTReverse53 : public DWTKernel<int> {
  public:
    explicit DWTReverse53(KernelInitInfoBase initInfo);
    virtual ~DWTReverse53(void) {}
  protected:
    bool irreversible() const override { return false; }
};
