// Drop-in of TC/DWT.h:13-34: per-tile inverse DWT of every tile-component.
// iwt replaces each tile-component's coefficient plane by its sample raster.
#pragma once
#include "DWTReverse53.h"
#include "DWTReverse97.h"
#include "codestream_image_types.h"

class DWT {
  public:
    explicit DWT(KernelInitInfoBase initInfo);
    ~DWT(void);
    void iwt(type_tile* tile);                   // TC/DWT.cpp:98-114
  private:
    KernelInitInfoBase initInfo;
    DWTReverse53* r53;
    DWTReverse97* r97;
};
