// Drop-in of TC/Preprocessor.h:19-79: inverse component transforms and DC
// level shift over the tile-component rasters, in place (int32 results).
#pragma once
#include "DeviceKernel.h"
#include "codestream_image_types.h"

typedef enum { RCT, TCR, ICT, TCI } color_trans_type;

class Preprocessor {
  public:
    explicit Preprocessor(KernelInitInfoBase initInfo) : initInfo(initInfo) {}
    ~Preprocessor(void) {}
    void idc_level_shifting(type_image* img);    // TC/Preprocessor.cpp:154-157
    int color_decoder_lossy(type_image* img);    // TC/Preprocessor.cpp:102-108
    int color_decoder_lossless(type_image* img); // TC/Preprocessor.cpp:109-115
  private:
    int color_trans_gpu(type_image* img, color_trans_type type);  // TC/Preprocessor.cpp:44-93
    KernelInitInfoBase initInfo;
};
