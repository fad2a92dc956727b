// Drop-in of TC/Decoder.h:14-33.  decode() runs the fused device pipeline
// (tier-1 with dequantization -> inverse DWT levels -> final level with the
// colour transform, DC shift and packing); decode_staged() runs the stage
// classes one after the other in the order of TC/Decoder.cpp:153-180.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "CoefficientCoder.h"
#include "DWT.h"
#include "Preprocessor.h"
#include "Quantizer.h"

class Decoder {
  public:
    explicit Decoder(ocl_args_d_t* ocl);
    ~Decoder(void);
    int decode(std::string fileName);                          // TC/Decoder.cpp:94-216
    int decode_memory(const uint8_t* data, size_t len);        // same, from memory
    int decode_staged(const uint8_t* data, size_t len);        // stage classes in reference order
    // interleaved output of the last decode (1 or 2 bytes per sample)
    const std::vector<uint8_t>& output() const { return out_; }
    int width() const { return width_; }
    int height() const { return height_; }
    int components() const { return ncomp_; }
    int bytes_per_sample() const { return bps_; }
    float last_decode_ms() const { return last_ms_; }

  private:
    tDeviceRC mapComponentToHost(type_tile_comp* tile_comp);   // TC/Decoder.cpp:81-91
    ocl_args_d_t* _ocl;
    CoefficientCoder* coder;
    Quantizer* quantizer;
    DWT* dwt;
    Preprocessor* preprocessor;
    void* fused_;  // tc_decoder of the fused path
    std::vector<uint8_t> out_;
    int width_ = 0, height_ = 0, ncomp_ = 0, bps_ = 1;
    float last_ms_ = 0.f;
};
