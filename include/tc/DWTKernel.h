// Drop-in of TC/DWTKernel.h:6-31: multi-level inverse DWT of one component
// plane held in Mallat layout (origin 0,0) into a raster, or (the forward
// classes) a raster into that layout.  One launch per level on the typed
// tc_launch_idwt_level / tc_launch_fdwt_level entries; no LL copy-back.
#pragma once
#include <vector>

#include "DeviceKernel.h"

template <typename T>
class DWTKernel : public DeviceKernel {
  public:
    DWTKernel(int waveletImpulseDiameter, KernelInitInfo initInfo);
    virtual ~DWTKernel(void);
    tDeviceRC run(tDeviceMem in, tDeviceMem out, int sizeX, int sizeY, int levels);  // TC/DWTKernel.cpp:158-178
    tDeviceRC run(T* in, int sizeX, int sizeY, int levels);                          // TC/DWTKernel.cpp:45-78
    T* mapOutputBufferToHost();                                                       // TC/DWTKernel.cpp:180-191
    // general form: tile-component origin (x0, y0) fixes the parity of every level
    // (inverse classes; the forward classes take origin (0, 0))
    tDeviceRC run_at(tDeviceMem in, tDeviceMem out, int x0, int y0, int sizeX, int sizeY, int levels);

  protected:
    virtual bool irreversible() const = 0;
    virtual bool forward() const { return false; }
    tDeviceRC run_forward(tDeviceMem in, tDeviceMem out, int sizeX, int sizeY, int levels);
    tDeviceMem srcMem = nullptr;
    tDeviceMem dstMem = nullptr;
    int dimX = 0, dimY = 0;
    bool ownsMemory = false;
    int waveletImpulseDiameter;
    tc_buf work_ = nullptr, jobs_ = nullptr;
    std::vector<T> host_out_;
};
