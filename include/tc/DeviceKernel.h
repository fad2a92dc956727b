// Drop-in of TC/DeviceKernel.h:10-31.  The reference object compiled one
// OpenCL kernel from source at construction (TC/DeviceKernel.cpp:68-150) and
// launched it with a generic enqueue(dim, gws, lws); here every kernel is
// precompiled for gfx950 and launched through its typed tc_launch_* entry, so
// this base keeps only the queue semantics (finish) that callers rely on.
#pragma once
#include "DeviceQueue.h"

class DeviceKernel {
  public:
    explicit DeviceKernel(KernelInitInfo initInfo) : queue_(initInfo.cmd_queue), deviceQueue(initInfo), name_(initInfo.kernelName) {}
    virtual ~DeviceKernel() {}
    tDeviceRC finish() { return deviceQueue.finish(); }
    const std::string& kernelName() const { return name_; }
  protected:
    tc_stream queue_;
    DeviceQueue deviceQueue;
    std::string name_;
};
