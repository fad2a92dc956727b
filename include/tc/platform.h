// Drop-in replacement of TC/platform.h:1-52 and the parts of TC/ocl_util.h:17-25
// the stage classes use: the device types become the C-ABI handles of
// tc_device.h, so the classes keep their signatures.
#pragma once
#include <string>

#include "../tc_device.h"

typedef tc_buf tDeviceMem;   // cl_mem
typedef tc_status tDeviceRC; // cl_int
#define DeviceSuccess TC_SUCCESS

// InitOpenCL's result (TC/ocl_util.h:17-25): the device and its in-order queue.
struct ocl_args_d_t {
    tc_device device = nullptr;
    tc_stream commandQueue = nullptr;
};

struct QueueInfo {
    explicit QueueInfo(tc_stream queue) : cmd_queue(queue) {}
    tc_stream cmd_queue;
};

// buildOptions is kept for interface parity; kernels are compiled ahead of time.
struct KernelInitInfoBase : QueueInfo {
    KernelInitInfoBase(tc_stream queue, std::string bldOptions) : QueueInfo(queue), buildOptions(bldOptions) {}
    KernelInitInfoBase(const KernelInitInfoBase& o) : QueueInfo(o.cmd_queue), buildOptions(o.buildOptions) {}
    std::string buildOptions;
};

struct KernelInitInfo : KernelInitInfoBase {
    KernelInitInfo(tc_stream queue, std::string progName, std::string knlName, std::string bldOptions)
        : KernelInitInfoBase(queue, bldOptions), programName(progName), kernelName(knlName) {}
    KernelInitInfo(KernelInitInfoBase base, std::string progName, std::string knlName)
        : KernelInitInfoBase(base), programName(progName), kernelName(knlName) {}
    std::string programName;
    std::string kernelName;
};

// Error convention of TC/basic.h:26-33,95-104: failures of calls the reference
// wrapped in SAMPLE_CHECK_ERRORS throw Error.
#include <stdexcept>
struct Error : public std::runtime_error {
    explicit Error(const std::string& msg) : std::runtime_error(msg) {}
};
