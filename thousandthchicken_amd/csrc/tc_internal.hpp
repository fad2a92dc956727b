// Internal helpers shared by the host-side translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/tc_device.h"

namespace tc {
void set_error(const std::string& m);
tc_status fail(tc_status code, const std::string& m);
tc_status hip_fail(hipError_t e, const char* what);
hipStream_t stream_handle(tc_stream s);
int device_ordinal(tc_device d);
int device_cu_count(tc_device d);
// device address of pinned host memory from tc_host_alloc_pinned (kernels write it over the bus)
void* host_device_ptr(void* host);
}  // namespace tc
