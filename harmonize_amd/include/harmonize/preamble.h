// gfx950-only platform layer. The reference selects CUDA or HIP here and, on
// HIP, replaces the warp intrinsics with constants (preamble/adapt.h:93-95);
// this layer has a single target and keeps only the GPUrt* names that
// specifications and host mains call (adapt.h:99-137), as thin forwards to HIP.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "portable.h"

namespace adapt {

size_t const WARP_SIZE = 64;

using GPUrtError_t = hipError_t;
using GPUrtEvent_t = hipEvent_t;
using GPUrtStream_t = hipStream_t;

auto const GPUrtSuccess = hipSuccess;
auto const GPUrtMemcpyDeviceToDevice = hipMemcpyDeviceToDevice;
auto const GPUrtMemcpyDeviceToHost = hipMemcpyDeviceToHost;
auto const GPUrtMemcpyHostToDevice = hipMemcpyHostToDevice;
auto const GPUrtMemcpyDefault = hipMemcpyDefault;

template <typename T>
inline hipError_t GPUrtMalloc(T** p, size_t bytes) { return hipMalloc((void**)p, bytes); }
template <typename T>
inline hipError_t GPUrtMallocManaged(T** p, size_t bytes) { return hipMallocManaged((void**)p, bytes); }
inline hipError_t GPUrtFree(void* p) { return hipFree(p); }
inline hipError_t GPUrtDeviceSynchronize() { return hipDeviceSynchronize(); }
inline const char* GPUrtGetErrorString(hipError_t e) { return hipGetErrorString(e); }
inline hipError_t GPUrtGetLastError() { return hipGetLastError(); }
inline hipError_t GPUrtEventCreate(hipEvent_t* e) { return hipEventCreate(e); }
inline hipError_t GPUrtEventRecord(hipEvent_t e, hipStream_t s = nullptr) { return hipEventRecord(e, s); }
inline hipError_t GPUrtEventSynchronize(hipEvent_t e) { return hipEventSynchronize(e); }
inline hipError_t GPUrtEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) { return hipEventElapsedTime(ms, a, b); }
inline hipError_t GPUrtMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    return hipMemcpy(dst, src, bytes, kind);
}
inline hipError_t GPUrtMemset(void* dst, int value, size_t bytes) { return hipMemset(dst, value, bytes); }
inline hipError_t GPUrtSetDevice(int device) { return hipSetDevice(device); }

}  // namespace adapt
