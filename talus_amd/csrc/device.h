// device.h -- HIP error handling shared by the device sources of libtalus_mi.
#pragma once

#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>

namespace tmi {

// A failed HIP call; the C ABI reports it as TALUS_MI_ERR_DEVICE.
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

}  // namespace tmi

#define HIPCHECK(expr)                                                                       \
  do {                                                                                       \
    const hipError_t status_ = (expr);                                                       \
    if (status_ != hipSuccess)                                                               \
      throw ::tmi::DeviceError(std::string(#expr) + " failed: " + hipGetErrorString(status_)); \
  } while (0)
