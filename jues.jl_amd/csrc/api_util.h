// Helpers for the C ABI translation units: status mapping and host-array uploads.
#pragma once

#include <cmath>

#include "jues_internal.h"

#define JUES_API_BEGIN(ctx)                         \
  if (!(ctx)) return JUES_BAD_ARG;                  \
  try {                                             \
    if (hipSetDevice((ctx)->device) != hipSuccess)  \
      throw ::jues::Error(JUES_HIP_ERROR, "hipSetDevice failed");

#define JUES_API_END(ctx)                                 \
  }                                                       \
  catch (const ::jues::Error& e) {                        \
    (ctx)->err = e.msg;                                   \
    (void)hipGetLastError();                              \
    return e.st;                                          \
  }                                                       \
  catch (const std::exception& e) {                       \
    (ctx)->err = e.what();                                \
    return JUES_HIP_ERROR;                                \
  }                                                       \
  return JUES_OK;

namespace jues {

// Device copy of a host FP64 array, converted to T.
template <typename T>
struct DeviceArray {
  DevBuf buf;
  T* p = nullptr;
  DeviceArray(jues_ctx* ctx, const double* host, size_t n) {
    if (n == 0) return;
    if (sizeof(T) == sizeof(double)) {
      buf.alloc(n * sizeof(double));
      h2d(ctx, buf.p, host, n * sizeof(double));
    } else {
      DevBuf tmp(n * sizeof(double));
      h2d(ctx, tmp.p, host, n * sizeof(double));
      buf.alloc(n * sizeof(T));
      convert<double, T>(ctx, tmp.as<double>(), buf.as<T>(), (long)n);
      sync(ctx);
    }
    p = buf.as<T>();
  }
};

}  // namespace jues
