// Error plumbing shared by the host side of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

int gg_fail(int code, const char* msg);
int gg_fail_hip(hipError_t e, const char* expr, const char* file, int line);
std::string gg_last_error_str();

// Packed device weights of a learned-likelihood handle (gg_learn.hip).
struct gg_mlp;
const void* gg_mlp_packed_weights(const gg_mlp* m);

#define GG_HIP(expr)                                                   \
  do {                                                                 \
    const hipError_t gg_e_ = (expr);                                   \
    if (gg_e_ != hipSuccess) return gg_fail_hip(gg_e_, #expr, __FILE__, __LINE__); \
  } while (0)
