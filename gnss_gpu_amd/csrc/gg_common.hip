// Library-wide entry points: error reporting, version, device query.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "gg_common_host.h"
#include "gnss_gpu_amd.h"

namespace {
thread_local std::string g_err;
}

int gg_fail(int code, const char* msg) {
  g_err = msg ? msg : "";
  return code;
}

int gg_fail_hip(hipError_t e, const char* expr, const char* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "HIP error %d (%s) at %s:%d in %s", (int)e, hipGetErrorString(e),
                file, line, expr);
  g_err = buf;
  return -1;
}

std::string gg_last_error_str() { return g_err; }

extern "C" {

const char* gg_last_error(void) { return g_err.c_str(); }

int gg_abi_version(void) { return 1; }

int gg_device_count(int32_t* out) {
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  if (out) *out = n;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Device self-test hook: evaluates the deterministic building blocks on the
// GPU so tests can compare them bit for bit with the oracle.
//   op 0: out[i] = gg_exp64(in[i])
//   op 1: out[i] = gg_exp32((float)in[i])
//   op 2: Philox draw with counter (in[i], 7, 3, 1), key (seed 0x123456789abcdef0);
//         out[4i..4i+3] = the four 32-bit words
// ---------------------------------------------------------------------------
#include "gg_device.h"

__global__ void gg_k_debug_math(int op, const double* in, double* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (op == 0) {
    out[i] = gg_exp64(in[i]);
  } else if (op == 1) {
    out[i] = (double)gg_exp32((float)in[i]);
  } else if (op == 2) {
    const gg_u32x4 r = gg_draw(0x123456789abcdef0ull, (uint32_t)in[i], 7u, 3u, 1u);
    for (int k = 0; k < 4; ++k) out[4 * i + k] = (double)r.v[k];
  } else {
    gg_xo x = gg_xo_seed(0x123456789abcdef0ull, (uint32_t)in[i], 7u, 3u, 1u);
    for (int k = 0; k < 8; ++k) out[8 * i + k] = (double)gg_xo_next(x);
  }
}

extern "C" int gg_debug_math(int32_t op, const double* in, double* out, int32_t n) {
  if (n <= 0) return 0;
  if (op < 0 || op > 3) return gg_fail(-2, "unknown op");
  const size_t nout = (op == 2) ? 4 * (size_t)n : (op == 3 ? 8 * (size_t)n : (size_t)n);
  double* d = nullptr;
  GG_HIP(hipMalloc((void**)&d, sizeof(double) * (n + nout)));
  int rc = 0;
  if (hipMemcpy(d, in, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) rc = -1;
  if (!rc) {
    hipLaunchKernelGGL(gg_k_debug_math, dim3((n + 255) / 256), dim3(256), 0, 0, op, d, d + n, n);
    if (hipGetLastError() != hipSuccess) rc = -1;
  }
  if (!rc && hipMemcpy(out, d + n, sizeof(double) * nout, hipMemcpyDeviceToHost) != hipSuccess)
    rc = -1;
  if (hipFree(d) != hipSuccess) rc = -1;
  return rc ? gg_fail(-1, "gg_debug_math: HIP failure") : 0;
}
