// Shared helpers for the gfx950 kernels behind include/sn_api.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

#include "../../include/sn_api.h"

namespace sn {

// Thread-local error text returned by sn_last_error().
void set_error(const char* fmt, ...);
void clear_error();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Returns SN_EHIP (with message) if the last launch failed.
inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return SN_EHIP;
  }
  return SN_OK;
}

constexpr int kWave = 64;

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Grid size for grid-stride loops: enough workgroups to fill 256 CUs a few
// times over, but never more than the work needs.
inline unsigned grid_for(int64_t n, int block, int64_t max_blocks = 256 * 16) {
  int64_t b = ceil_div(n, block);
  if (b > max_blocks) b = max_blocks;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// scan.hip
size_t scan_ws_bytes(int64_t n);
int fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
int fill_f32(float* p, int64_t n, float v, hipStream_t s);
int scan_flags_dev(const uint8_t* flags, int64_t n_cap, const int32_t* n_dev, int32_t* offsets,
                   int32_t* total, void* ws, size_t ws_bytes, hipStream_t s);

// mlp.hip: out[c] (+)= sum over rows of part[row][c], fixed order
int mlp_reduce_rows(const float* part, int64_t rows, int cols, float* out, int accumulate,
                    hipStream_t s);

__device__ __forceinline__ int64_t dev_count(int64_t n_cap, const int32_t* n_dev) {
  if (!n_dev) return n_cap;
  const int64_t v = *n_dev;
  return v < 0 ? 0 : (v < n_cap ? v : n_cap);
}

}  // namespace sn

#define SN_REQUIRE(cond, ...)      \
  do {                             \
    if (!(cond)) {                 \
      sn::set_error(__VA_ARGS__);  \
      return SN_EINVAL;            \
    }                              \
  } while (0)

#define SN_HIP(call)                                                              \
  do {                                                                            \
    hipError_t _e = (call);                                                       \
    if (_e != hipSuccess) {                                                       \
      sn::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, \
                    __LINE__);                                                    \
      return SN_EHIP;                                                             \
    }                                                                             \
  } while (0)
