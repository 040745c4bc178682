// Shared helpers for the MORESANE HIP library (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace pms {

// ---------------------------------------------------------------------------
// error reporting: every C-ABI entry point returns an int status and leaves a
// message in a thread-local buffer that pms_last_error() returns.
// ---------------------------------------------------------------------------
enum Status : int {
  PMS_OK = 0,
  PMS_ERR_ARG = -1,      // invalid argument (shape, scale, mode ...)
  PMS_ERR_HIP = -2,      // a HIP runtime call failed
  PMS_ERR_ALLOC = -3,    // device allocation failed / budget exceeded
  PMS_ERR_STATE = -4,    // call out of order (library not initialised ...)
  PMS_ERR_INTERNAL = -5,
};

void set_error(const std::string& msg);
const char* last_error();

struct Error {
  int code;
  std::string msg;
};

// Throwing helpers used by the C++ implementation; the C-ABI layer converts
// them into a status code and the thread-local message.
[[noreturn]] void fail(int code, const std::string& msg);

#define PMS_HIP(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      ::pms::fail(::pms::PMS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define PMS_REQUIRE(cond, msg)                               \
  do {                                                       \
    if (!(cond)) ::pms::fail(::pms::PMS_ERR_ARG, (msg));     \
  } while (0)

#define PMS_LAUNCH_CHECK() PMS_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// Half-sample symmetric extension of [0, n): ... c b a | a b c ... x y z | z y x ...
// Periodic with period 2n, so any integer index maps back into range.
__device__ __forceinline__ int reflect_idx(int i, int n) {
  const int p = 2 * n;
  int m = i % p;
  if (m < 0) m += p;
  return (m >= n) ? (p - 1 - m) : m;
}

// The same extension for indices within two periods of the range, i in [-2n, 3n), without a
// division: the a trous taps never reach further (offsets up to 4*2^s with 2^(s+1) <= n).
__device__ __forceinline__ int fold_idx(int i, int n) {
  i = i < 0 ? i + 2 * n : i;
  i = i >= 2 * n ? i - 2 * n : i;
  return i >= n ? 2 * n - 1 - i : i;
}

// XCD-aware block remap: hardware dispatches consecutive workgroups
// round-robin over the 8 XCDs; this gives each XCD a contiguous range of
// logical blocks so neighbouring tiles share one L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned hw, unsigned nblocks) {
  const unsigned nx = 8;
  if (nblocks < 2 * nx) return hw;
  const unsigned per = nblocks / nx;          // blocks per XCD in the even part
  const unsigned even = per * nx;
  if (hw >= even) return hw;                  // ragged tail keeps its id
  return (hw % nx) * per + hw / nx;
}

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }
inline bool is_pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }
inline int ilog2(long v) {
  int r = 0;
  while ((1L << r) < v) ++r;
  return r;
}

}  // namespace pms
