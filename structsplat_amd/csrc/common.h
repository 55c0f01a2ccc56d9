// common.h - pieces shared by the sources of libssplat_hip.so: error reporting, the launch
// check, the XCD-aware workgroup order and the field table of batched calls.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ssplat.h"

namespace ssplat_detail {

// Records a printf-style message for ssplat_last_error() and returns `code`.
int set_error(int code, const char* fmt, ...);
// Clears the message and returns SSPLAT_OK.
int success();
// The calling thread's last message.
const char* last_error();

// MI355X: 8 XCDs; workgroups are dispatched round-robin across them.
constexpr int kXcds = 8;

// Workgroup -> tile mapping that gives each XCD a contiguous run of tiles, so neighbouring
// tiles (which share most of their data) hit the same XCD-private L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  const int per = nb / kXcds;
  const int rem = nb % kXcds;
  const int xcd = bid % kXcds;
  const int local = bid / kXcds;
  return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Fields: one image and its Gaussians.  A batch concatenates the fields' Gaussians, tiles and
// pixels; every kernel finds its field by binary search over the (device) field table.  A single
// field travels by value in the kernel arguments instead.
struct FieldTable {
  const ssplat_field* fields;  // device table of num_fields entries, or NULL for `single`
  int num_fields;
  ssplat_field single;
};

// Field owning global tile `tile` (the last field whose first_tile <= tile).
__device__ __forceinline__ ssplat_field field_of_tile(const FieldTable& t, int tile) {
  if (t.fields == nullptr) return t.single;
  int lo = 0, hi = t.num_fields - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t.fields[mid].first_tile <= tile) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return t.fields[lo];
}

// Field owning global Gaussian `i` (the last field whose first_gaussian <= i; fields are
// contiguous, so an empty field never wins over the next one).
__device__ __forceinline__ ssplat_field field_of_gaussian(const FieldTable& t, int i) {
  if (t.fields == nullptr) return t.single;
  int lo = 0, hi = t.num_fields - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t.fields[mid].first_gaussian <= i) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return t.fields[lo];
}

}  // namespace ssplat_detail

#define SSPLAT_LAUNCH_CHECK(name)                                                             \
  do {                                                                                        \
    hipError_t err_ = hipGetLastError();                                                      \
    if (err_ != hipSuccess) {                                                                 \
      return ::ssplat_detail::set_error(SSPLAT_ERR_HIP, "%s launch failed: %s", name,         \
                                        hipGetErrorString(err_));                             \
    }                                                                                         \
  } while (0)
