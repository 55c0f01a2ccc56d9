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

// A table entry, read into registers before the lookup's branch ends: otherwise the compiler
// merges the table read and the read of `single` into one load through a selected pointer,
// which copies the kernel arguments to scratch memory.
__device__ __forceinline__ ssplat_field load_field(const ssplat_field* p) {
  static_assert(sizeof(ssplat_field) == 48, "ssplat_field is six 8-byte words");
  const int2* w = reinterpret_cast<const int2*>(p);
  const int2 w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4], w5 = w[5];
  auto i64 = [](int2 v) {
    return static_cast<int64_t>((uint64_t(uint32_t(v.y)) << 32) | uint32_t(v.x));
  };
  ssplat_field f;
  f.height = w0.x;
  f.width = w0.y;
  f.first_gaussian = i64(w1);
  f.num_gaussians = i64(w2);
  f.first_pixel = i64(w3);
  f.first_tile = i64(w4);
  f.tiles_x = w5.x;
  f.num_tiles = w5.y;
  // The words are in registers here, so the reads cannot be moved past the caller's branch.
  asm volatile("" ::"v"(w0), "v"(w1), "v"(w2), "v"(w3), "v"(w4), "v"(w5));
  return f;
}

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
  return load_field(t.fields + lo);
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
  return load_field(t.fields + lo);
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
