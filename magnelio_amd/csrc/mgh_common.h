// Shared helpers for the magnelio-hip C-ABI: status codes, the thread-local
// error message, and the device layout of the Yee field components.
//
// Device layout ("x-fastest"): every field component keeps the reference's
// component block order [Ex|Ey|Ez] and [Hx|Hy|Hz] with the reference's block
// lengths, but inside a block the element (i, j, k) lives at
//     k * (n1 * n0) + j * n0 + i
// where (n0, n1, n2) is the component's shape along (x, y, z).  The reference
// stores the same block in C order with z fastest
// (src/magnelio/_fields/field_arrays.py:190-196, mesh/indexing.py:17-63);
// the host side owns the exact permutation between the two.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>

#define MGH_OK 0
#define MGH_EINVAL 1
#define MGH_EHIP 2
#define MGH_ERANGE 3

namespace mgh {

void set_error(const char* fmt, ...);

// Field-component shapes of an Nx x Ny x Nz cell grid (x, y, z extents).
struct Dims {
  int nx, ny, nz;
  // E components
  __host__ __device__ int64_t n_ex() const { return (int64_t)nx * (ny + 1) * (nz + 1); }
  __host__ __device__ int64_t n_ey() const { return (int64_t)(nx + 1) * ny * (nz + 1); }
  __host__ __device__ int64_t n_ez() const { return (int64_t)(nx + 1) * (ny + 1) * nz; }
  // H components
  __host__ __device__ int64_t n_hx() const { return (int64_t)(nx + 1) * ny * nz; }
  __host__ __device__ int64_t n_hy() const { return (int64_t)nx * (ny + 1) * nz; }
  __host__ __device__ int64_t n_hz() const { return (int64_t)nx * ny * (nz + 1); }
  __host__ __device__ int64_t n_e() const { return n_ex() + n_ey() + n_ez(); }
  __host__ __device__ int64_t n_h() const { return n_hx() + n_hy() + n_hz(); }
};

}  // namespace mgh

#define MGH_CHECK_HIP(expr)                                                        \
  do {                                                                              \
    hipError_t err_ = (expr);                                                       \
    if (err_ != hipSuccess) {                                                       \
      mgh::set_error("%s failed: %s", #expr, hipGetErrorString(err_));             \
      return MGH_EHIP;                                                              \
    }                                                                               \
  } while (0)

#define MGH_CHECK_LAUNCH(name)                                                     \
  do {                                                                              \
    hipError_t err_ = hipGetLastError();                                            \
    if (err_ != hipSuccess) {                                                       \
      mgh::set_error("launch of %s failed: %s", name, hipGetErrorString(err_));    \
      return MGH_EHIP;                                                              \
    }                                                                               \
  } while (0)

#define MGH_REQUIRE(cond, ...)                                                     \
  do {                                                                              \
    if (!(cond)) {                                                                  \
      mgh::set_error(__VA_ARGS__);                                                  \
      return MGH_EINVAL;                                                            \
    }                                                                               \
  } while (0)
