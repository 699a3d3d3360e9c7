// Weighted square sums with float64 accumulation (the stored-energy check).
//
// Reference: FITTimeDomainSolver.run energy check
// (src/magnelio/solver/fit_td.py:1014-1019):
//     0.5 * (sum(M_eps * e * e, dtype=float64) + sum(M_mu * h * h, dtype=float64))
// The element product (M * x) * x is formed in the field precision and each
// product is widened to float64 before it is accumulated, as the reference's
// float64-accumulated sum of a field-dtype temporary does.  The reduction is
// deterministic: fixed per-block partial sums, then one ordered final pass.
#include "mgh_common.h"

#include <atomic>

namespace {

constexpr int RB = 256;           // threads per block
constexpr int RMAX_BLOCKS = 1024; // partial sums per field set

template <typename T>
__global__ __launch_bounds__(RB) void k_wsq_partial(const T* __restrict__ x, int64_t set_stride,
                                                    const T* __restrict__ w, int64_t n,
                                                    double* __restrict__ part) {
  const int s = blockIdx.y;
  const T* xs = x + (int64_t)s * set_stride;
  double acc = 0.0;
  for (int64_t idx = (int64_t)blockIdx.x * RB + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * RB) {
    const T v = xs[idx];
    const T p = (w[idx] * v) * v;
    acc += (double)p;
  }
  __shared__ double red[RB];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = RB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)s * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(RB) void k_wsq_final(const double* __restrict__ part, int nblk,
                                                  double* __restrict__ out, int accumulate,
                                                  double scale) {
  const int s = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblk; b += RB) acc += part[(int64_t)s * nblk + b];
  __shared__ double red[RB];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = RB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[s] = accumulate ? out[s] + scale * red[0] : scale * red[0];
}

// Launch the two reduction passes with the partial sums in work.
int wsq_launch(int dtype, long long n, int nsets, const void* x, long long set_stride,
               const void* w, double* out, int accumulate, double scale, double* work,
               hipStream_t st) {
  int nblk = (int)((n + RB - 1) / RB);
  if (nblk > RMAX_BLOCKS) nblk = RMAX_BLOCKS;
  if (nblk < 1) nblk = 1;
  dim3 g1(nblk, nsets, 1);
  if (dtype == 4)
    hipLaunchKernelGGL(k_wsq_partial<float>, g1, dim3(RB), 0, st, (const float*)x, set_stride,
                       (const float*)w, n, work);
  else
    hipLaunchKernelGGL(k_wsq_partial<double>, g1, dim3(RB), 0, st, (const double*)x, set_stride,
                       (const double*)w, n, work);
  MGH_CHECK_LAUNCH("k_wsq_partial");
  hipLaunchKernelGGL(k_wsq_final, dim3(nsets), dim3(RB), 0, st, (const double*)work, nblk, out,
                     accumulate, scale);
  MGH_CHECK_LAUNCH("k_wsq_final");
  return MGH_OK;
}

constexpr int NOALLOC_MAX_SETS = 64;

}  // namespace

// Scratch of the allocation-free reduction: the per-block partial sums of
// up to NOALLOC_MAX_SETS field sets (512 KiB).  Every GPU the library runs
// on holds its own copy until the process ends.  In host code the name is
// only a stand-in: kernels get the device address that hipGetSymbolAddress
// returns for the current device, resolved once per device.
static __device__ __align__(256) double scratch[RMAX_BLOCKS * NOALLOC_MAX_SETS];

namespace {

constexpr int MAX_DEVICES = 64;
std::atomic<double*> g_scratch[MAX_DEVICES];

int scratch_on_current_device(double** out) {
  int dev = 0;
  MGH_CHECK_HIP(hipGetDevice(&dev));
  MGH_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "device %d beyond the %d this library tracks", dev,
              MAX_DEVICES);
  double* p = g_scratch[dev].load(std::memory_order_acquire);
  if (p == nullptr) {
    void* q = nullptr;
    MGH_CHECK_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
    MGH_REQUIRE(q != nullptr, "no device address for the reduction scratch on device %d", dev);
    p = (double*)q;
    // two threads resolving the same device store the same address
    g_scratch[dev].store(p, std::memory_order_release);
  }
  *out = p;
  return MGH_OK;
}

}  // namespace

extern "C" {

// Number of float64 workspace words mgh_weighted_sq_sum needs for nsets sets.
long long mgh_weighted_sq_sum_workspace(int nsets) { return (long long)RMAX_BLOCKS * nsets; }

// out[s] (device, float64) = scale * sum_i (w[i] * x_s[i]) * x_s[i] for each
// of nsets sets (x_s = x + s*set_stride); with accumulate != 0 the value is
// added to out[s].  work must hold mgh_weighted_sq_sum_workspace(nsets)
// doubles.  Nothing is synchronised: the result stays on the device.
int mgh_weighted_sq_sum(int dtype, long long n, int nsets, const void* x, long long set_stride,
                        const void* w, double* out, int accumulate, double scale, double* work,
                        long long work_len, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8, got %d", dtype);
  MGH_REQUIRE(n >= 0 && nsets >= 1, "bad sizes n=%lld nsets=%d", n, nsets);
  MGH_REQUIRE(nsets <= 65535, "too many field sets");
  MGH_REQUIRE(x && w && out && work, "null device pointer");
  MGH_REQUIRE(work_len >= (long long)RMAX_BLOCKS * nsets, "workspace too small (%lld < %lld)",
              work_len, (long long)RMAX_BLOCKS * nsets);
  return wsq_launch(dtype, n, nsets, x, set_stride, w, out, accumulate, scale, work,
                    (hipStream_t)stream);
}

// mgh_weighted_sq_sum with its partial sums in the library's static scratch
// (at most NOALLOC_MAX_SETS sets): no workspace argument, no allocation.
// Every call on one GPU shares that GPU's scratch; calls that overlap on one
// GPU (different streams or threads) race.  Each partial is written before
// it is read, so nothing needs clearing between calls.
int mgh_weighted_sq_sum_noalloc(int dtype, long long n, int nsets, const void* x,
                                long long set_stride, const void* w, double* out, int accumulate,
                                double scale, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8, got %d", dtype);
  MGH_REQUIRE(n >= 0 && nsets >= 1, "bad sizes n=%lld nsets=%d", n, nsets);
  MGH_REQUIRE(nsets <= NOALLOC_MAX_SETS,
              "%d field sets exceed the static scratch of the allocation-free reduction (%d)",
              nsets, NOALLOC_MAX_SETS);
  MGH_REQUIRE(x && w && out, "null device pointer");
  double* work = nullptr;
  const int rc = scratch_on_current_device(&work);
  if (rc != MGH_OK) return rc;
  return wsq_launch(dtype, n, nsets, x, set_stride, w, out, accumulate, scale, work,
                    (hipStream_t)stream);
}

}  // extern "C"
