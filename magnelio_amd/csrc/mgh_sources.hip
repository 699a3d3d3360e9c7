// TF/SF plane-wave face corrections.
//
// Reference: PlaneWaveSource._apply / inject_E / inject_H
// (src/magnelio/sources/plane_wave.py:710-734): after the E (H) update, every
// face patch adds  coef * f(t - delay)  to its field entries, patch after
// patch, with coef = beta * (factor * metric) folded at attach time and the
// waveform sample cast to the field dtype before the product.
//
// Device form.  The patches are flattened into one correction list grouped
// by target element (CSR): target t owns terms [ptr[t], ptr[t+1]) in patch
// order, so an element that two patches touch (a box edge) gets its two
// additions in the reference's order, one thread per element, no atomics.
// The waveform is a host-built table wave[n * n_delays + d] of samples
// already cast to the field dtype (the retardations k.r / c0 take few
// distinct values for axis-aligned incidence); the step index n comes from
// the device step counter, so the launch can sit in a captured graph.
#include "mgh_common.h"

namespace {

template <typename T>
__global__ __launch_bounds__(256) void k_tfsf(T* __restrict__ f, int64_t set_stride, int n_targets,
                                              const long long* __restrict__ targets,
                                              const int* __restrict__ ptr,
                                              const T* __restrict__ coef,
                                              const int* __restrict__ did,
                                              const T* __restrict__ wave, int n_delays,
                                              long long tab_len, const long long* __restrict__ ctr,
                                              long long step_offset) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_targets) return;
  const long long n = *ctr + step_offset;
  if (n < 0 || n >= tab_len) return;  // outside the tabulated run: no injection
  T* fs = f + (int64_t)blockIdx.y * set_stride;
  const long long idx = targets[t];
  const T* row = wave + n * (long long)n_delays;
  T v = fs[idx];
  for (int q = ptr[t]; q < ptr[t + 1]; ++q) v = v + (coef[q] * row[did[q]]);
  fs[idx] = v;
}

}  // namespace

extern "C" {

// Adds the TF/SF corrections of one phase to nsets field vectors (stride
// set_stride elements).  targets: n_targets int64 device-layout indices into
// the E (or H) vector; ptr: n_targets + 1 int32 offsets into coef / did;
// wave: tab_len x n_delays samples; row n = *step_ctr + step_offset.
int mgh_tfsf_apply(int dtype, int nsets, void* field, long long set_stride, int n_targets,
                   const long long* targets, const int* ptr, const void* coef, const int* did,
                   const void* wave, int n_delays, long long tab_len, const long long* step_ctr,
                   long long step_offset, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8");
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "bad nsets %d", nsets);
  MGH_REQUIRE(n_targets >= 0 && n_delays >= 1 && tab_len >= 0, "bad table sizes");
  if (n_targets == 0) return MGH_OK;
  MGH_REQUIRE(field && targets && ptr && coef && did && wave && step_ctr, "null pointer");
  dim3 grd((n_targets + 255) / 256, nsets, 1);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    hipLaunchKernelGGL(k_tfsf<float>, grd, dim3(256), 0, st, (float*)field, set_stride, n_targets,
                       targets, ptr, (const float*)coef, did, (const float*)wave, n_delays, tab_len,
                       step_ctr, step_offset);
  else
    hipLaunchKernelGGL(k_tfsf<double>, grd, dim3(256), 0, st, (double*)field, set_stride,
                       n_targets, targets, ptr, (const double*)coef, did, (const double*)wave,
                       n_delays, tab_len, step_ctr, step_offset);
  MGH_CHECK_LAUNCH("k_tfsf");
  return MGH_OK;
}

}  // extern "C"
