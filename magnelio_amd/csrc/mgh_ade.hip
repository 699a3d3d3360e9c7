// Trapezoidal ADE update of pole-residue dielectrics (E side).
//
// Reference: DispersionOperator.save_field / update_field
// (src/magnelio/solver/_dispersion.py:466-533), NumPy branch, with the
// coefficient W folded into alpha/beta on the host (fit_td.py:382-403):
//   top of the step:   f_prev = f[idx]
//   after the curl:    j = sum_q hw_q Re((1 + k_q) J_q)      (first term seeds)
//                      f = f - beta j
//                      gd = g (f - f_prev)
//                      J_q = J_q k_q;  J_q = J_q + c_q gd
// in the field precision T (k, c, 1 + k and hw are stored as T: the
// reference's Python scalars take the state arrays' precision).  A complex
// product a*b rounds as NumPy's SIMD multiply loop does on FMA hosts:
// re = fma(a.re, b.re, -(a.im b.im)), im = fma(a.re, b.im, a.im b.re), with
// a the left operand ((1 + k) J: a = 1 + k; J *= k: a = J); c gd (complex
// times real) is two plain products.  Each dispersive edge is one
// thread; edges are disjoint across blocks, so the history subtraction and
// the pole advance run in one pass.
#include "mgh_common.h"

namespace {

constexpr int ADE_TAB = 8;  // per (block, pole): kind, hw, opk_re, opk_im, k_re, k_im, c_re, c_im

__device__ __forceinline__ float fused(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fused(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__global__ __launch_bounds__(256) void k_ade_save(const T* __restrict__ e, int64_t e_set, int n,
                                                  const long long* __restrict__ idx,
                                                  T* __restrict__ fprev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = blockIdx.y;
  fprev[(int64_t)s * n + i] = e[(int64_t)s * e_set + idx[i]];
}

template <typename T>
__global__ __launch_bounds__(256) void k_ade_update(T* __restrict__ e, int64_t e_set, int n,
                                                    const long long* __restrict__ idx,
                                                    const T* __restrict__ g,
                                                    const T* __restrict__ beta,
                                                    const T* __restrict__ fprev,
                                                    const int* __restrict__ blk,
                                                    const int* __restrict__ npoles,
                                                    const T* __restrict__ tab, int pmax,
                                                    T* __restrict__ J) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = blockIdx.y;
  T* es = e + (int64_t)s * e_set;
  const int64_t a = idx[i];
  const int b = blk[i];
  const int np = npoles[b];
  T* Js = J + (int64_t)s * pmax * n * 2;
  T jh = T(0);
  for (int q = 0; q < np; ++q) {
    const T* t = tab + ((int64_t)b * pmax + q) * ADE_TAB;
    const T* jq = Js + ((int64_t)q * n + i) * 2;
    T term;
    if (t[0] == T(0)) {
      term = t[1] * (t[2] * jq[0]);
    } else {
      const T re = fused(t[2], jq[0], -(t[3] * jq[1]));
      term = t[1] * re;
    }
    jh = (q == 0) ? term : (jh + term);
  }
  T f = es[a];
  f = f - (beta[i] * jh);
  es[a] = f;
  const T gd = g[i] * (f - fprev[(int64_t)s * n + i]);
  for (int q = 0; q < np; ++q) {
    const T* t = tab + ((int64_t)b * pmax + q) * ADE_TAB;
    T* jq = Js + ((int64_t)q * n + i) * 2;
    if (t[0] == T(0)) {
      T x = jq[0] * t[4];
      x = x + (t[6] * gd);
      jq[0] = x;
    } else {
      const T xr = jq[0], xi = jq[1];
      T nr = fused(xr, t[4], -(xi * t[5]));
      T ni = fused(xr, t[5], xi * t[4]);
      nr = nr + (t[6] * gd);
      ni = ni + (t[7] * gd);
      jq[0] = nr;
      jq[1] = ni;
    }
  }
}

}  // namespace

extern "C" {

// States: n dispersive E edges (device-layout indices idx, int64), their
// g = EPS0 A_dual/L and beta in the field dtype, block id blk (int32).
// fprev: nsets x n.  Pole table tab: n_blocks x pmax x 8 field-dtype values
// (kind 0 real / 1 complex, hw = weight/2, 1+k, k, c as re/im pairs), npoles
// per block (int32).  J: nsets x pmax x n complex states (re, im interleaved;
// real poles use re).
int mgh_ade_save(int dtype, int nsets, const void* e, long long e_set_stride, int n,
                 const long long* idx, void* fprev, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8");
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535 && n >= 0, "bad sizes");
  if (n == 0) return MGH_OK;
  MGH_REQUIRE(e && idx && fprev, "null pointer");
  dim3 grd((n + 255) / 256, nsets, 1);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    hipLaunchKernelGGL(k_ade_save<float>, grd, dim3(256), 0, st, (const float*)e, e_set_stride, n,
                       idx, (float*)fprev);
  else
    hipLaunchKernelGGL(k_ade_save<double>, grd, dim3(256), 0, st, (const double*)e, e_set_stride,
                       n, idx, (double*)fprev);
  MGH_CHECK_LAUNCH("k_ade_save");
  return MGH_OK;
}

int mgh_ade_update(int dtype, int nsets, void* e, long long e_set_stride, int n,
                   const long long* idx, const void* g, const void* beta, const void* fprev,
                   const int* blk, const int* npoles, const void* tab, int pmax, void* J,
                   void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8");
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535 && n >= 0 && pmax >= 1, "bad sizes");
  if (n == 0) return MGH_OK;
  MGH_REQUIRE(e && idx && g && beta && fprev && blk && npoles && tab && J, "null pointer");
  dim3 grd((n + 255) / 256, nsets, 1);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    hipLaunchKernelGGL(k_ade_update<float>, grd, dim3(256), 0, st, (float*)e, e_set_stride, n, idx,
                       (const float*)g, (const float*)beta, (const float*)fprev, blk, npoles,
                       (const float*)tab, pmax, (float*)J);
  else
    hipLaunchKernelGGL(k_ade_update<double>, grd, dim3(256), 0, st, (double*)e, e_set_stride, n,
                       idx, (const double*)g, (const double*)beta, (const double*)fprev, blk,
                       npoles, (const double*)tab, pmax, (double*)J);
  MGH_CHECK_LAUNCH("k_ade_update");
  return MGH_OK;
}

}  // extern "C"
