// Lumped RLC elements and discrete ports on grid-edge chains.
//
// Reference: LumpedElementOperator.update_e (src/magnelio/ports/_lumped/
// operator.py:134-166) with the trapezoidal companions SeriesRLC /
// ParallelRLC (src/magnelio/circuit/companion.py:45-180):
//   v_total = sum_q sign_q e_q                         (float64, chain order)
//   i       = ((v_src - v_hist) - v_total) / (r_eq + sum beta)
//   e_q    += (sign_q beta_q) i                        (float64, cast back)
//   companion advance with (i, r_eq i + v_hist)
// then the recorder's V = sum_q sign_q e_q (after the update), I = i.
//
// One 64-lane work-group per (element, field set): lanes gather and scatter
// the chain's edges, lane 0 runs the scalar recursion in float64 in the
// reference's operation order.  The companion state (float64) lives on the
// device; the step index comes from the device step counter.
#include "mgh_common.h"

namespace {

constexpr int LUMPED_WORDS = 12;
constexpr int LUMPED_MAX_EDGES = 1024;
constexpr int LUMPED_PER_LAUNCH = 16;

struct Elem {
  int ne;
  const long long* idx;
  const double* sign;
  const double* beta;
  const double* prm;
  double* state;
  const int* exc;
  const double* src;
  double* vrec;
  double* irec;
  long long tab_len, cap;
};

struct ElemSet {
  Elem el[LUMPED_PER_LAUNCH];
};

template <typename T>
__global__ __launch_bounds__(64) void k_lumped(T* __restrict__ e, int64_t e_set, ElemSet set,
                                               const long long* __restrict__ ctr) {
  __shared__ double sv[LUMPED_MAX_EDGES];
  __shared__ double s_i;
  const Elem& E = set.el[blockIdx.x];
  const int ne = E.ne;
  const long long* idx = E.idx;
  const double* sign = E.sign;
  const double* beta = E.beta;
  const double* prm = E.prm;
  double* state = E.state + (int64_t)blockIdx.y * 5;
  const int* exc = E.exc;
  const double* src = E.src;
  const long long tab_len = E.tab_len;
  double* vrec = E.vrec;
  double* irec = E.irec;
  const long long cap = E.cap;
  const int s = blockIdx.y;
  const int lane = threadIdx.x;
  T* es = e + (int64_t)s * e_set;
  const long long n = *ctr;

  for (int q = lane; q < ne; q += 64) sv[q] = sign[q] * (double)es[idx[q]];
  __syncthreads();
  if (lane == 0) {
    double v_total = 0.0;
    for (int q = 0; q < ne; ++q) v_total += sv[q];
    double v_src = 0.0;
    if (exc[s] && n >= 0 && n < tab_len) v_src = src[(int64_t)s * tab_len + n];
    const bool parallel = prm[0] != 0.0, has_l = prm[1] != 0.0, has_c = prm[2] != 0.0;
    const double r_eq = prm[3], beta_sum = prm[4], kl = prm[5], kc = prm[6], g = prm[7];
    double v_hist;
    if (!parallel) {  // series: x0 = I^n, x1 = V_L^n, x2 = V_C^n
      v_hist = 0.0;
      if (has_l) v_hist += (-kl) * state[0] - state[1];
      if (has_c) v_hist += state[2] + kc * state[0];
    } else {  // parallel: x0 = V^n, x1 = I_L^n, x2 = I_C^n
      double ih = 0.0;
      if (has_l) ih += state[1] + kl * state[0];
      if (has_c) ih += (-kc) * state[0] - state[2];
      v_hist = -ih / g;
    }
    const double z_eff = r_eq + beta_sum;
    const double i = ((v_src - v_hist) - v_total) / z_eff;
    if (!parallel) {
      if (has_l) state[1] = kl * (i - state[0]) - state[1];
      if (has_c) state[2] = state[2] + kc * (i + state[0]);
      state[0] = i;
    } else {
      const double v = r_eq * i + v_hist;
      if (has_l) state[1] = state[1] + kl * (v + state[0]);
      if (has_c) state[2] = kc * (v - state[0]) - state[2];
      state[0] = v;
    }
    state[3] = v_total + i * beta_sum;
    state[4] = i;
    s_i = i;
  }
  __syncthreads();
  const double i = s_i;
  for (int q = lane; q < ne; q += 64) {
    const int64_t a = idx[q];
    const T v = (T)((double)es[a] + (sign[q] * beta[q]) * i);
    es[a] = v;
    sv[q] = sign[q] * (double)v;
  }
  if (vrec == nullptr) return;
  __syncthreads();
  if (lane == 0 && n >= 0 && n < cap) {
    double v = 0.0;
    for (int q = 0; q < ne; ++q) v += sv[q];
    vrec[(int64_t)s * cap + n] = v;
    irec[(int64_t)s * cap + n] = i;
  }
}

}  // namespace

extern "C" {

// desc: host array of n_elems records of 12 int64 words:
//   0 n_edges (<= 1024), 1 idx (int64 device-layout E indices),
//   2 sign (f64), 3 beta (f64), 4 params (f64[8]: parallel, has_L, has_C,
//   r_eq, sum beta, kL, kC, G), 5 state (f64, nsets x 5), 6 exc (int32 per
//   set, nonzero: the set drives this element), 7 src (f64, nsets x tab_len,
//   v_src per step), 8 tab_len, 9 V_rec, 10 I_rec (f64 nsets x cap, or 0),
//   11 cap.
int mgh_lumped_step(int dtype, int nsets, void* e, long long e_set_stride, const long long* desc,
                    int n_elems, const long long* step_ctr, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8");
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "bad nsets %d", nsets);
  MGH_REQUIRE(n_elems >= 0, "bad element count");
  if (n_elems == 0) return MGH_OK;
  MGH_REQUIRE(e && desc && step_ctr, "null pointer");
  hipStream_t st = (hipStream_t)stream;
  // elements apply concurrently within a launch (their edge chains are
  // disjoint, checked by the caller); launches of 16 follow in order
  for (int q0 = 0; q0 < n_elems; q0 += LUMPED_PER_LAUNCH) {
    const int cnt = (n_elems - q0) < LUMPED_PER_LAUNCH ? (n_elems - q0) : LUMPED_PER_LAUNCH;
    ElemSet es;
    std::memset(&es, 0, sizeof(es));
    for (int q = 0; q < cnt; ++q) {
      const long long* r = desc + (int64_t)(q0 + q) * LUMPED_WORDS;
      MGH_REQUIRE(r[0] >= 1 && r[0] <= LUMPED_MAX_EDGES, "element %d: %lld edges (1..%d)", q0 + q,
                  r[0], LUMPED_MAX_EDGES);
      MGH_REQUIRE(r[1] && r[2] && r[3] && r[4] && r[5] && r[6] && r[7],
                  "element %d: null pointer", q0 + q);
      MGH_REQUIRE((r[9] == 0) == (r[10] == 0), "element %d: V/I recording pointers", q0 + q);
      Elem& E = es.el[q];
      E.ne = (int)r[0];
      E.idx = (const long long*)r[1];
      E.sign = (const double*)r[2];
      E.beta = (const double*)r[3];
      E.prm = (const double*)r[4];
      E.state = (double*)r[5];
      E.exc = (const int*)r[6];
      E.src = (const double*)r[7];
      E.tab_len = r[8];
      E.vrec = (double*)r[9];
      E.irec = (double*)r[10];
      E.cap = r[11];
    }
    dim3 grd(cnt, nsets, 1);
    if (dtype == 4)
      hipLaunchKernelGGL(k_lumped<float>, grd, dim3(64), 0, st, (float*)e, e_set_stride, es,
                         step_ctr);
    else
      hipLaunchKernelGGL(k_lumped<double>, grd, dim3(64), 0, st, (double*)e, e_set_stride, es,
                         step_ctr);
    MGH_CHECK_LAUNCH("k_lumped");
  }
  return MGH_OK;
}

}  // extern "C"
