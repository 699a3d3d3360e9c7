// Modal waveguide port step on the device: interior projection, exact DTBC
// recursion (or modal Mur-1), modal write-back of the port plane and the V/I
// recording — one work-group per (port, field set), float64 throughout.
//
// Reference (src/magnelio/ports/_modal/operator.py):
//   projections        _V_from_samples / _I_from_samples :1277-1321
//                      V_m = sum M_eps p_m e,  I_m = sum M_mu h_m h
//   update_e           :1351-1536  Mur total-field update, excited-Mur TF/SF,
//                      DTBC advance with the PREVIOUS step's interior
//                      projection, write e_port = sum_m V_m p_m (in mode order,
//                      float64, cast to the field dtype on store), then save
//                      V_port_prev / V_interior_prev and the |V| envelope
//   DTBC advance       src/magnelio/ports/_modal/dtbc.py:417-463
//   recorder           src/magnelio/ports/recorder.py:216-250 (V at t^{n+1}
//                      from the written plane, I at t^{n+1/2}, times
//                      record_scale)
// The step index n is read from a device counter, so the launch can sit in a
// captured graph.  Sums are block reductions in float64; their association
// differs from the reference's BLAS dots only at float64 round-off.
#include "mgh_common.h"
#include "mgh_ports.h"

namespace {

constexpr int PB = 256;

__device__ double block_sum(double v, double* sh) {
  // wave64 tree in LDS (PB = 256 threads -> 4 waves)
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = PB / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

template <typename T>
__global__ __launch_bounds__(PB) void k_port_step(T* __restrict__ e, int64_t e_set,
                                                  const T* __restrict__ h, int64_t h_set,
                                                  mgh::PortSet ps, const long long* step_ctr) {
  const mgh::PortDev& P = ps.p[blockIdx.x];
  const int s = blockIdx.y;
  const int n_pts = P.n_u + P.n_v;
  const int M = P.M;
  const long long n = *step_ctr;
  T* es = e + (int64_t)s * e_set;
  const T* hs = h + (int64_t)s * h_set;

  __shared__ double red[PB];
  __shared__ double vint[MGH_PORT_MAX_MODES];
  __shared__ double vcorr[MGH_PORT_MAX_MODES];
  __shared__ double conv_w[MGH_PORT_MAX_MODES];
  __shared__ double u_inc[MGH_PORT_MAX_MODES];

  // 1. interior-plane projection, u and v parts summed separately
  double acc[2 * MGH_PORT_MAX_MODES];
#pragma unroll
  for (int q = 0; q < 2 * MGH_PORT_MAX_MODES; ++q) acc[q] = 0.0;
  for (int p = threadIdx.x; p < n_pts; p += PB) {
    const double ev = (double)es[P.idx_int[p]];
    const double me = P.me_int[p];
    const int part = p < P.n_u ? 0 : 1;
#pragma unroll
    for (int m = 0; m < MGH_PORT_MAX_MODES; ++m)
      if (m < M) acc[2 * m + part] += me * (P.prof_e[(int64_t)m * n_pts + p] * ev);
  }
  for (int m = 0; m < M; ++m) {
    const double su = block_sum(acc[2 * m], red);
    const double sv = block_sum(acc[2 * m + 1], red);
    if (threadIdx.x == 0) vint[m] = su + sv;
  }

  // 2. DTBC history convolutions sum_{j=1..n} l_j hist[n-j]
  double* st = P.state + ((int64_t)s * M) * 4;
  for (int m = 0; m < M; ++m) {
    const double* prm = P.params + 8 * m;
    if (prm[0] != 1.0 || n == 0) {
      if (threadIdx.x == 0) { conv_w[m] = 0.0; u_inc[m] = 0.0; }
      continue;
    }
    // kernel version in force: smallest k_init*2^v >= n+1
    long long len = P.k_init, off = 0;
    while (len < n + 1) { off += len; len *= 2; }
    if (n > P.cap || off + len > P.kern_stride) {
      // history or kernel buffer too short for this step: poison the channel
      // (the host sizes both for the run; this guards every access)
      if (threadIdx.x == 0) { conv_w[m] = __builtin_nan(""); u_inc[m] = 0.0; }
      continue;
    }
    const double* l = P.kern + (int64_t)m * P.kern_stride + off;
    const double* wh = P.w_hist + ((int64_t)s * M + m) * P.cap;
    const double* sh = P.s_hist + ((int64_t)s * M + m) * P.cap;
    double cw = 0.0, ci = 0.0;
    for (long long j = 1 + threadIdx.x; j <= n; j += PB) {
      const double lj = l[j];
      cw += lj * wh[n - j];
      ci += lj * sh[n - j];
    }
    cw = block_sum(cw, red);
    ci = block_sum(ci, red);
    if (threadIdx.x == 0) { conv_w[m] = cw; u_inc[m] = ci; }
  }
  __syncthreads();

  // 3. per-mode termination (scalar float64, one thread)
  if (threadIdx.x == 0) {
    const int exc = P.exc_mode ? P.exc_mode[s] : -1;
    double env = 0.0;
    for (int m = 0; m < M; ++m) {
      const double* prm = P.params + 8 * m;
      double* sm = st + 4 * m;  // u, u_prev, V_port_prev, V_int_prev
      double v;
      if (prm[0] == 1.0) {
        const double src = (m == exc && n < P.tab_len) ? P.src_tab[(int64_t)s * P.tab_len + n] : 0.0;
        const double ghost = src + conv_w[m];
        double* wh = P.w_hist + ((int64_t)s * M + m) * P.cap;
        double* shh = P.s_hist + ((int64_t)s * M + m) * P.cap;
        if (n < P.cap) {
          wh[n] = sm[0] - u_inc[m];
          shh[n] = src;
        }
        const double u_new = (prm[5] * sm[0] - sm[1]) + prm[1] * (ghost + sm[3]);
        sm[1] = sm[0];
        sm[0] = u_new;
        v = u_new;
      } else if (m == exc && P.mur_tab != nullptr && n < P.tab_len) {
        const double* t4 = P.mur_tab + ((int64_t)s * P.tab_len + n) * 4;
        const double sin_now = vint[m] - t4[1];
        const double sip = sm[3] - t4[2];
        const double sfp = sm[2] - t4[3];
        v = (t4[0] + sip) + prm[3] * (sin_now - sfp);
      } else {
        v = sm[3] + prm[3] * (vint[m] - sm[2]);
      }
      vcorr[m] = v;
      sm[2] = v;
      sm[3] = vint[m];
      const double a = fabs(v) * prm[4];
      env = a > env ? a : env;
    }
    if (P.absmax && env > P.absmax[s]) P.absmax[s] = env;
  }
  __syncthreads();

  // 4. modal write-back and 5. recording
  const bool rec = P.V_rec != nullptr && n < P.rec_cap;
#pragma unroll
  for (int q = 0; q < 2 * MGH_PORT_MAX_MODES; ++q) acc[q] = 0.0;
  double acci[2 * MGH_PORT_MAX_MODES];
#pragma unroll
  for (int q = 0; q < 2 * MGH_PORT_MAX_MODES; ++q) acci[q] = 0.0;
  for (int p = threadIdx.x; p < n_pts; p += PB) {
    double val = 0.0;
    for (int m = 0; m < M; ++m) val += vcorr[m] * P.prof_e[(int64_t)m * n_pts + p];
    const T tv = (T)val;
    es[P.idx_port[p]] = tv;
    if (rec) {
      const double ev = (double)tv;
      const double hv = (double)hs[P.idx_h[p]];
      const double me = P.me_port[p], mh = P.mh[p];
      const int part = p < P.n_u ? 0 : 1;
#pragma unroll
      for (int m = 0; m < MGH_PORT_MAX_MODES; ++m) {
        if (m < M) {
          acc[2 * m + part] += me * (P.prof_e[(int64_t)m * n_pts + p] * ev);
          acci[2 * m + part] += mh * (P.prof_h[(int64_t)m * n_pts + p] * hv);
        }
      }
    }
  }
  if (rec) {
    for (int m = 0; m < M; ++m) {
      const double vu = block_sum(acc[2 * m], red);
      const double vv = block_sum(acc[2 * m + 1], red);
      // reference order: the u-edge part (h_u on the v edges) first
      const double iu = block_sum(acci[2 * m + 1], red);
      const double iv = block_sum(acci[2 * m], red);
      if (threadIdx.x == 0) {
        const double scale = P.params[8 * m + 4];
        const int64_t o = ((int64_t)s * M + m) * P.rec_cap + n;
        P.V_rec[o] = (vu + vv) * scale;
        P.I_rec[o] = (iu + iv) * scale;
      }
    }
  }
}

__global__ void k_step_advance(long long* ctr, long long by) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *ctr += by;
}

}  // namespace

extern "C" {

int mgh_port_step(int dtype, int nsets, void* e, long long e_set_stride, const void* h,
                  long long h_set_stride, const long long* port_desc, int n_ports,
                  const long long* step_ctr, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8, got %d", dtype);
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "bad nsets %d", nsets);
  MGH_REQUIRE(n_ports >= 1 && n_ports <= MGH_PORT_MAX_PORTS, "n_ports must be 1..%d, got %d",
              MGH_PORT_MAX_PORTS, n_ports);
  MGH_REQUIRE(e && h && port_desc && step_ctr, "null pointer");
  mgh::PortSet ps{};
  ps.n = n_ports;
  for (int i = 0; i < n_ports; ++i) {
    int rc = mgh::load_port(port_desc + (int64_t)i * MGH_PORT_WORDS, &ps.p[i]);
    if (rc) return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  dim3 grd(n_ports, nsets, 1);
  if (dtype == 4)
    hipLaunchKernelGGL(k_port_step<float>, grd, dim3(PB), 0, st, (float*)e, e_set_stride,
                       (const float*)h, h_set_stride, ps, step_ctr);
  else
    hipLaunchKernelGGL(k_port_step<double>, grd, dim3(PB), 0, st, (double*)e, e_set_stride,
                       (const double*)h, h_set_stride, ps, step_ctr);
  MGH_CHECK_LAUNCH("k_port_step");
  return MGH_OK;
}

int mgh_step_advance(long long* step_ctr, long long by, void* stream) {
  MGH_REQUIRE(step_ctr, "null step counter");
  hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, (hipStream_t)stream, step_ctr, by);
  MGH_CHECK_LAUNCH("k_step_advance");
  return MGH_OK;
}

}  // extern "C"
