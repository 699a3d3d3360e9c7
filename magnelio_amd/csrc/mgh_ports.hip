// Modal waveguide port step on the device: interior projection, exact DTBC
// recursion (or modal Mur-1), modal write-back of the port plane and the V/I
// recording, float64 throughout.  Four launches per step, so that a large
// port plane spreads over many work-groups:
//   k_port_project  (chunk, port, set)  interior-plane projections and DTBC
//                   history convolutions, chunk partials to the work area
//   k_port_advance  (port, set)         fixed-order sum of the partials, the
//                   scalar termination update, corrected modal voltages
//   k_port_write    (chunk, port, set)  modal write-back of the plane and the
//                   V/I recording partials
//   k_port_record   (port, set)         V/I of the step
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
// captured graph.  Sums are fixed-order reductions in float64 (thread, wave
// butterfly, waves, chunks): deterministic, and different from the
// reference's BLAS dots only at float64 round-off.
#include <algorithm>

#include "mgh_common.h"
#include "mgh_ports.h"

namespace {

constexpr int PB = 256;
constexpr int PR = MGH_PORT_CHUNK / PB;  // points per thread per chunk
constexpr int PQ = MGH_PORT_PQ;
constexpr int PH = 2 * MGH_PORT_MAX_MODES;  // offset of the second partial group

// Sum the live partial words of a work-group (wave butterfly, then the four
// waves in order) and store them to out[0..PQ).
__device__ void store_partials(double (&acc)[PQ], int M, double* __restrict__ out) {
  __shared__ double sh[PB / 64][PQ];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < PQ; ++q) {
    if ((q % PH) < 2 * M) {
      double v = acc[q];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) sh[w][q] = v;
    }
  }
  __syncthreads();
  const int q = threadIdx.x;
  if (q < PQ) out[q] = ((q % PH) < 2 * M) ? (sh[0][q] + sh[1][q]) + (sh[2][q] + sh[3][q]) : 0.0;
}

// Work area of a port: projection partials [set][chunk][PQ], record
// partials (same shape), corrected voltages [set][MGH_PORT_MAX_MODES].
__device__ inline int64_t rec_offset(int nch, int nsets) { return (int64_t)nsets * nch * PQ; }
__device__ inline int64_t vc_offset(int nch, int nsets) { return 2LL * nsets * nch * PQ; }

// DTBC kernel version in force at step n: smallest k_init * 2^v >= n + 1.
__device__ inline bool kernel_slice(const mgh::PortDev& P, long long n, long long* off) {
  long long len = P.k_init, o = 0;
  while (len < n + 1) { o += len; len *= 2; }
  *off = o;
  // history or kernel buffer too short for this step (the host sizes both
  // for the run; this guards every access)
  return n <= P.cap && o + len <= P.kern_stride;
}

template <typename T>
__global__ __launch_bounds__(PB) void k_port_project(mgh::PortSet ps) {
  const mgh::PortDev& P = ps.p[blockIdx.y];
  const int s = blockIdx.z, c = blockIdx.x;
  const int n_pts = P.n_u + P.n_v, M = P.M;
  const int nch = mgh::port_chunks(n_pts, P.cap);
  if (c >= nch || s >= P.nsets) return;  // whole work-group
  const long long n = *P.ctr;
  const T* es = (const T*)P.e + (int64_t)s * P.e_set;
  double acc[PQ];
#pragma unroll
  for (int q = 0; q < PQ; ++q) acc[q] = 0.0;
  // interior-plane projection, u and v parts kept apart
#pragma unroll
  for (int r = 0; r < PR; ++r) {
    const int p = c * MGH_PORT_CHUNK + r * PB + threadIdx.x;
    if (p < n_pts) {
      const double ev = (double)es[P.idx_int[p]];
      const double me = P.me_int[p];
      const int part = p < P.n_u ? 0 : 1;
#pragma unroll
      for (int m = 0; m < MGH_PORT_MAX_MODES; ++m) {
        // modes past M read row 0 into partial words that are never stored,
        // so every mode's load is straight-line code and they issue together
        const int mm = m < M ? m : 0;
        acc[2 * m + part] += me * (P.prof_e[(int64_t)mm * n_pts + p] * ev);
      }
    }
  }
  // DTBC history convolutions sum_{j=1..n} l_j hist[n-j], terms of this chunk.
  // The kernel slice is the same for every mode and the modes' flags are read
  // up front, so the loads of all modes are predicated lanes of one straight
  // stretch of code and issue together (a term outside j <= n, or of a mode
  // without the DTBC, adds +0).
  long long off;
  if (n >= 1 && (long long)c * MGH_PORT_CHUNK < n && kernel_slice(P, n, &off)) {
    bool dtbc[MGH_PORT_MAX_MODES];
#pragma unroll
    for (int m = 0; m < MGH_PORT_MAX_MODES; ++m) dtbc[m] = m < M && P.params[8 * m] == 1.0;
#pragma unroll
    for (int m = 0; m < MGH_PORT_MAX_MODES; ++m) {
      const double* l = P.kern + (int64_t)m * P.kern_stride + off;
      const double* wh = P.w_hist + ((int64_t)s * M + m) * P.cap;
      const double* sh = P.s_hist + ((int64_t)s * M + m) * P.cap;
#pragma unroll
      for (int r = 0; r < PR; ++r) {
        const long long j = 1 + (long long)c * MGH_PORT_CHUNK + r * PB + threadIdx.x;
        const bool ok = dtbc[m] && j <= n;
        const double lj = ok ? l[j] : 0.0;
        const double wv = ok ? wh[n - j] : 0.0;
        const double sv = ok ? sh[n - j] : 0.0;
        acc[PH + 2 * m] += lj * wv;
        acc[PH + 2 * m + 1] += lj * sv;
      }
    }
  }
  store_partials(acc, M, P.work + ((int64_t)s * nch + c) * PQ);
}

__global__ __launch_bounds__(64) void k_port_advance(mgh::PortSet ps) {
  const mgh::PortDev& P = ps.p[blockIdx.x];
  const int s = blockIdx.y;
  if (s >= P.nsets) return;  // whole work-group
  const int n_pts = P.n_u + P.n_v, M = P.M;
  const int nch = mgh::port_chunks(n_pts, P.cap);
  const long long n = *P.ctr;
  __shared__ double tot[PQ];
  const int q = threadIdx.x;
  if (q < PQ) {
    // chunk order kept; the unrolled loop issues the partials' loads together
    // (they were written on other XCDs, so each is a trip past this L2)
    double v = 0.0;
#pragma unroll 16
    for (int c = 0; c < nch; ++c) v += P.work[((int64_t)s * nch + c) * PQ + q];
    tot[q] = v;
  }
  __syncthreads();
  // one lane per mode (the modes' recursions are independent); lane 0 then
  // takes the envelope maximum in mode order
  __shared__ double amp[MGH_PORT_MAX_MODES];
  const int m = threadIdx.x;
  if (m < M) {
    double* st = P.state + ((int64_t)s * M) * 4;
    double* vcs = P.work + vc_offset(nch, P.nsets) + (int64_t)s * MGH_PORT_MAX_MODES;
    const int exc = P.exc_mode ? P.exc_mode[s] : -1;
    const double* prm = P.params + 8 * m;
    double* sm = st + 4 * m;  // u, u_prev, V_port_prev, V_int_prev
    const double vint = tot[2 * m] + tot[2 * m + 1];
    double v;
    if (prm[0] == 1.0) {
      double conv_w = 0.0, u_inc = 0.0;
      if (n >= 1) {
        long long off;
        if (kernel_slice(P, n, &off)) {
          conv_w = tot[PH + 2 * m];
          u_inc = tot[PH + 2 * m + 1];
        } else {
          conv_w = __builtin_nan("");
        }
      }
      const double src = (m == exc && n < P.tab_len) ? P.src_tab[(int64_t)s * P.tab_len + n] : 0.0;
      const double ghost = src + conv_w;
      double* wh = P.w_hist + ((int64_t)s * M + m) * P.cap;
      double* shh = P.s_hist + ((int64_t)s * M + m) * P.cap;
      if (n < P.cap) {
        wh[n] = sm[0] - u_inc;
        shh[n] = src;
      }
      const double u_new = (prm[5] * sm[0] - sm[1]) + prm[1] * (ghost + sm[3]);
      sm[1] = sm[0];
      sm[0] = u_new;
      v = u_new;
    } else if (m == exc && P.mur_tab != nullptr && n < P.tab_len) {
      const double* t4 = P.mur_tab + ((int64_t)s * P.tab_len + n) * 4;
      const double sin_now = vint - t4[1];
      const double sip = sm[3] - t4[2];
      const double sfp = sm[2] - t4[3];
      v = (t4[0] + sip) + prm[3] * (sin_now - sfp);
    } else {
      v = sm[3] + prm[3] * (vint - sm[2]);
    }
    vcs[m] = v;
    sm[2] = v;
    sm[3] = vint;
    amp[m] = fabs(v) * prm[4];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double env = 0.0;
  for (int q = 0; q < M; ++q) env = amp[q] > env ? amp[q] : env;
  if (P.absmax && env > P.absmax[s]) P.absmax[s] = env;
}
template <typename T>
__global__ __launch_bounds__(PB) void k_port_write(mgh::PortSet ps) {
  const mgh::PortDev& P = ps.p[blockIdx.y];
  const int s = blockIdx.z, c = blockIdx.x;
  const int n_pts = P.n_u + P.n_v, M = P.M;
  if (c >= mgh::port_plane_chunks(n_pts) || s >= P.nsets) return;  // whole work-group
  const int nch = mgh::port_chunks(n_pts, P.cap);
  const long long n = *P.ctr;
  const bool rec = P.V_rec != nullptr && n < P.rec_cap;
  const double* vcs = P.work + vc_offset(nch, P.nsets) + (int64_t)s * MGH_PORT_MAX_MODES;
  double vc[MGH_PORT_MAX_MODES];
#pragma unroll
  for (int m = 0; m < MGH_PORT_MAX_MODES; ++m) vc[m] = m < M ? vcs[m] : 0.0;
  T* es = (T*)P.e + (int64_t)s * P.e_set;
  const T* hs = (const T*)P.h + (int64_t)s * P.h_set;
  // complement absorber: the interior projections of this step (V_int_prev
  // after the advance) remove the port's channels from the interior plane
  const bool comp = P.comp_r != nullptr;
  double vi[MGH_PORT_MAX_MODES];
#pragma unroll
  for (int m = 0; m < MGH_PORT_MAX_MODES; ++m)
    vi[m] = (comp && m < M) ? P.state[((int64_t)s * M + m) * 4 + 3] : 0.0;
  double* cst = comp ? P.comp_state + (int64_t)s * 2 * n_pts : nullptr;
  double acc[PQ];
#pragma unroll
  for (int q = 0; q < PQ; ++q) acc[q] = 0.0;
#pragma unroll
  for (int r = 0; r < PR; ++r) {
    const int p = c * MGH_PORT_CHUNK + r * PB + threadIdx.x;
    if (p < n_pts) {
      double val = 0.0;
#pragma unroll
      for (int m = 0; m < MGH_PORT_MAX_MODES; ++m) {
        // modes past M: vc = 0, an exact zero term (loads issue together)
        const int mm = m < M ? m : 0;
        val += vc[m] * P.prof_e[(int64_t)mm * n_pts + p];
      }
      if (comp) {
        // remainder at the interior plane, channels removed in mode order,
        // advanced to the plane by the edge's Mur-1 and added to the write
        double rem = (double)es[P.idx_int[p]];
        for (int m = 0; m < M; ++m) rem -= vi[m] * P.prof_e[(int64_t)m * n_pts + p];
        const double cp = P.comp_live[p] * (cst[p] + P.comp_r[p] * (rem - cst[n_pts + p]));
        val += cp;
        cst[p] = rem;
        cst[n_pts + p] = cp;
      }
      const T tv = (T)val;
      es[P.idx_port[p]] = tv;
      if (rec) {
        const double ev = (double)tv;
        const double hv = (double)hs[P.idx_h[p]];
        const double me = P.me_port[p], mh = P.mh[p];
        const int part = p < P.n_u ? 0 : 1;
#pragma unroll
        for (int m = 0; m < MGH_PORT_MAX_MODES; ++m) {
          // modes past M: row 0 into partial words that are never stored
          const int mm = m < M ? m : 0;
          acc[2 * m + part] += me * (P.prof_e[(int64_t)mm * n_pts + p] * ev);
          acc[PH + 2 * m + part] += mh * (P.prof_h[(int64_t)mm * n_pts + p] * hv);
        }
      }
    }
  }
  if (rec) store_partials(acc, M, P.work + rec_offset(nch, P.nsets) + ((int64_t)s * nch + c) * PQ);
}

__global__ __launch_bounds__(64) void k_port_record(mgh::PortSet ps) {
  const mgh::PortDev& P = ps.p[blockIdx.x];
  const int s = blockIdx.y;
  if (s >= P.nsets) return;  // whole work-group
  const long long n = *P.ctr;
  if (P.V_rec == nullptr || n >= P.rec_cap) return;  // whole work-group
  const int n_pts = P.n_u + P.n_v, M = P.M;
  const int nch = mgh::port_chunks(n_pts, P.cap);
  const int npc = mgh::port_plane_chunks(n_pts);
  const double* part = P.work + rec_offset(nch, P.nsets);
  __shared__ double tot[PQ];
  const int q = threadIdx.x;
  if (q < PQ) {
    double v = 0.0;
#pragma unroll 16
    for (int c = 0; c < npc; ++c) v += part[((int64_t)s * nch + c) * PQ + q];
    tot[q] = v;
  }
  __syncthreads();
  const int m = threadIdx.x;  // one lane per mode
  if (m >= M) return;
  const double scale = P.params[8 * m + 4];
  const int64_t o = ((int64_t)s * M + m) * P.rec_cap + n;
  // reference order: the u-edge part (h_u on the v edges) first
  P.V_rec[o] = (tot[2 * m] + tot[2 * m + 1]) * scale;
  P.I_rec[o] = (tot[PH + 2 * m + 1] + tot[PH + 2 * m]) * scale;
}

__global__ void k_step_advance(long long* ctr, long long by) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *ctr += by;
}

// the step counters of packed solvers, one lane each
constexpr int MGH_PACK_CTRS = 64;
struct CtrTable {
  long long* c[MGH_PACK_CTRS];
};

__global__ void k_step_advance_packed(CtrTable t, int n, long long by) {
  if (blockIdx.x == 0 && (int)threadIdx.x < n) *t.c[threadIdx.x] += by;
}

// The four launches of a port step over the ports of ps (each carrying its
// field sets and step counter); max_sets: the largest set count among them.
int launch_port_step(int dtype, mgh::PortSet& ps, int max_sets, hipStream_t st) {
  int max_ch = 1, max_pch = 1;
  for (int i = 0; i < ps.n; ++i) {
    const int n_pts = ps.p[i].n_u + ps.p[i].n_v;
    max_ch = std::max(max_ch, mgh::port_chunks(n_pts, ps.p[i].cap));
    max_pch = std::max(max_pch, mgh::port_plane_chunks(n_pts));
  }
  const dim3 g_proj(max_ch, ps.n, max_sets), g_write(max_pch, ps.n, max_sets),
      g_one(ps.n, max_sets, 1);
  if (dtype == 4)
    hipLaunchKernelGGL(k_port_project<float>, g_proj, dim3(PB), 0, st, ps);
  else
    hipLaunchKernelGGL(k_port_project<double>, g_proj, dim3(PB), 0, st, ps);
  MGH_CHECK_LAUNCH("k_port_project");
  hipLaunchKernelGGL(k_port_advance, g_one, dim3(64), 0, st, ps);
  MGH_CHECK_LAUNCH("k_port_advance");
  if (dtype == 4)
    hipLaunchKernelGGL(k_port_write<float>, g_write, dim3(PB), 0, st, ps);
  else
    hipLaunchKernelGGL(k_port_write<double>, g_write, dim3(PB), 0, st, ps);
  MGH_CHECK_LAUNCH("k_port_write");
  hipLaunchKernelGGL(k_port_record, g_one, dim3(64), 0, st, ps);
  MGH_CHECK_LAUNCH("k_port_record");
  return MGH_OK;
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
    mgh::PortDev& P = ps.p[i];
    int rc = mgh::load_port(port_desc + (int64_t)i * MGH_PORT_WORDS, &P, nsets);
    if (rc) return rc;
    P.e = e;
    P.h = h;
    P.e_set = e_set_stride;
    P.h_set = h_set_stride;
    P.ctr = step_ctr;
    P.nsets = nsets;
  }
  return launch_port_step(dtype, ps, nsets, (hipStream_t)stream);
}

// The port steps of several solvers (a serving loop's packed small jobs) in
// the four launches of one step: port_desc holds n_ports descriptors, jobs
// per port 6 words: E, E set stride, H, H set stride, field sets, step counter
// of the solver the port belongs to.
int mgh_port_step_packed(int dtype, int n_ports, const long long* port_desc,
                         const long long* jobs, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8, got %d", dtype);
  MGH_REQUIRE(n_ports >= 1 && n_ports <= MGH_PORT_MAX_PORTS, "n_ports must be 1..%d, got %d",
              MGH_PORT_MAX_PORTS, n_ports);
  MGH_REQUIRE(port_desc && jobs, "null pointer");
  mgh::PortSet ps{};
  ps.n = n_ports;
  int max_sets = 1;
  for (int i = 0; i < n_ports; ++i) {
    const long long* w = jobs + (int64_t)i * 6;
    const int nsets = (int)w[4];
    MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "port %d: bad nsets %d", i, nsets);
    MGH_REQUIRE(w[0] && w[2] && w[5], "port %d: null field or counter pointer", i);
    mgh::PortDev& P = ps.p[i];
    int rc = mgh::load_port(port_desc + (int64_t)i * MGH_PORT_WORDS, &P, nsets);
    if (rc) return rc;
    P.e = (void*)w[0];
    P.e_set = w[1];
    P.h = (const void*)w[2];
    P.h_set = w[3];
    P.nsets = nsets;
    P.ctr = (const long long*)w[5];
    max_sets = std::max(max_sets, nsets);
  }
  return launch_port_step(dtype, ps, max_sets, (hipStream_t)stream);
}

int mgh_step_advance(long long* step_ctr, long long by, void* stream) {
  MGH_REQUIRE(step_ctr, "null step counter");
  hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, (hipStream_t)stream, step_ctr, by);
  MGH_CHECK_LAUNCH("k_step_advance");
  return MGH_OK;
}

// Advance the step counters of several solvers (packed jobs) by `by`.
int mgh_step_advance_packed(int n, const long long* ctrs, long long by, void* stream) {
  MGH_REQUIRE(n >= 1 && n <= MGH_PACK_CTRS, "n must be 1..%d, got %d", MGH_PACK_CTRS, n);
  MGH_REQUIRE(ctrs, "null counter table");
  CtrTable t{};
  for (int i = 0; i < n; ++i) {
    MGH_REQUIRE(ctrs[i], "null step counter %d", i);
    t.c[i] = (long long*)ctrs[i];
  }
  hipLaunchKernelGGL(k_step_advance_packed, dim3(1), dim3(64), 0, (hipStream_t)stream, t, n, by);
  MGH_CHECK_LAUNCH("k_step_advance_packed");
  return MGH_OK;
}

}  // extern "C"
