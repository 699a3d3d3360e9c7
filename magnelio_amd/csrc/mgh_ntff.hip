// Near-to-far-field transform of a Huygens patch set (FP64).
//
// Reference: ntff_transform (src/magnelio/post/far_field.py:139-271).  The
// reference conjugates the library's e^{-jwt} phasors on entry, sums
// e^{+jk r.r'} J dS' and conjugates the amplitudes back; the same numbers
// come from summing the unconjugated currents against e^{-jk r.r'} with the
// conjugated prefactor, which is what the kernels do:
//   N(r) = sum_p e^{-jk r.c_p} wJ_p,  L(r) = sum_p e^{-jk r.c_p} wM_p
//   E_theta = pre (eta N.theta + L.phi),  E_phi = pre (eta N.phi - L.theta)
// with wJ = area (n x H), wM = -area (n x E) and pre = +jk/(4 pi) (host).
//
// Two forms of the sums:
// * any patch set (mgh_ntff): k_ntff_partial, one thread per direction,
//   blockIdx.y a chunk of patches staged through LDS in tiles (every lane
//   reads the same patch: LDS broadcast), one FP64 sincos per (direction,
//   patch);
// * tensor faces (mgh_ntff_faces): a Huygens-box face, or its image, is a
//   grid c1 x c2 on the plane x_a = p, so the phase factors as
//   e^{-jk r_a p} e^{-jk r_1 c1_i} e^{-jk r_2 c2_j}.  k_ntff_face_tile holds
//   e^{-jk r_2 c2_j} of a 16-column tile in registers (two directions per
//   thread), sums each row T_i = sum_j B_j W_ij of the four currents (W
//   wave-uniform: scalar-cache loads, each feeding both directions), then
//   S += e^{-jk r_1 c1_i} T_i: one sincos per (direction, row, tile) instead
//   of per patch, 16 FP64 FMA per (direction, patch).
// Partial sums per (chunk or tile job, direction) go to a work array and
// the finish kernels add them in job order (the result does not depend on
// the launch) before the projections.
#include "mgh_common.h"

namespace {

constexpr int NTFF_THREADS = 256;
constexpr int NTFF_TILE = 128;  // patches per LDS tile
constexpr int PATCH_WORDS = 15;  // centre (3), wJ (3 complex), wM (3 complex)

__global__ __launch_bounds__(NTFF_THREADS) void k_ntff_partial(int np, const double* __restrict__ pat,
                                                               int nd, const double* __restrict__ dirs,
                                                               double k, int chunk,
                                                               double* __restrict__ part) {
  __shared__ double tile[NTFF_TILE * PATCH_WORDS];
  const int d = blockIdx.x * NTFF_THREADS + threadIdx.x;
  const int c = blockIdx.y;
  const long long p0 = (long long)c * chunk;
  const long long p1 = p0 + chunk < np ? p0 + chunk : np;
  double rx = 0.0, ry = 0.0, rz = 0.0;
  if (d < nd) {
    rx = dirs[(long long)d * 9 + 0];
    ry = dirs[(long long)d * 9 + 1];
    rz = dirs[(long long)d * 9 + 2];
  }
  double acc[12];
#pragma unroll
  for (int v = 0; v < 12; ++v) acc[v] = 0.0;
  for (long long t0 = p0; t0 < p1; t0 += NTFF_TILE) {
    const int nt = (int)(p1 - t0 < NTFF_TILE ? p1 - t0 : NTFF_TILE);
    __syncthreads();
    for (int x = threadIdx.x; x < nt * PATCH_WORDS; x += NTFF_THREADS)
      tile[x] = pat[t0 * PATCH_WORDS + x];
    __syncthreads();
    if (d < nd) {
      for (int p = 0; p < nt; ++p) {
        const double* q = tile + p * PATCH_WORDS;
        const double a = k * ((rx * q[0] + ry * q[1]) + rz * q[2]);
        double s, co;
        sincos(a, &s, &co);
        // e^{-ja} w = (co - j s)(wr + j wi)
#pragma unroll
        for (int v = 0; v < 6; ++v) {
          const double wr = q[3 + 2 * v], wi = q[4 + 2 * v];
          acc[2 * v] += co * wr + s * wi;
          acc[2 * v + 1] += co * wi - s * wr;
        }
      }
    }
  }
  if (d < nd) {
#pragma unroll
    for (int v = 0; v < 12; ++v) part[((long long)c * 12 + v) * nd + d] = acc[v];
  }
}

__global__ __launch_bounds__(NTFF_THREADS) void k_ntff_finish(int nd, int nchunks,
                                                              const double* __restrict__ part,
                                                              const double* __restrict__ dirs,
                                                              double pre_re, double pre_im, double eta,
                                                              double* __restrict__ out) {
  const int d = blockIdx.x * NTFF_THREADS + threadIdx.x;
  if (d >= nd) return;
  double s[12];
#pragma unroll
  for (int v = 0; v < 12; ++v) s[v] = 0.0;
  for (int c = 0; c < nchunks; ++c) {
#pragma unroll
    for (int v = 0; v < 12; ++v) s[v] += part[((long long)c * 12 + v) * nd + d];
  }
  const double* th = dirs + (long long)d * 9 + 3;
  const double* ph = dirs + (long long)d * 9 + 6;
  // N = (s0 + j s1, s2 + j s3, s4 + j s5), L likewise from s6
  double nth_r = 0.0, nth_i = 0.0, nph_r = 0.0, nph_i = 0.0;
  double lth_r = 0.0, lth_i = 0.0, lph_r = 0.0, lph_i = 0.0;
  for (int a = 0; a < 3; ++a) {
    nth_r += s[2 * a] * th[a];
    nth_i += s[2 * a + 1] * th[a];
    nph_r += s[2 * a] * ph[a];
    nph_i += s[2 * a + 1] * ph[a];
    lth_r += s[6 + 2 * a] * th[a];
    lth_i += s[7 + 2 * a] * th[a];
    lph_r += s[6 + 2 * a] * ph[a];
    lph_i += s[7 + 2 * a] * ph[a];
  }
  const double at_r = eta * nth_r + lph_r, at_i = eta * nth_i + lph_i;
  const double ap_r = eta * nph_r - lth_r, ap_i = eta * nph_i - lth_i;
  out[(long long)d * 4 + 0] = pre_re * at_r - pre_im * at_i;
  out[(long long)d * 4 + 1] = pre_re * at_i + pre_im * at_r;
  out[(long long)d * 4 + 2] = pre_re * ap_r - pre_im * ap_i;
  out[(long long)d * 4 + 3] = pre_re * ap_i + pre_im * ap_r;
}

constexpr int FACE_JT = 16;  // columns per tile job
constexpr int FACE_DESC = 6;  // axis, n1, n2, offset of c1, offset of c2, offset of W

__device__ inline void cmac(double& sr, double& si, double ar, double ai, double br, double bi) {
  sr += ar * br - ai * bi;
  si += ar * bi + ai * br;
}

// FACE_DPT directions per thread share every W load (scalar-cache loads:
// W is wave-uniform); directions d0 + q * blockDim for q < FACE_DPT
template <int DPT>
__global__ __launch_bounds__(NTFF_THREADS) void k_ntff_face_tile(
    const long long* __restrict__ desc, const double* __restrict__ planes,
    const double* __restrict__ coords, const double* __restrict__ W, const int* __restrict__ jobs,
    int nd, const double* __restrict__ dirs, double k, double* __restrict__ part) {
  const int d0 = blockIdx.x * NTFF_THREADS * DPT + threadIdx.x;
  const int job = blockIdx.y;
  const int f = jobs[job * 3 + 0];
  const int j0 = jobs[job * 3 + 1];
  const int nj = jobs[job * 3 + 2];
  const long long* fd = desc + (long long)f * FACE_DESC;
  const int axis = (int)fd[0];
  const int n1 = (int)fd[1];
  const int n2 = (int)fd[2];
  const double* c1 = coords + fd[3];
  const double* c2 = coords + fd[4];
  const double* w = W + fd[5];
  if (d0 >= nd) return;
  const int t1 = axis == 0 ? 1 : 0;
  const int t2 = axis == 2 ? 1 : 2;
  double ra[DPT], r1[DPT];
  double br[DPT][FACE_JT], bi[DPT][FACE_JT];
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const int d = d0 + q * NTFF_THREADS;
    const bool live = d < nd;
    const long long dd = live ? d : d0;
    ra[q] = dirs[dd * 9 + axis];
    r1[q] = dirs[dd * 9 + t1];
    const double r2 = dirs[dd * 9 + t2];
#pragma unroll
    for (int j = 0; j < FACE_JT; ++j) {
      if (j < nj) {
        double s, c;
        sincos(k * (r2 * c2[j0 + j]), &s, &c);
        br[q][j] = c;
        bi[q][j] = -s;
      } else {
        br[q][j] = 0.0;
        bi[q][j] = 0.0;
      }
    }
  }
  double S[DPT][8];
#pragma unroll
  for (int q = 0; q < DPT; ++q)
#pragma unroll
    for (int v = 0; v < 8; ++v) S[q][v] = 0.0;
  for (int i = 0; i < n1; ++i) {
    const double* wr = w + ((long long)i * n2 + j0) * 8;
    double T[DPT][8];
#pragma unroll
    for (int q = 0; q < DPT; ++q)
#pragma unroll
      for (int v = 0; v < 8; ++v) T[q][v] = 0.0;
#pragma unroll
    for (int j = 0; j < FACE_JT; ++j) {
      if (j < nj) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const double xr = wr[j * 8 + 2 * c], xi = wr[j * 8 + 2 * c + 1];
#pragma unroll
          for (int q = 0; q < DPT; ++q) cmac(T[q][2 * c], T[q][2 * c + 1], br[q][j], bi[q][j], xr, xi);
        }
      }
    }
    const double ci = c1[i];
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      double s, co;
      sincos(k * (r1[q] * ci), &s, &co);
#pragma unroll
      for (int c = 0; c < 4; ++c) cmac(S[q][2 * c], S[q][2 * c + 1], co, -s, T[q][2 * c], T[q][2 * c + 1]);
    }
  }
  const double pl = planes[f];
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const int d = d0 + q * NTFF_THREADS;
    if (d >= nd) continue;
    double s, co;
    sincos(k * (ra[q] * pl), &s, &co);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double xr = co * S[q][2 * c] + s * S[q][2 * c + 1];
      const double xi = co * S[q][2 * c + 1] - s * S[q][2 * c];
      part[((long long)job * 8 + 2 * c) * nd + d] = xr;
      part[((long long)job * 8 + 2 * c + 1) * nd + d] = xi;
    }
  }
}

constexpr int FACE_DPT = 2;

__global__ __launch_bounds__(NTFF_THREADS) void k_ntff_face_finish(
    int nd, int njobs, const long long* __restrict__ desc, const int* __restrict__ jobs,
    const double* __restrict__ part, const double* __restrict__ dirs, double pre_re, double pre_im,
    double eta, double* __restrict__ out) {
  const int d = blockIdx.x * NTFF_THREADS + threadIdx.x;
  if (d >= nd) return;
  double N[6], L[6];
#pragma unroll
  for (int v = 0; v < 6; ++v) {
    N[v] = 0.0;
    L[v] = 0.0;
  }
  for (int job = 0; job < njobs; ++job) {
    const int axis = (int)desc[(long long)jobs[job * 3] * FACE_DESC];
    const int t1 = axis == 0 ? 1 : 0;
    const int t2 = axis == 2 ? 1 : 2;
    const double* q = part + (long long)job * 8 * nd + d;
    N[2 * t1] += q[0];
    N[2 * t1 + 1] += q[(long long)nd];
    N[2 * t2] += q[2 * (long long)nd];
    N[2 * t2 + 1] += q[3 * (long long)nd];
    L[2 * t1] += q[4 * (long long)nd];
    L[2 * t1 + 1] += q[5 * (long long)nd];
    L[2 * t2] += q[6 * (long long)nd];
    L[2 * t2 + 1] += q[7 * (long long)nd];
  }
  const double* th = dirs + (long long)d * 9 + 3;
  const double* ph = dirs + (long long)d * 9 + 6;
  double nth_r = 0.0, nth_i = 0.0, nph_r = 0.0, nph_i = 0.0;
  double lth_r = 0.0, lth_i = 0.0, lph_r = 0.0, lph_i = 0.0;
  for (int a = 0; a < 3; ++a) {
    nth_r += N[2 * a] * th[a];
    nth_i += N[2 * a + 1] * th[a];
    nph_r += N[2 * a] * ph[a];
    nph_i += N[2 * a + 1] * ph[a];
    lth_r += L[2 * a] * th[a];
    lth_i += L[2 * a + 1] * th[a];
    lph_r += L[2 * a] * ph[a];
    lph_i += L[2 * a + 1] * ph[a];
  }
  const double at_r = eta * nth_r + lph_r, at_i = eta * nth_i + lph_i;
  const double ap_r = eta * nph_r - lth_r, ap_i = eta * nph_i - lth_i;
  out[(long long)d * 4 + 0] = pre_re * at_r - pre_im * at_i;
  out[(long long)d * 4 + 1] = pre_re * at_i + pre_im * at_r;
  out[(long long)d * 4 + 2] = pre_re * ap_r - pre_im * ap_i;
  out[(long long)d * 4 + 3] = pre_re * ap_i + pre_im * ap_r;
}

}  // namespace

extern "C" {

// patches: n_patches x 15 doubles (centre xyz, wJ xyz as (re, im), wM xyz as
// (re, im)); dirs: n_dirs x 9 doubles (r, theta, phi unit vectors); work:
// at least n_chunks x 12 x n_dirs doubles; out: n_dirs x 4 doubles
// (E_theta re, im, E_phi re, im).
int mgh_ntff(int n_patches, const double* patches, int n_dirs, const double* dirs, double k,
             double pre_re, double pre_im, double eta, int n_chunks, double* work, double* out,
             void* stream) {
  MGH_REQUIRE(n_patches >= 0 && n_dirs >= 1 && n_chunks >= 1 && n_chunks <= 65535, "bad sizes");
  MGH_REQUIRE((n_patches == 0 || patches) && dirs && work && out, "null pointer");
  hipStream_t st = (hipStream_t)stream;
  const int chunk = n_patches == 0 ? 0 : (n_patches + n_chunks - 1) / n_chunks;
  const int nb = (n_dirs + NTFF_THREADS - 1) / NTFF_THREADS;
  hipLaunchKernelGGL(k_ntff_partial, dim3(nb, n_chunks, 1), dim3(NTFF_THREADS), 0, st, n_patches,
                     patches, n_dirs, dirs, k, chunk, work);
  MGH_CHECK_LAUNCH("k_ntff_partial");
  hipLaunchKernelGGL(k_ntff_finish, dim3(nb, 1, 1), dim3(NTFF_THREADS), 0, st, n_dirs, n_chunks, work,
                     dirs, pre_re, pre_im, eta, out);
  MGH_CHECK_LAUNCH("k_ntff_finish");
  return MGH_OK;
}

// Tensor faces: face f is the grid c1 x c2 (coords[desc[f,3] ..], coords[desc[f,4] ..],
// n1 = desc[f,1], n2 = desc[f,2]) on the plane x_axis = planes[f], axis = desc[f,0]
// and its other two axes in increasing order; W[desc[f,5] + (i n2 + j) 8 ..] holds
// the four currents area (n x H)_t1, (n x H)_t2, -area (n x E)_t1, -(n x E)_t2 as
// (re, im).  jobs: n_jobs x 3 int32 (face, first column, columns <= 16), every
// column of every face in exactly one job.  work: n_jobs x 8 x n_dirs doubles.
int mgh_ntff_faces(int n_faces, const long long* face_desc, const double* planes,
                   const double* coords, const double* W, int n_jobs, const int* jobs, int n_dirs,
                   const double* dirs, double k, double pre_re, double pre_im, double eta,
                   double* work, double* out, void* stream) {
  MGH_REQUIRE(n_faces >= 1 && n_jobs >= 1 && n_jobs <= 65535 && n_dirs >= 1, "bad sizes");
  MGH_REQUIRE(face_desc && planes && coords && W && jobs && dirs && work && out, "null pointer");
  hipStream_t st = (hipStream_t)stream;
  const int nb = (n_dirs + NTFF_THREADS - 1) / NTFF_THREADS;
  const int nbt = (n_dirs + NTFF_THREADS * FACE_DPT - 1) / (NTFF_THREADS * FACE_DPT);
  hipLaunchKernelGGL(k_ntff_face_tile<FACE_DPT>, dim3(nbt, n_jobs, 1), dim3(NTFF_THREADS), 0, st,
                     face_desc, planes, coords, W, jobs, n_dirs, dirs, k, work);
  MGH_CHECK_LAUNCH("k_ntff_face_tile");
  hipLaunchKernelGGL(k_ntff_face_finish, dim3(nb, 1, 1), dim3(NTFF_THREADS), 0, st, n_dirs, n_jobs,
                     face_desc, jobs, work, dirs, pre_re, pre_im, eta, out);
  MGH_CHECK_LAUNCH("k_ntff_face_finish");
  return MGH_OK;
}

}  // extern "C"
