// Material matrices and update coefficients built on the device from the
// cell material-id grid.
//
// Reference: build_M_eps / build_M_mu / build_M_sigma / build_M_sigma_m
// (src/magnelio/_operators/material_matrices.py:334-757, category-0 cells:
// owning cell by the clamped one-sided lookup, dual widths from
// _build_avg_d), the port-plane flattening of M_eps / M_mu / the PEC mask
// (:106-331), the PEC OR rule and bbox walls (src/magnelio/mesh/indexing.py:
// 168-257, mesher.py:1370-1430) and the semi-implicit coefficients of
// FITTimeDomainSolver.setup (src/magnelio/solver/fit_td.py:375-513):
//   M_eps   = (EPS0 eps_r[c]) ((A_y A_z) / L_x)   (Ex; cyclic for Ey, Ez)
//   M_sigma = sigma[c] ((A_y A_z) / L_x)          (PEC cells: 0)
//   M_mu    = (MU0 mu_r[c]) ((L_y L_z) / A_x)     (Hx; cyclic)
//   alpha = (M - (dt/2) S) / (M + (dt/2) S), beta = dt / (M + (dt/2) S),
//   M <= 0: alpha = 1, beta = 0;  PEC edge: alpha = beta = 0
// evaluated in float64 in that order (no contraction) and rounded once to
// the field dtype, so the result equals the host construction bit for bit.
// One thread per edge / face, device layout (x-fastest blocks).
#include "mgh_common.h"

namespace {

constexpr int TABW = 16;  // per material: eps xyz, mu xyz, sigma xyz, sigma_m xyz, is_pec

struct Geo {
  const double *dx, *dy, *dz;  // primal lengths (n)
  const double *ax, *ay, *az;  // dual widths (n + 1)
};

__device__ inline int clampc(int v, int n) { return v < n - 1 ? v : (n > 0 ? n - 1 : 0); }
__device__ inline int lowc(int v) { return v > 0 ? v - 1 : 0; }

// face bits: xmin 1, xmax 2, ymin 4, ymax 8, zmin 16, zmax 32
__device__ inline bool on_face(int bits, int axis, int pos, int n) {
  return ((bits >> (2 * axis)) & 1 && pos == 0) || ((bits >> (2 * axis + 1)) & 1 && pos == n);
}

// position of an entity after port flattening: the boundary plane of a port
// face reads the first interior plane.  The reference flattens face by face,
// each a copy of the first interior slab into the boundary slab; copies
// along different axes commute, so an entity on the boundary planes of two
// port axes reads the entity one plane in along both, whatever the order.
__device__ inline int flat_pos(int bits, int axis, int pos, int n) {
  if ((bits >> (2 * axis)) & 1 && pos == 0) return 1;
  if ((bits >> (2 * axis + 1)) & 1 && pos == n) return n - 1;
  return pos;
}

__device__ inline bool cell_pec(const int* __restrict__ mid, const double* __restrict__ tab,
                                int ny, int nz, int i, int j, int k) {
  return tab[(int64_t)mid[((int64_t)i * ny + j) * nz + k] * TABW + 12] != 0.0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_coeff_e(int nx, int ny, int nz, const int* __restrict__ mid,
                                                 const double* __restrict__ tab, Geo g, double dt,
                                                 double eps0, int walls, int ports,
                                                 T* __restrict__ alpha, T* __restrict__ beta,
                                                 T* __restrict__ mass) {
  const int64_t n0 = (int64_t)nx * (ny + 1) * (nz + 1);
  const int64_t n1 = (int64_t)(nx + 1) * ny * (nz + 1);
  const int64_t n2 = (int64_t)(nx + 1) * (ny + 1) * nz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n0 + n1 + n2) return;
  int c, i, j, k;
  if (t < n0) {
    c = 0; i = (int)(t % nx); j = (int)((t / nx) % (ny + 1)); k = (int)(t / ((int64_t)nx * (ny + 1)));
  } else if (t < n0 + n1) {
    const int64_t r = t - n0;
    c = 1; i = (int)(r % (nx + 1)); j = (int)((r / (nx + 1)) % ny); k = (int)(r / ((int64_t)(nx + 1) * ny));
  } else {
    const int64_t r = t - n0 - n1;
    c = 2; i = (int)(r % (nx + 1)); j = (int)((r / (nx + 1)) % (ny + 1));
    k = (int)(r / ((int64_t)(nx + 1) * (ny + 1)));
  }
  // flattened position (tangential components of port faces only)
  int fi = i, fj = j, fk = k;
  if (c != 0) fi = flat_pos(ports, 0, i, nx);
  if (c != 1) fj = flat_pos(ports, 1, j, ny);
  if (c != 2) fk = flat_pos(ports, 2, k, nz);
  double geo_f, geo_o;
  int64_t cell_f, cell_o;
  if (c == 0) {
    geo_f = (g.ay[fj] * g.az[fk]) / g.dx[fi];
    geo_o = (g.ay[j] * g.az[k]) / g.dx[i];
    cell_f = ((int64_t)fi * ny + clampc(fj, ny)) * nz + clampc(fk, nz);
    cell_o = ((int64_t)i * ny + clampc(j, ny)) * nz + clampc(k, nz);
  } else if (c == 1) {
    geo_f = (g.ax[fi] * g.az[fk]) / g.dy[fj];
    geo_o = (g.ax[i] * g.az[k]) / g.dy[j];
    cell_f = ((int64_t)clampc(fi, nx) * ny + fj) * nz + clampc(fk, nz);
    cell_o = ((int64_t)clampc(i, nx) * ny + j) * nz + clampc(k, nz);
  } else {
    geo_f = (g.ax[fi] * g.ay[fj]) / g.dz[fk];
    geo_o = (g.ax[i] * g.ay[j]) / g.dz[k];
    cell_f = ((int64_t)clampc(fi, nx) * ny + clampc(fj, ny)) * nz + fk;
    cell_o = ((int64_t)clampc(i, nx) * ny + clampc(j, ny)) * nz + k;
  }
  const double M = (eps0 * tab[(int64_t)mid[cell_f] * TABW + c]) * geo_f;
  const double S = tab[(int64_t)mid[cell_o] * TABW + 6 + c] * geo_o;
  // PEC: any adjacent cell (OR rule) or a wall face, at the flattened position
  bool pec = false;
  {
    int is[2] = {fi, fi}, js[2] = {fj, fj}, ks[2] = {fk, fk};
    if (c != 0) { is[0] = lowc(fi); is[1] = clampc(fi, nx); }
    if (c != 1) { js[0] = lowc(fj); js[1] = clampc(fj, ny); }
    if (c != 2) { ks[0] = lowc(fk); ks[1] = clampc(fk, nz); }
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 2; ++d) pec = pec || cell_pec(mid, tab, ny, nz, is[a], js[b], ks[d]);
    if (c != 0) pec = pec || on_face(walls, 0, fi, nx);
    if (c != 1) pec = pec || on_face(walls, 1, fj, ny);
    if (c != 2) pec = pec || on_face(walls, 2, fk, nz);
  }
  const double hd = 0.5 * dt;
  double a = 1.0, b = 0.0;
  if (M > 0.0) {
    const double den = M + hd * S;
    a = (M - hd * S) / den;
    b = dt / den;
  }
  if (pec) { a = 0.0; b = 0.0; }
  alpha[t] = (T)a;
  beta[t] = (T)b;
  if (mass) mass[t] = (T)M;
}

template <typename T>
__global__ __launch_bounds__(256) void k_coeff_h(int nx, int ny, int nz, const int* __restrict__ mid,
                                                 const double* __restrict__ tab, Geo g, double dt,
                                                 double mu0, int ports, T* __restrict__ alpha,
                                                 T* __restrict__ beta, T* __restrict__ mass) {
  const int64_t n0 = (int64_t)(nx + 1) * ny * nz;
  const int64_t n1 = (int64_t)nx * (ny + 1) * nz;
  const int64_t n2 = (int64_t)nx * ny * (nz + 1);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n0 + n1 + n2) return;
  int c, i, j, k;
  if (t < n0) {
    c = 0; i = (int)(t % (nx + 1)); j = (int)((t / (nx + 1)) % ny); k = (int)(t / ((int64_t)(nx + 1) * ny));
  } else if (t < n0 + n1) {
    const int64_t r = t - n0;
    c = 1; i = (int)(r % nx); j = (int)((r / nx) % (ny + 1)); k = (int)(r / ((int64_t)nx * (ny + 1)));
  } else {
    const int64_t r = t - n0 - n1;
    c = 2; i = (int)(r % nx); j = (int)((r / nx) % ny); k = (int)(r / ((int64_t)nx * ny));
  }
  // normal component of a port face: mass from the first interior plane
  int fi = i, fj = j, fk = k;
  if (c == 0) fi = flat_pos(ports, 0, i, nx);
  if (c == 1) fj = flat_pos(ports, 1, j, ny);
  if (c == 2) fk = flat_pos(ports, 2, k, nz);
  double geo_f, geo_o;
  int64_t cell_f, cell_o;
  if (c == 0) {
    geo_f = (g.dy[fj] * g.dz[fk]) / g.ax[fi];
    geo_o = (g.dy[j] * g.dz[k]) / g.ax[i];
    cell_f = ((int64_t)clampc(fi, nx) * ny + fj) * nz + fk;
    cell_o = ((int64_t)clampc(i, nx) * ny + j) * nz + k;
  } else if (c == 1) {
    geo_f = (g.dx[fi] * g.dz[fk]) / g.ay[fj];
    geo_o = (g.dx[i] * g.dz[k]) / g.ay[j];
    cell_f = ((int64_t)fi * ny + clampc(fj, ny)) * nz + fk;
    cell_o = ((int64_t)i * ny + clampc(j, ny)) * nz + k;
  } else {
    geo_f = (g.dx[fi] * g.dy[fj]) / g.az[fk];
    geo_o = (g.dx[i] * g.dy[j]) / g.az[k];
    cell_f = ((int64_t)fi * ny + fj) * nz + clampc(fk, nz);
    cell_o = ((int64_t)i * ny + j) * nz + clampc(k, nz);
  }
  const double M = (mu0 * tab[(int64_t)mid[cell_f] * TABW + 3 + c]) * geo_f;
  const double Sm = tab[(int64_t)mid[cell_o] * TABW + 9 + c] * geo_o;
  const double hd = 0.5 * dt;
  double a = 1.0, b = 0.0;
  if (M > 0.0) {
    const double den = M + hd * Sm;
    a = (M - hd * Sm) / den;
    b = dt / den;
  }
  alpha[t] = (T)a;
  beta[t] = (T)b;
  if (mass) mass[t] = (T)M;
}

}  // namespace

extern "C" {

int mgh_build_coefficients(int dtype, int nx, int ny, int nz, const int* mat_id, int n_mat,
                           const double* tab, const double* geo, double dt, double eps0, double mu0,
                           int wall_faces, int port_faces, void* alpha_e, void* beta_e, void* m_eps,
                           void* alpha_h, void* beta_h, void* m_mu, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8, got %d", dtype);
  MGH_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "bad grid %d x %d x %d", nx, ny, nz);
  MGH_REQUIRE(n_mat >= 1, "empty material table");
  MGH_REQUIRE(mat_id && tab && geo && alpha_e && beta_e && alpha_h && beta_h, "null pointer");
  MGH_REQUIRE(dt > 0.0, "dt must be positive");
  MGH_REQUIRE((wall_faces & ~63) == 0 && (port_faces & ~63) == 0, "bad face bits");
  const int dims[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a)
    MGH_REQUIRE(!((port_faces >> (2 * a)) & 3) || dims[a] >= 2, "port axis needs >= 2 cells");
  Geo g;
  g.dx = geo;
  g.dy = g.dx + nx;
  g.dz = g.dy + ny;
  g.ax = g.dz + nz;
  g.ay = g.ax + nx + 1;
  g.az = g.ay + ny + 1;
  const int64_t ne = (int64_t)nx * (ny + 1) * (nz + 1) + (int64_t)(nx + 1) * ny * (nz + 1) +
                     (int64_t)(nx + 1) * (ny + 1) * nz;
  const int64_t nh = (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz +
                     (int64_t)nx * ny * (nz + 1);
  hipStream_t st = (hipStream_t)stream;
  const dim3 blk(256), ge((unsigned)((ne + 255) / 256)), gh((unsigned)((nh + 255) / 256));
  if (dtype == 4) {
    hipLaunchKernelGGL(k_coeff_e<float>, ge, blk, 0, st, nx, ny, nz, mat_id, tab, g, dt, eps0,
                       wall_faces, port_faces, (float*)alpha_e, (float*)beta_e, (float*)m_eps);
    MGH_CHECK_LAUNCH("k_coeff_e");
    hipLaunchKernelGGL(k_coeff_h<float>, gh, blk, 0, st, nx, ny, nz, mat_id, tab, g, dt, mu0,
                       port_faces, (float*)alpha_h, (float*)beta_h, (float*)m_mu);
  } else {
    hipLaunchKernelGGL(k_coeff_e<double>, ge, blk, 0, st, nx, ny, nz, mat_id, tab, g, dt, eps0,
                       wall_faces, port_faces, (double*)alpha_e, (double*)beta_e, (double*)m_eps);
    MGH_CHECK_LAUNCH("k_coeff_e");
    hipLaunchKernelGGL(k_coeff_h<double>, gh, blk, 0, st, nx, ny, nz, mat_id, tab, g, dt, mu0,
                       port_faces, (double*)alpha_h, (double*)beta_h, (double*)m_mu);
  }
  MGH_CHECK_LAUNCH("k_coeff_h");
  return MGH_OK;
}

}  // extern "C"
