/*
 * Oracle / CPU baseline: C + OpenMP restatement of the reference's Numba
 * fused kernels — TEST INFRASTRUCTURE ONLY (used by tests/ and by the
 * cpu_baseline leg of bench.py, never by the product).
 *
 * Follows src/magnelio/_operators/numba_kernels.py:36-127 (update_E_fused,
 * update_H_fused): reference layout (C order, z fastest), the outer loop over
 * i parallel (Numba prange), inner j and k loops, and on the E side a double
 * curl accumulator that starts from 0.0 even for float32 fields (:50), with
 * E = (float)((double)(a*E) + (double)b*curl).  The H side is field-dtype
 * arithmetic ((t1 - t2) - t3) + t4, H = (a*H) - (b*curl).
 * Compiled with -ffp-contract=off so every product and sum rounds separately,
 * as Numba (fastmath off) does.
 */
#include <stddef.h>
#include <stdint.h>

#define IDX(i, j, k, nj, nk) (((int64_t)(i) * (nj) + (j)) * (nk) + (k))

void fit_omp_update_e_f32(float* e, const float* h, const float* ae, const float* be, int nx,
                          int ny, int nz) {
  const int64_t nEx = (int64_t)nx * (ny + 1) * (nz + 1);
  const int64_t nEy = (int64_t)(nx + 1) * ny * (nz + 1);
  const int64_t nHx = (int64_t)(nx + 1) * ny * nz;
  const int64_t nHy = (int64_t)nx * (ny + 1) * nz;
  float *Ex = e, *Ey = e + nEx, *Ez = e + nEx + nEy;
  const float *Hx = h, *Hy = h + nHx, *Hz = h + nHx + nHy;
  const float *aEx = ae, *aEy = ae + nEx, *aEz = ae + nEx + nEy;
  const float *bEx = be, *bEy = be + nEx, *bEz = be + nEx + nEy;
  int ni, nj, nk;

  ni = nx; nj = ny + 1; nk = nz + 1; /* Ex */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ni; ++i)
    for (int j = 0; j < nj; ++j)
      for (int k = 0; k < nk; ++k) {
        double c = 0.0;
        if (j < nj - 1) c += (double)Hz[IDX(i, j, k, ny, nz + 1)];
        if (j > 0) c -= (double)Hz[IDX(i, j - 1, k, ny, nz + 1)];
        if (k < nk - 1) c -= (double)Hy[IDX(i, j, k, ny + 1, nz)];
        if (k > 0) c += (double)Hy[IDX(i, j, k - 1, ny + 1, nz)];
        const int64_t q = IDX(i, j, k, nj, nk);
        const float p = aEx[q] * Ex[q];
        Ex[q] = (float)((double)p + (double)bEx[q] * c);
      }
  ni = nx + 1; nj = ny; nk = nz + 1; /* Ey */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ni; ++i)
    for (int j = 0; j < nj; ++j)
      for (int k = 0; k < nk; ++k) {
        double c = 0.0;
        if (k < nk - 1) c += (double)Hx[IDX(i, j, k, ny, nz)];
        if (k > 0) c -= (double)Hx[IDX(i, j, k - 1, ny, nz)];
        if (i < ni - 1) c -= (double)Hz[IDX(i, j, k, ny, nz + 1)];
        if (i > 0) c += (double)Hz[IDX(i - 1, j, k, ny, nz + 1)];
        const int64_t q = IDX(i, j, k, nj, nk);
        const float p = aEy[q] * Ey[q];
        Ey[q] = (float)((double)p + (double)bEy[q] * c);
      }
  ni = nx + 1; nj = ny + 1; nk = nz; /* Ez */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ni; ++i)
    for (int j = 0; j < nj; ++j)
      for (int k = 0; k < nk; ++k) {
        double c = 0.0;
        if (i < ni - 1) c += (double)Hy[IDX(i, j, k, ny + 1, nz)];
        if (i > 0) c -= (double)Hy[IDX(i - 1, j, k, ny + 1, nz)];
        if (j < nj - 1) c -= (double)Hx[IDX(i, j, k, ny, nz)];
        if (j > 0) c += (double)Hx[IDX(i, j - 1, k, ny, nz)];
        const int64_t q = IDX(i, j, k, nj, nk);
        const float p = aEz[q] * Ez[q];
        Ez[q] = (float)((double)p + (double)bEz[q] * c);
      }
}

void fit_omp_update_h_f32(const float* e, float* h, const float* ah, const float* bh, int nx,
                          int ny, int nz) {
  const int64_t nEx = (int64_t)nx * (ny + 1) * (nz + 1);
  const int64_t nEy = (int64_t)(nx + 1) * ny * (nz + 1);
  const int64_t nHx = (int64_t)(nx + 1) * ny * nz;
  const int64_t nHy = (int64_t)nx * (ny + 1) * nz;
  const float *Ex = e, *Ey = e + nEx, *Ez = e + nEx + nEy;
  float *Hx = h, *Hy = h + nHx, *Hz = h + nHx + nHy;
  const float *aHx = ah, *aHy = ah + nHx, *aHz = ah + nHx + nHy;
  const float *bHx = bh, *bHy = bh + nHx, *bHz = bh + nHx + nHy;

#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx + 1; ++i) /* Hx (nx+1, ny, nz) */
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const float c = ((Ez[IDX(i, j + 1, k, ny + 1, nz)] - Ez[IDX(i, j, k, ny + 1, nz)]) -
                         Ey[IDX(i, j, k + 1, ny, nz + 1)]) + Ey[IDX(i, j, k, ny, nz + 1)];
        const int64_t q = IDX(i, j, k, ny, nz);
        Hx[q] = (aHx[q] * Hx[q]) - (bHx[q] * c);
      }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) /* Hy (nx, ny+1, nz) */
    for (int j = 0; j < ny + 1; ++j)
      for (int k = 0; k < nz; ++k) {
        const float c = ((Ex[IDX(i, j, k + 1, ny + 1, nz + 1)] - Ex[IDX(i, j, k, ny + 1, nz + 1)]) -
                         Ez[IDX(i + 1, j, k, ny + 1, nz)]) + Ez[IDX(i, j, k, ny + 1, nz)];
        const int64_t q = IDX(i, j, k, ny + 1, nz);
        Hy[q] = (aHy[q] * Hy[q]) - (bHy[q] * c);
      }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) /* Hz (nx, ny, nz+1) */
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz + 1; ++k) {
        const float c = ((Ey[IDX(i + 1, j, k, ny, nz + 1)] - Ey[IDX(i, j, k, ny, nz + 1)]) -
                         Ex[IDX(i, j + 1, k, ny + 1, nz + 1)]) + Ex[IDX(i, j, k, ny + 1, nz + 1)];
        const int64_t q = IDX(i, j, k, ny, nz + 1);
        Hz[q] = (aHz[q] * Hz[q]) - (bHz[q] * c);
      }
}

int fit_omp_max_threads(void);
#ifdef _OPENMP
#include <omp.h>
int fit_omp_max_threads(void) { return omp_get_max_threads(); }
void fit_omp_set_threads(int n) { omp_set_num_threads(n); }
#else
int fit_omp_max_threads(void) { return 1; }
void fit_omp_set_threads(int n) { (void)n; }
#endif
