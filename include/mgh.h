/*
 * mgh.h — C-ABI of libmgh.so, the MI355X (gfx950) FIT-TD stepping engine.
 *
 * Conventions
 *   - Every entry point returns an int status: 0 = ok, 1 = invalid argument,
 *     2 = HIP runtime error, 3 = size out of range.  mgh_last_error() returns
 *     a thread-local message for the last failure.  No C++ exception crosses
 *     the boundary.
 *   - All device memory is caller-owned (plain device pointers), except where
 *     a workspace size query says otherwise.  Every launch takes an explicit
 *     hipStream_t passed as void*; nothing synchronises the host.
 *   - dtype is the element width in bytes: 4 = float32, 8 = float64.
 *   - Field vectors use the reference's component blocks [Ex|Ey|Ez] and
 *     [Hx|Hy|Hz] with the reference's block lengths; inside a block the
 *     element (i,j,k) of a component of shape (n0,n1,n2) sits at
 *     k*n1*n0 + j*n0 + i ("x-fastest").  The reference orders the same
 *     block z-fastest (src/magnelio/_fields/field_arrays.py:190-196).
 *   - K field sets ("nsets") are stored one after another with a caller-given
 *     element stride; material coefficients are shared by all sets.
 *
 * Reference paths below are relative to the reference repository root.
 */
#ifndef MGH_H
#define MGH_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library ------------------------------------------------------------ */

/* Thread-local message of the last failing call. */
const char* mgh_last_error(void);

/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int mgh_version(void);

/* Name, compute-unit count, HBM bytes and gfx arch of the current device. */
int mgh_device_info(char* name, int name_len, int* cu_count, long long* hbm_bytes, char* arch,
                    int arch_len);

/* ---- (a3) E phase --------------------------------------------------------
 * Replaces update_E_fused_cuda / update_E_fused / update_E_stencil
 * (src/magnelio/_operators/numba_kernels.py:518-536, :36-93, :137-171),
 * called from FITTimeDomainSolver.run e_phase_device
 * (src/magnelio/solver/fit_td.py:843-852).
 * E_s <- alpha_e * E_s + beta_e * (C^T H_s) for s = 0..nsets-1.
 * curl64 = 0: array-stencil operation order in the field dtype;
 * curl64 = 1 (float32 only): the Numba tier's float64 curl accumulator.
 * kchunk: z planes marched per thread (launch shape; results do not depend
 * on it). */
int mgh_update_e(int dtype, int nx, int ny, int nz, int nsets, void* e, long long e_set_stride,
                 const void* h, long long h_set_stride, const void* alpha_e, const void* beta_e,
                 int curl64, int kchunk, void* stream);

/* ---- (a4) H phase --------------------------------------------------------
 * Replaces update_H_fused_cuda / update_H_fused / update_H_stencil
 * (numba_kernels.py:538-552, :95-127, :174-205), called from
 * h_phase_device (fit_td.py:893-902).
 * H_s <- alpha_h * H_s - beta_h * (C E_s). */
int mgh_update_h(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                 long long e_set_stride, void* h, long long h_set_stride, const void* alpha_h,
                 const void* beta_h, int kchunk, void* stream);

/* ---- (a3)+(a7) / (a4)+(a7): curl phases with fused CPML -----------------
 * Same as mgh_update_e / mgh_update_h, plus the CPML auxiliary recursion of
 * every listed face fused into the boundary slabs.  Replaces
 * CPMLBoundary.update_E / update_H (src/magnelio/boundaries/cpml.py:343-516),
 * called right after the curl update in e_phase_device / h_phase_device
 * (src/magnelio/solver/fit_td.py:870-876, :918-921).
 * cpml_desc: host pointer to int64 words: [F, F records of 19 words]
 *   axis, k0, n, b, c, ck (device ptrs, n values in the field dtype),
 *   psi_e0, psi_e1, psi_h0, psi_h1 (device ptrs, field set 0),
 *   stride_e0, stride_e1, stride_h0, stride_h1 (elements between sets),
 *   air_e0, air_e1, air_h0, air_h1 (device uint8 0/1 masks or 0), spare.
 * Term components per normal axis: x (Ey,Ez | Hy,Hz), y (Ex,Ez | Hx,Hz),
 * z (Ex,Ey | Hx,Hy); psi slabs are x-fastest with the normal extent cut to
 * n layers.  Faces apply in record order.  cpml_desc = NULL: no CPML. */
int mgh_update_e_pml(int dtype, int nx, int ny, int nz, int nsets, void* e,
                     long long e_set_stride, const void* h, long long h_set_stride,
                     const void* alpha_e, const void* beta_e, int curl64, int kchunk,
                     const void* cpml_desc, void* stream);
int mgh_update_h_pml(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                     long long e_set_stride, void* h, long long h_set_stride,
                     const void* alpha_h, const void* beta_h, int kchunk, const void* cpml_desc,
                     void* stream);

/* ---- (a18) stored-energy sum ---------------------------------------------
 * Replaces the float64-accumulated reductions of the energy check
 * (src/magnelio/solver/fit_td.py:1014-1019).
 * out[s] = scale * sum_i (w[i]*x_s[i])*x_s[i]  (or += with accumulate != 0);
 * out and work are float64 device arrays; work holds
 * mgh_weighted_sq_sum_workspace(nsets) doubles. Deterministic. */
long long mgh_weighted_sq_sum_workspace(int nsets);
int mgh_weighted_sq_sum(int dtype, long long n, int nsets, const void* x, long long set_stride,
                        const void* w, double* out, int accumulate, double scale, double* work,
                        long long work_len, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MGH_H */
