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

/* HBM streaming probe (no counterpart in the reference: a measurement aid
 * for callers that report the curl launches against what the device can
 * stream).  Streams n_read input arrays and n_write output arrays of n floats
 * (in: the inputs back to back, out: the outputs back to back) the way a curl
 * launch streams its algorithmic words, with nothing else in the kernel:
 * out[w][i] = sum of in[r][i] over r = w mod n_write.  Mixes: 12 + 3 (the
 * E or H launch of one field set) and 42 + 18 (six sets). */
int mgh_stream_probe(int n_read, int n_write, const void* in, void* out, long long n,
                     void* stream);

/* ---- (a3) E phase --------------------------------------------------------
 * Replaces update_E_fused_cuda / update_E_fused / update_E_stencil
 * (src/magnelio/_operators/numba_kernels.py:518-536, :36-93, :137-171),
 * called from FITTimeDomainSolver.run e_phase_device
 * (src/magnelio/solver/fit_td.py:843-852).
 * E_s <- alpha_e * E_s + beta_e * (C^T H_s) for s = 0..nsets-1.
 * curl64 = 0: array-stencil operation order in the field dtype;
 * curl64 = 1 (float32 only): the Numba tier's float64 curl accumulator.
 * kchunk: the launch shape word (results do not depend on it): z planes
 * marched per thread in bits 0-15; optionally field sets per thread (1-3)
 * in bits 16-19 and the XCD band in launch units (1-255) in bits 20-27,
 * 0 leaving each to the library (sets: 3 when nsets is a multiple of 3,
 * else 2, 1 for one set; band: 8).  Every curl entry point below takes the
 * same word. */
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

/* ---- (a5) listed launches of the curl phases -------------------------------
 * Replace the live-block launches of update_E_fused_cuda / update_H_fused_cuda
 * (src/magnelio/_operators/numba_kernels.py:249-262, :459-474) driven by the
 * dead-tile plan (src/magnelio/solver/_tile_skip.py:187-256).  Same update as
 * mgh_update_e / mgh_update_h over the listed launch units only: a unit is
 * 256 consecutive nodes (i, j) of the x-fastest node plane (tile
 * tp = (j (nx+1) + i) / 256) times kchunk z planes (kc = k / kchunk), id
 * kc * ceil((nx+1)(ny+1)/256) + tp; live: device int32 ids (ascending),
 * n_live of them (0 launches nothing; live = NULL with n_live = 0 runs every
 * unit).  Skipped units must hold only elements whose update leaves them
 * unchanged (dead E edges already zero, no-op or lossless curl-dead H). */
int mgh_update_e_tiles(int dtype, int nx, int ny, int nz, int nsets, void* e,
                       long long e_set_stride, const void* h, long long h_set_stride,
                       const void* alpha_e, const void* beta_e, int curl64, int kchunk,
                       const int* live, int n_live, void* stream);
int mgh_update_h_tiles(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                       long long e_set_stride, void* h, long long h_set_stride,
                       const void* alpha_h, const void* beta_h, int kchunk, const int* live,
                       int n_live, void* stream);

/* ---- (a3)+(a7) / (a4)+(a7): curl phases with CPML ------------------------
 * Same as mgh_update_e / mgh_update_h, followed by the CPML auxiliary
 * recursion of every listed face over its slab, in record order (same values
 * as mgh_update_* then mgh_cpml_*, bit for bit): the curl launch applies the
 * faces' terms to each slab element before its store (mgh_update_*_pml_core;
 * its z planes split into interior-column and row-seam tiles so the x-slab
 * elements fill waves of their own); when the x slabs are not both at the
 * domain ends it leaves them to a launch over the x-slab columns
 * (mgh_cpml_xcols_*, a no-op otherwise).
 * Replaces CPMLBoundary.update_E /
 * update_H (src/magnelio/boundaries/cpml.py:343-516), called right after the
 * curl update in e_phase_device / h_phase_device
 * (src/magnelio/solver/fit_td.py:870-876, :918-921).
 * cpml_desc: host pointer to int64 words: [F, F records of 9 words]
 *   axis, k0, n,
 *   coef   (device ptr: b[n] | c[n] | ck[n] in the field dtype),
 *   psi_e, psi_h (device ptrs: term 0 of all K sets, then term 1),
 *   air_e, air_h (device uint8 0/1 masks [term 0 | term 1] or 0),
 *   K     (field sets held by the psi buffers, >= nsets).
 * Term components per normal axis: x (Ey,Ez | Hy,Hz), y (Ex,Ez | Hx,Hz),
 * z (Ex,Ey | Hx,Hy); psi slabs are x-fastest with the normal extent cut to
 * n layers.  At most two faces per axis, not overlapping.
 * cpml_desc = NULL: no CPML. */
int mgh_update_e_pml(int dtype, int nx, int ny, int nz, int nsets, void* e,
                     long long e_set_stride, const void* h, long long h_set_stride,
                     const void* alpha_e, const void* beta_e, int curl64, int kchunk,
                     const void* cpml_desc, void* stream);
int mgh_update_h_pml(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                     long long e_set_stride, void* h, long long h_set_stride,
                     const void* alpha_h, const void* beta_h, int kchunk, const void* cpml_desc,
                     void* stream);
/* The two launches of mgh_update_e_pml / mgh_update_h_pml, callable apart
 * (to time the curl launch alone): _pml_core, then _xcols on the same
 * stream, equals the composed call. */
int mgh_update_e_pml_core(int dtype, int nx, int ny, int nz, int nsets, void* e,
                          long long e_set_stride, const void* h, long long h_set_stride,
                          const void* alpha_e, const void* beta_e, int curl64, int kchunk,
                          const void* cpml_desc, void* stream);
int mgh_update_h_pml_core(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                          long long e_set_stride, void* h, long long h_set_stride,
                          const void* alpha_h, const void* beta_h, int kchunk,
                          const void* cpml_desc, void* stream);
int mgh_cpml_xcols_e(int dtype, int nx, int ny, int nz, int nsets, void* e,
                     long long e_set_stride, const void* h, long long h_set_stride,
                     const void* beta_e, const void* cpml_desc, void* stream);
int mgh_cpml_xcols_h(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                     long long e_set_stride, void* h, long long h_set_stride,
                     const void* beta_h, const void* cpml_desc, void* stream);

/* ---- (a7) CPML slab passes alone -------------------------------------------
 * The CPML part of mgh_update_e_pml / mgh_update_h_pml: for each face in
 * record order, psi and the tangential components of its slab are updated
 * from the fields as they stand (call right after mgh_update_e / _h).
 * Replaces CPMLBoundary.update_E / update_H (cpml.py:343-516). */
int mgh_cpml_e(int dtype, int nx, int ny, int nz, int nsets, void* e, long long e_set_stride,
               const void* h, long long h_set_stride, const void* beta_e, const void* cpml_desc,
               void* stream);
int mgh_cpml_h(int dtype, int nx, int ny, int nz, int nsets, const void* e,
               long long e_set_stride, void* h, long long h_set_stride, const void* beta_h,
               const void* cpml_desc, void* stream);

/* ---- (a8)-(a11) modal port step ------------------------------------------
 * Replaces PortOperatorModal.update_e + project_V/project_I
 * (src/magnelio/ports/_modal/operator.py:1351-1536, :1209-1321),
 * DTBCTermination.advance (src/magnelio/ports/_modal/dtbc.py:417-463) and
 * PortSignalRecorder.record (src/magnelio/ports/recorder.py:216-250), called
 * after the E phase in FITTimeDomainSolver.run (fit_td.py:975-982).
 * Float64 projections, history convolutions and recursion in four
 * launches over (plane chunk, port, field set); the step index n is read
 * from *step_ctr.  With the complement absorber (operator.py:1395-1427,
 * :1499-1522, live while a channel of the port runs on Mur) the write adds,
 * per plane edge, live (c_int_prev + r (c - c_port_prev)) of the interior
 * remainder c = e_int - sum_m V_int,m p_m.
 * port_desc: n_ports records of 32 int64 words (see csrc/mgh_ports.h):
 *   0 n_u, 1 n_v, 2 M, 3-5 idx_port/idx_int/idx_h (int64 device-layout
 *   indices), 6-8 me_port/me_int/mh, 9-10 prof_e/prof_h (M x n),
 *   11 params (M x 8: kind, r^2, q^2, mur_r, record_scale, 2-2r^2-q^2),
 *   12-14 kern/kern_stride/k_init (packed DTBC kernels), 15 state
 *   (nsets x M x 4), 16-18 w_hist/s_hist/cap, 19 exc_mode (int32 nsets),
 *   20-21 src_tab/tab_len, 22 mur_tab, 23-25 V_rec/I_rec/rec_cap,
 *   26 absmax, 27-28 work/work_len (scratch), 29-31 comp_r/comp_live (n)
 *   and comp_state (nsets x 2 x n) of the complement absorber, or 0. */
int mgh_port_step(int dtype, int nsets, void* e, long long e_set_stride, const void* h,
                  long long h_set_stride, const long long* port_desc, int n_ports,
                  const long long* step_ctr, void* stream);

/* ---- step counter ----------------------------------------------------------
 * *step_ctr += by on the device (keeps per-step launches graph-capturable).
 * No reference counterpart: the reference's loop index n lives on the host
 * (fit_td.py:950). */
int mgh_step_advance(long long* step_ctr, long long by, void* stream);

/* ---- packed small jobs ----------------------------------------------------
 * No reference counterpart (the reference runs one solver at a time; its
 * unit of work is AnalysisScatteringTD.run, scattering_td.py:1252-1330): the
 * per-step launches of several small solvers, as a serving loop holds them,
 * in one launch each.  Results equal each solver stepped alone, bit for bit.
 *
 * mgh_update_e_packed / mgh_update_h_packed: the E / H curl of n_jobs
 * (1..16) grids, plain (no CPML, dense units), one field set per thread;
 * jobs = n_jobs x 10 int64 words: nx, ny, nz, nsets, E, E set stride, H,
 * H set stride, alpha and beta of the phase's side; shape: the launch shape
 * word of mgh_update_e with a sets-per-thread part of 0 or 1. */
int mgh_update_e_packed(int dtype, int n_jobs, const long long* jobs, int shape, void* stream);
int mgh_update_h_packed(int dtype, int n_jobs, const long long* jobs, int shape, void* stream);

/* mgh_port_step_packed: mgh_port_step for ports of several solvers (1..8
 * ports in all): port_desc as for mgh_port_step, jobs = n_ports x 6 int64
 * words: E, E set stride, H, H set stride, field sets and step counter of
 * the solver each port belongs to. */
int mgh_port_step_packed(int dtype, int n_ports, const long long* port_desc,
                         const long long* jobs, void* stream);

/* mgh_step_advance_packed: *ctrs[i] += by for n (1..64) step counters. */
int mgh_step_advance_packed(int n, const long long* ctrs, long long by, void* stream);

/* ---- (a2)+(a6) material matrices and coefficients on the device ----------
 * Replaces build_M_eps / build_M_mu / build_M_sigma / build_M_sigma_m
 * (src/magnelio/_operators/material_matrices.py:334-757, category-0 cells),
 * the port-plane flattening (:106-331), the PEC OR rule with bbox walls
 * (src/magnelio/mesh/indexing.py:168-257) and the alpha/beta construction
 * of FITTimeDomainSolver.setup (src/magnelio/solver/fit_td.py:375-513),
 * in float64 in the same operation order, rounded once to dtype.
 * mat_id: device int32 (nx, ny, nz), z fastest (numpy C order); tab: device
 * float64 n_mat x 16 (eps xyz, mu xyz, sigma xyz, sigma_m xyz with 0 for PEC
 * materials, is_pec, 3 unused); geo: device float64 dx[nx] dy[ny] dz[nz]
 * then dual widths ax[nx+1] ay[ny+1] az[nz+1]; face bits xmin 1, xmax 2,
 * ymin 4, ymax 8, zmin 16, zmax 32 for PEC walls and port faces (on any
 * axes).  Outputs in the device layout: alpha/beta/M_eps on the E
 * side, alpha/beta/M_mu on the H side (m_eps / m_mu may be NULL).  No ADE
 * W term (dispersive meshes use the host construction). */
int mgh_build_coefficients(int dtype, int nx, int ny, int nz, const int* mat_id, int n_mat,
                           const double* tab, const double* geo, double dt, double eps0, double mu0,
                           int wall_faces, int port_faces, void* alpha_e, void* beta_e, void* m_eps,
                           void* alpha_h, void* beta_h, void* m_mu, void* stream);

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

/* mgh_weighted_sq_sum_noalloc: mgh_weighted_sq_sum without a workspace
 * argument and without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial sums live in a static device array of the library (512 KiB per
 * GPU, held until the process ends), so nsets must be <= 64 (error
 * otherwise).  All calls on one GPU share it: calls that overlap on one GPU
 * race. */
int mgh_weighted_sq_sum_noalloc(int dtype, long long n, int nsets, const void* x,
                                long long set_stride, const void* w, double* out, int accumulate,
                                double scale, void* stream);

/* ---- (a12) lumped RLC elements and discrete ports -----------------------
 * Replaces LumpedElementOperator.update_e (src/magnelio/ports/_lumped/
 * operator.py:134-166) with SeriesRLC / ParallelRLC companions
 * (src/magnelio/circuit/companion.py:45-180), run in the port loop after the
 * E phase (fit_td.py:975-976), and the recorder's project_V / project_I for
 * lumped ports (recorder.py:216-250).  One work-group per (element, set):
 *   v_total = sum sign*e (float64, chain order)
 *   i = ((v_src - v_hist) - v_total) / (r_eq + sum beta)
 *   e += (sign*beta)*i;  companion advance;  V = sum sign*e, I = i.
 * desc: host array of n_elems records of 12 int64 words:
 *   0 n_edges (<= 1024), 1 idx (int64 device-layout E indices), 2 sign (f64),
 *   3 beta (f64), 4 params (f64[8]: parallel, has_L, has_C, r_eq, sum beta,
 *   kL, kC, G), 5 state (f64 nsets x 5), 6 exc (int32 per set), 7 src (f64
 *   nsets x tab_len: v_src of step n), 8 tab_len, 9 V_rec, 10 I_rec (f64
 *   nsets x cap, or 0), 11 cap.  Edge chains of different elements must be
 *   disjoint. */
int mgh_lumped_step(int dtype, int nsets, void* e, long long e_set_stride, const long long* desc,
                    int n_elems, const long long* step_ctr, void* stream);

/* ---- (a14) TF/SF plane-wave corrections ---------------------------------
 * Replaces PlaneWaveSource.inject_E / inject_H / _apply
 * (src/magnelio/sources/plane_wave.py:710-734), called after the E phase and
 * after the H phase (fit_td.py:965-967, :991-993).  The face patches
 * (plane_wave.py:318-707, coef = beta * (factor * metric)) arrive grouped by
 * target element: targets[t] (int64 device-layout index) owns terms
 * [ptr[t], ptr[t+1]) (int32), applied in order:
 *   f[set][targets[t]] += coef[q] * wave[n * n_delays + did[q]]
 * with wave a tab_len x n_delays table of waveform samples in the field
 * dtype and n = *step_ctr + step_offset (no injection outside the table). */
int mgh_tfsf_apply(int dtype, int nsets, void* field, long long set_stride, int n_targets,
                   const long long* targets, const int* ptr, const void* coef, const int* did,
                   const void* wave, int n_delays, long long tab_len, const long long* step_ctr,
                   long long step_offset, void* stream);

/* ---- (a15) ADE pole-residue dielectrics, E side -------------------------
 * Replace DispersionOperator.save_field / update_field
 * (src/magnelio/solver/_dispersion.py:504-569, NumPy branch), called at the
 * top of the step and right after the E curl (fit_td.py:833-834, :858-859);
 * the coefficient W is folded into alpha_E / beta_E on the host
 * (fit_td.py:382-403).  n dispersive E edges idx (int64, device layout,
 * disjoint), g = EPS0 A_dual/L and beta_E at those edges (field dtype), the
 * edge's block (material) blk (int32).  fprev: nsets x n.
 * The same two calls serve the mu(w) mirror (the H side of the same
 * operator, _dispersion.py:45-76; fit_td.py:835-836, :906-907): e is then
 * the H vector, idx its dispersive faces, g = MU0 A_primal/L_dual and beta
 * beta_H, the stash at the top of the step and the update right after the
 * H curl (before the CPML H corrections).
 *   save:   fprev = e[idx]
 *   update: j = sum_q hw_q Re((1 + k_q) J_q) (first term seeds);
 *           e[idx] -= beta j;  gd = g (e[idx] - fprev);
 *           J_q = J_q k_q;  J_q += c_q gd
 * complex a*b rounded as NumPy's SIMD loop on FMA hosts:
 *   (fma(a.re, b.re, -(a.im b.im)), fma(a.re, b.im, a.im b.re)).
 * tab: n_blocks x pmax x 8 field-dtype values per pole (kind 0 real /
 * 1 conjugate pair, hw = weight/2, 1+k, k, c as (re, im) pairs); npoles per
 * block (int32); J: nsets x pmax x n complex states, (re, im) interleaved. */
int mgh_ade_save(int dtype, int nsets, const void* e, long long e_set_stride, int n,
                 const long long* idx, void* fprev, void* stream);
int mgh_ade_update(int dtype, int nsets, void* e, long long e_set_stride, int n,
                   const long long* idx, const void* g, const void* beta, const void* fprev,
                   const int* blk, const int* npoles, const void* tab, int pmax, void* J,
                   void* stream);

/* ---- Far field: near-to-far-field transform of a Huygens patch set --------
 * Replaces the radiation-vector sums and projections of ntff_transform
 * (src/magnelio/post/far_field.py:139-271); the host builds the patch set
 * (the monitor's faces divided by the source spectrum, doubled by each image
 * plane, monitors/far_field.py:325-355, post/far_field.py:100-122) and the
 * direction grid.  patches: n_patches x 15 doubles (centre xyz, area (n x H)
 * xyz and -area (n x E) xyz as (re, im) pairs, library e^{-jwt} phasors);
 * dirs: n_dirs x 9 doubles (r, theta, phi unit vectors); FP64 throughout:
 *   N = sum_p e^{-jk r.c_p} wJ_p,  L = sum_p e^{-jk r.c_p} wM_p
 *   E_theta = pre (eta N.theta + L.phi),  E_phi = pre (eta N.phi - L.theta)
 * with pre = +jk/(4 pi) — the reference's conjugate-in / conjugate-out
 * amplitudes.  The patches split into n_chunks ranges whose partial sums
 * (work: n_chunks x 12 x n_dirs doubles) are added in chunk order.
 * out: n_dirs x 4 doubles (E_theta re, im, E_phi re, im). */
int mgh_ntff(int n_patches, const double* patches, int n_dirs, const double* dirs, double k,
             double pre_re, double pre_im, double eta, int n_chunks, double* work, double* out,
             void* stream);
/* The same transform for tensor faces — a Huygens-box face or its image: the
 * grid c1 x c2 on the plane x_axis = planes[f] (face_desc: n_faces x 6 int64:
 * axis, n1, n2, offsets of c1 and c2 into coords, offset of W), whose phase
 * factors into e^{-jk r_a p} e^{-jk r_1 c1_i} e^{-jk r_2 c2_j} (t1 < t2 the
 * other two axes).  W: per face n1 x n2 x 4 (re, im) pairs: area (n x H)_t1,
 * area (n x H)_t2, -area (n x E)_t1, -area (n x E)_t2.  jobs: n_jobs x 3 int32
 * (face, first column, columns <= 16), each column of each face in one job;
 * work: n_jobs x 8 x n_dirs doubles, added in job order; out as above. */
int mgh_ntff_faces(int n_faces, const long long* face_desc, const double* planes,
                   const double* coords, const double* W, int n_jobs, const int* jobs, int n_dirs,
                   const double* dirs, double k, double pre_re, double pre_im, double eta,
                   double* work, double* out, void* stream);

/* ---- SIBC walls: lossy walls in the time-domain H update ------------------
 * Replaces SIBCOperator.update_field (src/magnelio/solver/_sibc.py:213-249),
 * called right after the H curl and the mu(w) completion, before the CPML H
 * corrections (fit_td.py:909-916); the stash at the top of the step
 * (save_field, :194-199, fit_td.py:839-840) is mgh_ade_save over the entry
 * indices, and W = g r_inst is folded into M_sigma_m on the host
 * (fit_td.py:419-429).  Entries: the wall blocks' faces concatenated in
 * block order (n_entries; device-layout H indices entry_idx int64, block id
 * blk int32; g = A/l_dual^2, beta_H and hprev (nsets x n_entries) in the
 * field dtype).  Distinct faces face_idx (n_faces, int64) with their
 * entries face_entries[face_ptr[f] .. face_ptr[f + 1]) in block order
 * (int32), so a face booked by two walls takes both corrections in the
 * reference's order:
 *   h[f] = h[f] + beta_e (g_e sum_p hw_p u_p,e)     for each entry e of f in order
 *   h_mid = 0.5 (h[idx_e] + hprev_e);  u_p,e = u_p,e k_p;  u_p,e += q_p h_mid
 * tab: n_blocks x bmax x 4 field-dtype values (hw = (1 + k)/2, k, q, unused),
 * nbr branches per block (int32); u: nsets x bmax x n_entries. */
int mgh_sibc_update(int dtype, int nsets, void* h, long long h_set_stride, int n_faces,
                    const long long* face_idx, const int* face_ptr, const int* face_entries,
                    int n_entries, const long long* entry_idx, const int* blk, const void* g,
                    const void* beta, const void* hprev, const int* nbr, const void* tab, int bmax,
                    void* u, void* stream);

/* ---- (a16) DFT monitors ----------------------------------------------------
 * Replace DFTAccumulator.accumulate (src/magnelio/monitors/_dft.py:37-52) fed
 * with cell-centred samples (src/magnelio/monitors/base.py:51-173) from the
 * record hooks of MonitorFieldFrequency (field_frequency.py:249-294) and
 * MonitorFarField (far_field.py:230-248); E stamped t = n dt, H t + dt/2.
 *
 * mgh_dft_sample: row ((n - n_start) mod B) of X (nsets x B x S_pad float32,
 * B <= 512) receives — with x_buf_stride > 0 X holds two such buffers that
 * many floats apart and the row goes to buffer ((n - n_start) div B) mod 2 — for each of S samples, sum_t w[s,t] * field[idx[s,t]] over 8
 * terms; idx >= 0 indexes E, idx < 0 indexes H at -(idx + 1) (device
 * layout); weight 0 marks an unused term. n is read from *step_ctr. */
int mgh_dft_sample(int dtype, int nsets, const void* e, long long e_set_stride, const void* h,
                   long long h_set_stride, const int* idx, const float* w, int S, float* X,
                   int S_pad, int B, long long x_buf_stride, const long long* step_ctr,
                   long long n_start, void* stream);

/* mgh_dft_sample_groups: the same row of X from structured groups — the
 * samples of one monitor component on one face or box, whose terms are affine
 * in the sample's natural coordinates (c0 fastest): term t reads
 * field[base_t + c0 s0 + c1 s1 + c2 s2] with weight wtab[woff_t + c_{waxis_t}]
 * (waxis_t < 0: a constant).  grp: G descriptors of 40 int32 words sorted by
 * start — [0] start [1] count [2] n0 [3] n1 [4] field (0 E, 1 H) [5] terms
 * [6..8] strides [9..16] bases [17..24] weight coordinates [25..32] weight
 * offsets [33] first work-group [34] mapping (0: 256 consecutive samples per
 * work-group, 1: 16 x 16 patches of (c0, c1)) [35] work-groups; n_blocks is
 * their total.  Built by the host from the mgh_dft_sample map and checked
 * against it; the values are those of mgh_dft_sample bit for bit.  Samples
 * outside every group (padding) are not written. */
int mgh_dft_sample_groups(int dtype, int nsets, const void* e, long long e_set_stride,
                          const void* h, long long h_set_stride, const int* grp, int G,
                          int n_blocks, const float* wtab, int S, float* X, int S_pad, int B,
                          long long x_buf_stride, const long long* step_ctr, long long n_start,
                          void* stream);

/* mgh_dft_block: bins[set, f, s] += exp(j w_f t0) * sum_{b < nb_valid}
 * (Pc + j Ps)[f, b] * X[set, b, s], t0 = n0 dt for s < s_h_start and
 * n0 dt + dt/2 otherwise, n0 the first step of the block holding *step_ctr.
 * Pc/Ps: F_pad x Bp float32 tables cos/sin(w_f b dt) dt in the kernel's lane
 * order (Bp: B rounded up to 64 steps up to 256, 512 beyond). bins: nsets x F x S
 * complex float64 (re/im interleaved); lo (same shape in float32, or NULL)
 * holds the TwoSum compensation terms. blk: F x 4 float64 scratch. Contraction on the MFMA
 * units (v_mfma_f32_16x16x4f32). */
int mgh_dft_block(int nsets, const float* X, int S, int S_pad, int B, int nb_valid,
                  const float* Pc, const float* Ps, const double* omega, int F, int F_pad,
                  double dt, double* blk, int s_h_start, double* bins, float* lo,
                  const long long* step_ctr, long long n_start, void* stream);

/* mgh_dft_blocks: mgh_dft_block over nsub (1, 2 or 4) consecutive B-step
 * blocks in one launch — X: nsets x nsub*B x S_pad, rows q*B .. q*B + B - 1
 * are block q (first nb_valid rows valid), blk: nsub x F x 4 float64
 * scratch; each block is its own contraction against the same Pc/Ps with its
 * own exp(j w_f t0_q), the float64 contributions are summed and the bins read
 * and written once (nsub = 1 is mgh_dft_block).  Requires Bp * nsub <= 256.
 * Replaces the same DFTAccumulator.accumulate steps (monitors/_dft.py:37-52)
 * as mgh_dft_block; the reference accumulates every step. */
int mgh_dft_blocks(int nsets, const float* X, int S, int S_pad, int B, int nsub, int nb_valid,
                   const float* Pc, const float* Ps, const double* omega, int F, int F_pad,
                   double dt, double* blk, int s_h_start, double* bins, float* lo,
                   const long long* step_ctr, long long n_start, void* stream);

/* mgh_dft_blocks_range: mgh_dft_blocks over the sample waves [wave_lo,
 * wave_hi) (16 samples each, W = ceil(S / 16) in all) with the first step
 * of block 0 given as n0 instead of read from a counter, frequencies split
 * over fsplit groups of work-groups.  Flushes what a pipelined monitor has
 * not contracted when a run ends (the same accumulate steps). */
int mgh_dft_blocks_range(int nsets, const float* X, int S, int S_pad, int B, int nsub,
                         int nb_valid, int wave_lo, int wave_hi, long long n0, const float* Pc,
                         const float* Ps, const double* omega, int F, int F_pad, double dt,
                         double* blk, int s_h_start, double* bins, float* lo, int fsplit,
                         void* stream);

/* mgh_dft_phase_table: the block phases exp(j w t0) (E and H stamps) of
 * n_periods consecutive periods of nsub B steps from step n0, nsub x F x 4
 * float64 per period — the table mgh_dft_slice reads. */
int mgh_dft_phase_table(const double* omega, int F, double dt, long long n0, int B, int nsub,
                        int n_periods, double* blk, void* stream);

/* mgh_dft_slice: one slice of a pipelined contraction.  X holds two sample
 * buffers (x_buf_stride floats apart, each nsets x P x S_pad, P = nsub B);
 * period p (P steps from n_start) is sampled into buffer p mod 2.  Called at
 * the steps n (*step_ctr) with ((n - n_start) mod P + 1) mod cad = 0 (cad
 * divides P), it contracts slice j = ((n - n_start) mod P) div cad of the
 * P / cad slices of period p - 1 — waves [j W / (P/cad), (j + 1) W /
 * (P/cad)), W = ceil(S / 16) — over all P rows of buffer (p - 1) mod 2 with
 * that period's phases (phase_table from mgh_dft_phase_table from n_start,
 * n_periods periods).  The P steps of a period contract every wave of the
 * previous one once, so K steps carry K / P of a period's contraction
 * work.  Nothing is contracted in period 0.  Replaces the same
 * DFTAccumulator.accumulate steps (monitors/_dft.py:37-52). */
int mgh_dft_slice(int nsets, const float* X, long long x_buf_stride, int S, int S_pad, int B,
                  int nsub, const float* Pc, const float* Ps, int F, int F_pad,
                  const double* phase_table, int n_periods, int s_h_start, double* bins, float* lo,
                  const long long* step_ctr, long long n_start, int cad, int fsplit,
                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MGH_H */
