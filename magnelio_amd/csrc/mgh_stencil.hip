// Fused Yee curl + material update kernels for K field sets, with the CPML
// auxiliary recursion fused into the boundary slabs.
//
// Reference semantics (src/magnelio/_operators/numba_kernels.py):
//   E phase  E <- aE * E + bE * (C^T h)   update_E_stencil :137-171, update_E_fused :36-93
//   H phase  H <- aH * H - bH * (C e)     update_H_stencil :174-205, update_H_fused :95-127
// The dual curl of an E edge is guarded at the array bounds (a boundary edge
// has fewer than four H neighbours); the primal curl of an H face never is.
//
// Operation order follows the reference's array-stencil tier term by term:
//   curl_E = (((0 + t1) - t2) - t3) + t4  with absent terms skipped,
//   E      = (aE * E) + (bE * curl_E)
//   curl_H = ((t1 - t2) - t3) + t4,   H = (aH * H) - (bH * curl_H)
// and the library is compiled with -ffp-contract=off, so one step of this
// kernel is bit-identical to one step of that tier in the same precision.
// CURL64 selects the Numba tier's order instead: the E-side curl accumulates
// in double and E = (T)((double)(aE * E) + (double)bE * curl)  (:50, :59).
//
// CPML (src/magnelio/boundaries/cpml.py:343-516): after its curl update, an
// element inside the slab of a CPML face applies, face by face in the
// configured order,
//   psi = psi * b;  psi = psi + (c * d);  psi = psi * air;
//   F   = F +/- (beta * ((ck * d) + psi))
// where d is the one-sided difference of the neighbouring field along the
// face normal (the same two words the curl already read; the missing
// neighbour on a min-side slab is zero, cpml.py:375-379).  The reference runs
// these as separate array passes after the full curl update; fusing them
// changes no rounding, because the element's updated value is the register
// the pass would have reloaded and H (resp. E) is not written in the phase.
//
// Work decomposition (x-fastest device layout, see mgh_common.h): a 64 x 4
// thread tile covers (i, j) columns of the node lattice i in [0, Nx],
// j in [0, Ny]; each thread marches a chunk of z planes and keeps the
// z-neighbour values (H at k-1 in the E phase, E at k+1 in the H phase) in
// registers, so every field word is fetched from memory once per phase.  The
// in-plane neighbours (i-1, j-1 / i+1, j+1) are loads of words that adjacent
// lanes or the adjacent wave of the same work-group fetch in the same
// iteration.  One thread owns all three components at its node, so the
// shared H (or E) words are read once for the three updates, and the
// material coefficients are read once per node for all NS field sets.
#include "mgh_common.h"
#include "mgh_cpml.h"

namespace {

constexpr int TX = 64;
constexpr int TY = 4;

// Slab-local index of element (i, j, k) of a component with shape
// (n0, n1, n2) whose normal-axis extent is cut to the face's n layers.
__device__ __forceinline__ int64_t slab_index(const mgh::CpmlFace& f, int i, int j, int k, int n0,
                                              int n1) {
  if (f.axis == 0) return ((int64_t)k * n1 + j) * f.n + (i - f.k0);
  if (f.axis == 1) return ((int64_t)k * f.n + (j - f.k0)) * n0 + i;
  return ((int64_t)(k - f.k0) * n1 + j) * n0 + i;
}

template <typename T>
__device__ __forceinline__ T cpml_apply(const mgh::CpmlFace& f, int term, bool e_side,
                                        int64_t sidx, int s, int layer, T d, T beta, T val,
                                        int sign) {
  T* psi = (T*)(e_side ? f.psi_e[term] : f.psi_h[term]) +
           (int64_t)s * (e_side ? f.psi_e_stride[term] : f.psi_h_stride[term]);
  const unsigned char* air = e_side ? f.air_e[term] : f.air_h[term];
  const T b = ((const T*)f.b)[layer];
  const T c = ((const T*)f.c)[layer];
  const T ck = ((const T*)f.ck)[layer];
  T p = psi[sidx];
  p = p * b;
  p = p + (c * d);
  if (air != nullptr && air[sidx] == 0) p = p * T(0);
  psi[sidx] = p;
  const T corr = beta * ((ck * d) + p);
  return sign > 0 ? (val + corr) : (val - corr);
}

template <typename T, int NS, bool CURL64, bool PML>
__global__ __launch_bounds__(TX* TY) void k_update_e(T* __restrict__ e, const T* __restrict__ h,
                                                      const T* __restrict__ ae,
                                                      const T* __restrict__ be, int nx, int ny,
                                                      int nz, int kchunk, int64_t e_set,
                                                      int64_t h_set, int set0,
                                                      mgh::CpmlSet pml) {
  const int i = blockIdx.x * TX + threadIdx.x;
  const int j = blockIdx.y * TY + threadIdx.y;
  if (i > nx || j > ny) return;
  const int k0 = blockIdx.z * kchunk;
  const int k1 = min(k0 + kchunk, nz + 1);

  const mgh::Dims d{nx, ny, nz};
  const int64_t oEy = d.n_ex(), oEz = oEy + d.n_ey();
  const int64_t oHy = d.n_hx(), oHz = oHy + d.n_hy();
  // per-plane sizes (x-fastest: plane = k)
  const int64_t pEx = (int64_t)(ny + 1) * nx, pEy = (int64_t)ny * (nx + 1),
                pEz = (int64_t)(ny + 1) * (nx + 1);
  const int64_t pHx = (int64_t)ny * (nx + 1), pHy = (int64_t)(ny + 1) * nx,
                pHz = (int64_t)ny * nx;
  // in-plane offsets
  const int64_t rEx = (int64_t)j * nx + i;        // Ex (i<nx, j<=ny)
  const int64_t rEy = (int64_t)j * (nx + 1) + i;  // Ey (i<=nx, j<ny)
  const int64_t rEz = (int64_t)j * (nx + 1) + i;  // Ez (i<=nx, j<=ny)
  const int64_t rHx = (int64_t)j * (nx + 1) + i;  // Hx (i<=nx, j<ny)
  const int64_t rHy = (int64_t)j * nx + i;        // Hy (i<nx, j<=ny)
  const int64_t rHz = (int64_t)j * nx + i;        // Hz (i<nx, j<ny)

  const bool inx = i < nx, iny = j < ny;

  T hy_prev[NS], hx_prev[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    hy_prev[s] = T(0);
    hx_prev[s] = T(0);
    if (k0 > 0) {
      const T* hs = h + s * h_set;
      if (inx) hy_prev[s] = hs[oHy + (int64_t)(k0 - 1) * pHy + rHy];
      if (iny) hx_prev[s] = hs[(int64_t)(k0 - 1) * pHx + rHx];
    }
  }

  for (int k = k0; k < k1; ++k) {
    const bool knz = k < nz;
    const int64_t iEx = (int64_t)k * pEx + rEx;
    const int64_t iEy = oEy + (int64_t)k * pEy + rEy;
    const int64_t iEz = oEz + (int64_t)k * pEz + rEz;
    T a_x = T(0), b_x = T(0), a_y = T(0), b_y = T(0), a_z = T(0), b_z = T(0);
    if (inx) { a_x = ae[iEx]; b_x = be[iEx]; }
    if (iny) { a_y = ae[iEy]; b_y = be[iEy]; }
    if (knz) { a_z = ae[iEz]; b_z = be[iEz]; }

#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const T* hs = h + s * h_set;
      T* es = e + s * e_set;
      // H words at this node (zero where the face does not exist)
      const T hy = (inx && knz) ? hs[oHy + (int64_t)k * pHy + rHy] : T(0);
      const T hx = (iny && knz) ? hs[(int64_t)k * pHx + rHx] : T(0);
      const T hz = (inx && iny) ? hs[oHz + (int64_t)k * pHz + rHz] : T(0);

      if (inx) {  // Ex(i,j,k): dHz/dy - dHy/dz
        const T hz_jm = (j > 0) ? hs[oHz + (int64_t)k * pHz + rHz - nx] : T(0);
        T v;
        if constexpr (CURL64) {
          double c = 0.0;
          if (iny) c += (double)hz;
          if (j > 0) c -= (double)hz_jm;
          if (knz) c -= (double)hy;
          if (k > 0) c += (double)hy_prev[s];
          v = (T)((double)(a_x * es[iEx]) + (double)b_x * c);
        } else {
          T c = T(0);
          if (iny) c += hz;
          if (j > 0) c -= hz_jm;
          if (knz) c -= hy;
          if (k > 0) c += hy_prev[s];
          v = (a_x * es[iEx]) + (b_x * c);
        }
        if constexpr (PML) {
          for (int f = 0; f < pml.nfaces; ++f) {
            const mgh::CpmlFace& F = pml.f[f];
            if (F.axis == 2 && k >= F.k0 && k < F.k0 + F.n)
              v = cpml_apply<T>(F, 0, true, slab_index(F, i, j, k, nx, ny + 1), set0 + s, k - F.k0,
                                hy - hy_prev[s], b_x, v, +1);
            else if (F.axis == 1 && j >= F.k0 && j < F.k0 + F.n)
              v = cpml_apply<T>(F, 0, true, slab_index(F, i, j, k, nx, ny + 1), set0 + s, j - F.k0,
                                hz - hz_jm, b_x, v, -1);
          }
        }
        es[iEx] = v;
      }
      if (iny) {  // Ey(i,j,k): dHx/dz - dHz/dx
        const T hz_im = (i > 0) ? hs[oHz + (int64_t)k * pHz + rHz - 1] : T(0);
        T v;
        if constexpr (CURL64) {
          double c = 0.0;
          if (knz) c += (double)hx;
          if (k > 0) c -= (double)hx_prev[s];
          if (inx) c -= (double)hz;
          if (i > 0) c += (double)hz_im;
          v = (T)((double)(a_y * es[iEy]) + (double)b_y * c);
        } else {
          T c = T(0);
          if (knz) c += hx;
          if (k > 0) c -= hx_prev[s];
          if (inx) c -= hz;
          if (i > 0) c += hz_im;
          v = (a_y * es[iEy]) + (b_y * c);
        }
        if constexpr (PML) {
          for (int f = 0; f < pml.nfaces; ++f) {
            const mgh::CpmlFace& F = pml.f[f];
            if (F.axis == 2 && k >= F.k0 && k < F.k0 + F.n)
              v = cpml_apply<T>(F, 1, true, slab_index(F, i, j, k, nx + 1, ny), set0 + s, k - F.k0,
                                hx - hx_prev[s], b_y, v, -1);
            else if (F.axis == 0 && i >= F.k0 && i < F.k0 + F.n)
              v = cpml_apply<T>(F, 0, true, slab_index(F, i, j, k, nx + 1, ny), set0 + s, i - F.k0,
                                hz - hz_im, b_y, v, +1);
          }
        }
        es[iEy] = v;
      }
      if (knz) {  // Ez(i,j,k): dHy/dx - dHx/dy
        const T hy_im = (i > 0) ? hs[oHy + (int64_t)k * pHy + rHy - 1] : T(0);
        const T hx_jm = (j > 0) ? hs[(int64_t)k * pHx + rHx - (nx + 1)] : T(0);
        T v;
        if constexpr (CURL64) {
          double c = 0.0;
          if (inx) c += (double)hy;
          if (i > 0) c -= (double)hy_im;
          if (iny) c -= (double)hx;
          if (j > 0) c += (double)hx_jm;
          v = (T)((double)(a_z * es[iEz]) + (double)b_z * c);
        } else {
          T c = T(0);
          if (inx) c += hy;
          if (i > 0) c -= hy_im;
          if (iny) c -= hx;
          if (j > 0) c += hx_jm;
          v = (a_z * es[iEz]) + (b_z * c);
        }
        if constexpr (PML) {
          for (int f = 0; f < pml.nfaces; ++f) {
            const mgh::CpmlFace& F = pml.f[f];
            if (F.axis == 0 && i >= F.k0 && i < F.k0 + F.n)
              v = cpml_apply<T>(F, 1, true, slab_index(F, i, j, k, nx + 1, ny + 1), set0 + s,
                                i - F.k0, hy - hy_im, b_z, v, -1);
            else if (F.axis == 1 && j >= F.k0 && j < F.k0 + F.n)
              v = cpml_apply<T>(F, 1, true, slab_index(F, i, j, k, nx + 1, ny + 1), set0 + s,
                                j - F.k0, hx - hx_jm, b_z, v, +1);
          }
        }
        es[iEz] = v;
      }
      hy_prev[s] = hy;
      hx_prev[s] = hx;
    }
  }
}

template <typename T, int NS, bool PML>
__global__ __launch_bounds__(TX* TY) void k_update_h(const T* __restrict__ e, T* __restrict__ h,
                                                      const T* __restrict__ ah,
                                                      const T* __restrict__ bh, int nx, int ny,
                                                      int nz, int kchunk, int64_t e_set,
                                                      int64_t h_set, int set0,
                                                      mgh::CpmlSet pml) {
  const int i = blockIdx.x * TX + threadIdx.x;
  const int j = blockIdx.y * TY + threadIdx.y;
  if (i > nx || j > ny) return;
  const int k0 = blockIdx.z * kchunk;
  const int k1 = min(k0 + kchunk, nz + 1);

  const mgh::Dims d{nx, ny, nz};
  const int64_t oEy = d.n_ex(), oEz = oEy + d.n_ey();
  const int64_t oHy = d.n_hx(), oHz = oHy + d.n_hy();
  const int64_t pEx = (int64_t)(ny + 1) * nx, pEy = (int64_t)ny * (nx + 1),
                pEz = (int64_t)(ny + 1) * (nx + 1);
  const int64_t pHx = (int64_t)ny * (nx + 1), pHy = (int64_t)(ny + 1) * nx,
                pHz = (int64_t)ny * nx;
  const int64_t rEx = (int64_t)j * nx + i;
  const int64_t rEy = (int64_t)j * (nx + 1) + i;
  const int64_t rEz = (int64_t)j * (nx + 1) + i;
  const int64_t rHx = (int64_t)j * (nx + 1) + i;
  const int64_t rHy = (int64_t)j * nx + i;
  const int64_t rHz = (int64_t)j * nx + i;

  const bool inx = i < nx, iny = j < ny;

  // E at plane k (carried from the previous iteration's k+1 load)
  T ey_k[NS], ex_k[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const T* es = e + s * e_set;
    ey_k[s] = iny ? es[oEy + (int64_t)k0 * pEy + rEy] : T(0);
    ex_k[s] = inx ? es[(int64_t)k0 * pEx + rEx] : T(0);
  }

  for (int k = k0; k < k1; ++k) {
    const bool knz = k < nz;
    const int64_t iHx = (int64_t)k * pHx + rHx;
    const int64_t iHy = oHy + (int64_t)k * pHy + rHy;
    const int64_t iHz = oHz + (int64_t)k * pHz + rHz;
    const bool hasHx = iny && knz, hasHy = inx && knz, hasHz = inx && iny;
    T a_x = T(0), b_x = T(0), a_y = T(0), b_y = T(0), a_z = T(0), b_z = T(0);
    if (hasHx) { a_x = ah[iHx]; b_x = bh[iHx]; }
    if (hasHy) { a_y = ah[iHy]; b_y = bh[iHy]; }
    if (hasHz) { a_z = ah[iHz]; b_z = bh[iHz]; }

#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const T* es = e + s * e_set;
      T* hs = h + s * h_set;
      // E at plane k+1 (exists for k+1 <= nz)
      const T ey_kp = (iny && knz) ? es[oEy + (int64_t)(k + 1) * pEy + rEy] : T(0);
      const T ex_kp = (inx && knz) ? es[(int64_t)(k + 1) * pEx + rEx] : T(0);
      const T ez = knz ? es[oEz + (int64_t)k * pEz + rEz] : T(0);
      if (hasHx) {  // Hx(i,j,k): dEz/dy - dEy/dz
        const T ez_jp = es[oEz + (int64_t)k * pEz + rEz + (nx + 1)];
        const T c = ((ez_jp - ez) - ey_kp) + ey_k[s];
        T v = (a_x * hs[iHx]) - (b_x * c);
        if constexpr (PML) {
          for (int f = 0; f < pml.nfaces; ++f) {
            const mgh::CpmlFace& F = pml.f[f];
            if (F.axis == 2 && k >= F.k0 && k < F.k0 + F.n)
              v = cpml_apply<T>(F, 0, false, slab_index(F, i, j, k, nx + 1, ny), set0 + s, k - F.k0,
                                ey_kp - ey_k[s], b_x, v, -1);
            else if (F.axis == 1 && j >= F.k0 && j < F.k0 + F.n)
              v = cpml_apply<T>(F, 0, false, slab_index(F, i, j, k, nx + 1, ny), set0 + s, j - F.k0,
                                ez_jp - ez, b_x, v, +1);
          }
        }
        hs[iHx] = v;
      }
      if (hasHy) {  // Hy(i,j,k): dEx/dz - dEz/dx
        const T ez_ip = es[oEz + (int64_t)k * pEz + rEz + 1];
        const T c = ((ex_kp - ex_k[s]) - ez_ip) + ez;
        T v = (a_y * hs[iHy]) - (b_y * c);
        if constexpr (PML) {
          for (int f = 0; f < pml.nfaces; ++f) {
            const mgh::CpmlFace& F = pml.f[f];
            if (F.axis == 2 && k >= F.k0 && k < F.k0 + F.n)
              v = cpml_apply<T>(F, 1, false, slab_index(F, i, j, k, nx, ny + 1), set0 + s, k - F.k0,
                                ex_kp - ex_k[s], b_y, v, +1);
            else if (F.axis == 0 && i >= F.k0 && i < F.k0 + F.n)
              v = cpml_apply<T>(F, 0, false, slab_index(F, i, j, k, nx, ny + 1), set0 + s, i - F.k0,
                                ez_ip - ez, b_y, v, -1);
          }
        }
        hs[iHy] = v;
      }
      if (hasHz) {  // Hz(i,j,k): dEy/dx - dEx/dy
        const T ey_ip = es[oEy + (int64_t)k * pEy + rEy + 1];
        const T ex_jp = es[(int64_t)k * pEx + rEx + nx];
        const T c = ((ey_ip - ey_k[s]) - ex_jp) + ex_k[s];
        T v = (a_z * hs[iHz]) - (b_z * c);
        if constexpr (PML) {
          for (int f = 0; f < pml.nfaces; ++f) {
            const mgh::CpmlFace& F = pml.f[f];
            if (F.axis == 0 && i >= F.k0 && i < F.k0 + F.n)
              v = cpml_apply<T>(F, 1, false, slab_index(F, i, j, k, nx, ny), set0 + s, i - F.k0,
                                ey_ip - ey_k[s], b_z, v, +1);
            else if (F.axis == 1 && j >= F.k0 && j < F.k0 + F.n)
              v = cpml_apply<T>(F, 1, false, slab_index(F, i, j, k, nx, ny), set0 + s, j - F.k0,
                                ex_jp - ex_k[s], b_z, v, -1);
          }
        }
        hs[iHz] = v;
      }
      ey_k[s] = ey_kp;
      ex_k[s] = ex_kp;
    }
  }
}

template <typename T, int NS, bool C64, bool PML>
int launch_e(void* e, const void* h, const void* ae, const void* be, int nx, int ny, int nz,
             int kchunk, int64_t es, int64_t hs, int set0, const mgh::CpmlSet& pml,
             hipStream_t st) {
  dim3 blk(TX, TY, 1);
  dim3 grd((nx + 1 + TX - 1) / TX, (ny + 1 + TY - 1) / TY, (nz + 1 + kchunk - 1) / kchunk);
  hipLaunchKernelGGL((k_update_e<T, NS, C64, PML>), grd, blk, 0, st, (T*)e, (const T*)h,
                     (const T*)ae, (const T*)be, nx, ny, nz, kchunk, es, hs, set0, pml);
  MGH_CHECK_LAUNCH("k_update_e");
  return MGH_OK;
}

template <typename T, int NS, bool PML>
int launch_h(const void* e, void* h, const void* ah, const void* bh, int nx, int ny, int nz,
             int kchunk, int64_t es, int64_t hs, int set0, const mgh::CpmlSet& pml,
             hipStream_t st) {
  dim3 blk(TX, TY, 1);
  dim3 grd((nx + 1 + TX - 1) / TX, (ny + 1 + TY - 1) / TY, (nz + 1 + kchunk - 1) / kchunk);
  hipLaunchKernelGGL((k_update_h<T, NS, PML>), grd, blk, 0, st, (const T*)e, (T*)h, (const T*)ah,
                     (const T*)bh, nx, ny, nz, kchunk, es, hs, set0, pml);
  MGH_CHECK_LAUNCH("k_update_h");
  return MGH_OK;
}

#define MGH_NS_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8)

template <typename T, bool C64, bool PML>
int dispatch_e(int ns, void* e, const void* h, const void* ae, const void* be, int nx, int ny,
               int nz, int kc, int64_t es, int64_t hs, int set0, const mgh::CpmlSet& pml,
               hipStream_t st) {
  switch (ns) {
#define X(N) \
  case N: return launch_e<T, N, C64, PML>(e, h, ae, be, nx, ny, nz, kc, es, hs, set0, pml, st);
    MGH_NS_CASES(X)
#undef X
  }
  mgh::set_error("unsupported field-set group size %d", ns);
  return MGH_EINVAL;
}

template <typename T, bool PML>
int dispatch_h(int ns, const void* e, void* h, const void* ah, const void* bh, int nx, int ny,
               int nz, int kc, int64_t es, int64_t hs, int set0, const mgh::CpmlSet& pml,
               hipStream_t st) {
  switch (ns) {
#define X(N) \
  case N: return launch_h<T, N, PML>(e, h, ah, bh, nx, ny, nz, kc, es, hs, set0, pml, st);
    MGH_NS_CASES(X)
#undef X
  }
  mgh::set_error("unsupported field-set group size %d", ns);
  return MGH_EINVAL;
}

constexpr int kMaxGroup = 8;

int check_common(int dtype, int nx, int ny, int nz, int nsets, int kchunk) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 (float32) or 8 (float64), got %d",
              dtype);
  MGH_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "grid must have at least one cell per axis");
  MGH_REQUIRE(nsets >= 1, "nsets must be >= 1, got %d", nsets);
  MGH_REQUIRE(kchunk >= 1, "kchunk must be >= 1, got %d", kchunk);
  // Work-group grid limits: y dimension carries ceil((ny+1)/4) tiles.
  MGH_REQUIRE((int64_t)(ny + 1) / TY < 65535 && (int64_t)(nz + 1) / kchunk < 65535,
              "grid too large for the launch geometry");
  return MGH_OK;
}

}  // namespace

extern "C" {

int mgh_update_e_pml(int dtype, int nx, int ny, int nz, int nsets, void* e,
                     long long e_set_stride, const void* h, long long h_set_stride,
                     const void* alpha_e, const void* beta_e, int curl64, int kchunk,
                     const void* cpml, void* stream) {
  int rc = check_common(dtype, nx, ny, nz, nsets, kchunk);
  if (rc) return rc;
  MGH_REQUIRE(e && h && alpha_e && beta_e, "null device pointer");
  MGH_REQUIRE(!(curl64 && dtype == 8), "curl64 applies to float32 fields only");
  mgh::CpmlSet pml{};
  if (cpml) {
    rc = mgh::load_cpml_set(cpml, nx, ny, nz, &pml);
    if (rc) return rc;
  }
  const bool has_pml = pml.nfaces > 0;
  hipStream_t st = (hipStream_t)stream;
  const size_t w = (size_t)dtype;
  for (int s0 = 0; s0 < nsets; s0 += kMaxGroup) {
    const int ns = (nsets - s0) < kMaxGroup ? (nsets - s0) : kMaxGroup;
    char* eg = (char*)e + (size_t)s0 * (size_t)e_set_stride * w;
    const char* hg = (const char*)h + (size_t)s0 * (size_t)h_set_stride * w;
    if (dtype == 4) {
      if (curl64)
        rc = has_pml ? dispatch_e<float, true, true>(ns, eg, hg, alpha_e, beta_e, nx, ny, nz,
                                                     kchunk, e_set_stride, h_set_stride, s0, pml, st)
                     : dispatch_e<float, true, false>(ns, eg, hg, alpha_e, beta_e, nx, ny, nz,
                                                      kchunk, e_set_stride, h_set_stride, s0, pml,
                                                      st);
      else
        rc = has_pml ? dispatch_e<float, false, true>(ns, eg, hg, alpha_e, beta_e, nx, ny, nz,
                                                      kchunk, e_set_stride, h_set_stride, s0, pml,
                                                      st)
                     : dispatch_e<float, false, false>(ns, eg, hg, alpha_e, beta_e, nx, ny, nz,
                                                       kchunk, e_set_stride, h_set_stride, s0,
                                                       pml, st);
    } else {
      rc = has_pml ? dispatch_e<double, false, true>(ns, eg, hg, alpha_e, beta_e, nx, ny, nz,
                                                     kchunk, e_set_stride, h_set_stride, s0, pml,
                                                     st)
                   : dispatch_e<double, false, false>(ns, eg, hg, alpha_e, beta_e, nx, ny, nz,
                                                      kchunk, e_set_stride, h_set_stride, s0, pml,
                                                      st);
    }
    if (rc) return rc;
  }
  return MGH_OK;
}

int mgh_update_h_pml(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                     long long e_set_stride, void* h, long long h_set_stride,
                     const void* alpha_h, const void* beta_h, int kchunk, const void* cpml,
                     void* stream) {
  int rc = check_common(dtype, nx, ny, nz, nsets, kchunk);
  if (rc) return rc;
  MGH_REQUIRE(e && h && alpha_h && beta_h, "null device pointer");
  mgh::CpmlSet pml{};
  if (cpml) {
    rc = mgh::load_cpml_set(cpml, nx, ny, nz, &pml);
    if (rc) return rc;
  }
  const bool has_pml = pml.nfaces > 0;
  hipStream_t st = (hipStream_t)stream;
  const size_t w = (size_t)dtype;
  for (int s0 = 0; s0 < nsets; s0 += kMaxGroup) {
    const int ns = (nsets - s0) < kMaxGroup ? (nsets - s0) : kMaxGroup;
    const char* eg = (const char*)e + (size_t)s0 * (size_t)e_set_stride * w;
    char* hg = (char*)h + (size_t)s0 * (size_t)h_set_stride * w;
    if (dtype == 4)
      rc = has_pml ? dispatch_h<float, true>(ns, eg, hg, alpha_h, beta_h, nx, ny, nz, kchunk,
                                             e_set_stride, h_set_stride, s0, pml, st)
                   : dispatch_h<float, false>(ns, eg, hg, alpha_h, beta_h, nx, ny, nz, kchunk,
                                              e_set_stride, h_set_stride, s0, pml, st);
    else
      rc = has_pml ? dispatch_h<double, true>(ns, eg, hg, alpha_h, beta_h, nx, ny, nz, kchunk,
                                              e_set_stride, h_set_stride, s0, pml, st)
                   : dispatch_h<double, false>(ns, eg, hg, alpha_h, beta_h, nx, ny, nz, kchunk,
                                               e_set_stride, h_set_stride, s0, pml, st);
    if (rc) return rc;
  }
  return MGH_OK;
}

int mgh_update_e(int dtype, int nx, int ny, int nz, int nsets, void* e, long long e_set_stride,
                 const void* h, long long h_set_stride, const void* alpha_e, const void* beta_e,
                 int curl64, int kchunk, void* stream) {
  return mgh_update_e_pml(dtype, nx, ny, nz, nsets, e, e_set_stride, h, h_set_stride, alpha_e,
                          beta_e, curl64, kchunk, nullptr, stream);
}

int mgh_update_h(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                 long long e_set_stride, void* h, long long h_set_stride, const void* alpha_h,
                 const void* beta_h, int kchunk, void* stream) {
  return mgh_update_h_pml(dtype, nx, ny, nz, nsets, e, e_set_stride, h, h_set_stride, alpha_h,
                          beta_h, kchunk, nullptr, stream);
}

}  // extern "C"
