// Surface-impedance (SIBC) walls on the H update.
//
// Reference: SIBCOperator.save_field / update_field
// (src/magnelio/solver/_sibc.py:194-249), with W = g r_inst folded into
// M_sigma_m on the host (fit_td.py:419-429):
//   top of the step:  h_prev = h[idx]                 (mgh_ade_save, the same gather)
//   after the H curl and the mu(w) completion, wall block by wall block:
//                     hist = sum_p hw_p u_p           (first term seeds, hw = (1 + k)/2)
//                     h[idx] = h[idx] + beta (g hist)
//   then every block: h_mid = 0.5 (h[idx] + h_prev)
//                     u_p = u_p k_p;  u_p = u_p + q_p h_mid
// in the field precision T (hw, k, q stored as T: the reference's Python
// scalars take the state arrays' precision).  A face booked by two walls (a
// seam between two metals, or the edge where two boundary walls meet) takes
// both corrections in block order: the history pass runs one thread per
// distinct face over that face's entries, sorted by block; the advance pass
// one thread per entry, after every correction has landed.
#include "mgh_common.h"

namespace {

constexpr int SIBC_TAB = 4;  // per (block, branch): hw, k, q, unused

template <typename T>
__global__ __launch_bounds__(256) void k_sibc_hist(T* __restrict__ h, int64_t h_set, int nf,
                                                   const long long* __restrict__ face,
                                                   const int* __restrict__ ptr,
                                                   const int* __restrict__ ent,
                                                   const int* __restrict__ blk,
                                                   const T* __restrict__ g,
                                                   const T* __restrict__ beta,
                                                   const int* __restrict__ nbr,
                                                   const T* __restrict__ tab, int bmax,
                                                   const T* __restrict__ u, int ne) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int s = blockIdx.y;
  T* hs = h + (int64_t)s * h_set;
  const T* us = u + (int64_t)s * bmax * ne;
  T v = hs[face[f]];
  for (int x = ptr[f]; x < ptr[f + 1]; ++x) {
    const int e = ent[x];
    const int b = blk[e];
    const int nb = nbr[b];
    if (nb == 0) continue;
    const T* t = tab + (int64_t)b * bmax * SIBC_TAB;
    T hist = t[0] * us[e];
    for (int p = 1; p < nb; ++p) hist = hist + t[p * SIBC_TAB] * us[(int64_t)p * ne + e];
    v = v + beta[e] * (g[e] * hist);
  }
  hs[face[f]] = v;
}

template <typename T>
__global__ __launch_bounds__(256) void k_sibc_advance(const T* __restrict__ h, int64_t h_set, int ne,
                                                      const long long* __restrict__ idx,
                                                      const T* __restrict__ hprev,
                                                      const int* __restrict__ blk,
                                                      const int* __restrict__ nbr,
                                                      const T* __restrict__ tab, int bmax,
                                                      T* __restrict__ u) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int s = blockIdx.y;
  const int b = blk[e];
  const int nb = nbr[b];
  if (nb == 0) return;
  const T hmid = T(0.5) * (h[(int64_t)s * h_set + idx[e]] + hprev[(int64_t)s * ne + e]);
  T* us = u + (int64_t)s * bmax * ne;
  const T* t = tab + (int64_t)b * bmax * SIBC_TAB;
  for (int p = 0; p < nb; ++p) {
    T up = us[(int64_t)p * ne + e];
    up = up * t[p * SIBC_TAB + 1];
    up = up + t[p * SIBC_TAB + 2] * hmid;
    us[(int64_t)p * ne + e] = up;
  }
}

template <typename T>
void launch(int nsets, void* h, long long h_set, int nf, const long long* face, const int* ptr,
            const int* ent, int ne, const long long* idx, const int* blk, const void* g,
            const void* beta, const void* hprev, const int* nbr, const void* tab, int bmax, void* u,
            hipStream_t st) {
  if (nf > 0)
    hipLaunchKernelGGL(k_sibc_hist<T>, dim3((nf + 255) / 256, nsets, 1), dim3(256), 0, st, (T*)h,
                       (int64_t)h_set, nf, face, ptr, ent, blk, (const T*)g, (const T*)beta, nbr,
                       (const T*)tab, bmax, (const T*)u, ne);
  hipLaunchKernelGGL(k_sibc_advance<T>, dim3((ne + 255) / 256, nsets, 1), dim3(256), 0, st,
                     (const T*)h, (int64_t)h_set, ne, idx, (const T*)hprev, blk, nbr, (const T*)tab,
                     bmax, (T*)u);
}

}  // namespace

extern "C" {

// Entries: the wall blocks' faces concatenated in block order, ne of them
// (device-layout H indices idx, int64; block id blk, int32; g, beta_H and
// the stash hprev (nsets x ne) in the field dtype).  Distinct faces: nf of
// them (face, int64) with their entries ent[ptr[f] .. ptr[f + 1]) in block
// order (int32).  Ladder table tab: n_blocks x bmax x 4 field-dtype values
// (hw = (1 + k)/2, k, q, unused), nbr branches per block (int32).
// u: nsets x bmax x ne branch states.
int mgh_sibc_update(int dtype, int nsets, void* h, long long h_set_stride, int n_faces,
                    const long long* face_idx, const int* face_ptr, const int* face_entries,
                    int n_entries, const long long* entry_idx, const int* blk, const void* g,
                    const void* beta, const void* hprev, const int* nbr, const void* tab, int bmax,
                    void* u, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8");
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535 && n_faces >= 0 && n_entries >= 0 && bmax >= 1,
              "bad sizes");
  if (n_entries == 0) return MGH_OK;
  MGH_REQUIRE(n_faces >= 1 && n_faces <= n_entries, "n_faces must lie in [1, n_entries]");
  MGH_REQUIRE(h && face_idx && face_ptr && face_entries && entry_idx && blk && g && beta && hprev &&
                  nbr && tab && u,
              "null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    launch<float>(nsets, h, h_set_stride, n_faces, face_idx, face_ptr, face_entries, n_entries,
                  entry_idx, blk, g, beta, hprev, nbr, tab, bmax, u, st);
  else
    launch<double>(nsets, h, h_set_stride, n_faces, face_idx, face_ptr, face_entries, n_entries,
                   entry_idx, blk, g, beta, hprev, nbr, tab, bmax, u, st);
  MGH_CHECK_LAUNCH("k_sibc_hist / k_sibc_advance");
  return MGH_OK;
}

}  // extern "C"
