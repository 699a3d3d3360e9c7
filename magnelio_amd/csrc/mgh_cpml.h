// CPML face descriptors shared by the E/H kernels.
//
// The host passes a flat int64 descriptor (plain words, no structs across the
// C-ABI):
//   desc[0] = number of faces F (0..6), then F records of MGH_CPML_WORDS words:
//     axis, k0, n,                       normal axis (0/1/2), first layer, layers
//     b, c, ck,                          device pointers, n coefficients each
//     psi_e0, psi_e1, psi_h0, psi_h1,    device pointers to psi of field set 0
//     str_e0, str_e1, str_h0, str_h1,    element stride between field sets
//     air_e0, air_e1, air_h0, air_h1     device uint8 0/1 masks or 0 (all air)
// Terms per normal axis (reference src/magnelio/boundaries/cpml.py:343-516):
//   x: E (Ey, Ez)  H (Hy, Hz)     y: E (Ex, Ez)  H (Hx, Hz)
//   z: E (Ex, Ey)  H (Hx, Hy)
// Faces are applied in descriptor order (the reference's boundary dict order).
#pragma once

#include "mgh_common.h"

#define MGH_CPML_WORDS 19
#define MGH_CPML_MAX_FACES 6

namespace mgh {

struct CpmlFace {
  int axis, k0, n, pad;
  const void* b;
  const void* c;
  const void* ck;
  void* psi_e[2];
  void* psi_h[2];
  int64_t psi_e_stride[2];
  int64_t psi_h_stride[2];
  const unsigned char* air_e[2];
  const unsigned char* air_h[2];
};

struct CpmlSet {
  int nfaces;
  int pad;
  CpmlFace f[MGH_CPML_MAX_FACES];
};

inline int load_cpml_set(const void* desc_v, int nx, int ny, int nz, CpmlSet* out) {
  const long long* d = (const long long*)desc_v;
  const long long nf = d[0];
  MGH_REQUIRE(nf >= 0 && nf <= MGH_CPML_MAX_FACES, "CPML descriptor: %lld faces (max 6)", nf);
  out->nfaces = (int)nf;
  const int ncell[3] = {nx, ny, nz};
  for (int f = 0; f < nf; ++f) {
    const long long* r = d + 1 + (long long)f * MGH_CPML_WORDS;
    CpmlFace& F = out->f[f];
    F.axis = (int)r[0];
    F.k0 = (int)r[1];
    F.n = (int)r[2];
    MGH_REQUIRE(F.axis >= 0 && F.axis <= 2, "CPML face %d: bad axis %d", f, F.axis);
    MGH_REQUIRE(F.n >= 1 && F.k0 >= 0 && F.k0 + F.n <= ncell[F.axis],
                "CPML face %d: layers [%d, %d) outside 0..%d", f, F.k0, F.k0 + F.n,
                ncell[F.axis]);
    F.b = (const void*)r[3];
    F.c = (const void*)r[4];
    F.ck = (const void*)r[5];
    for (int t = 0; t < 2; ++t) {
      F.psi_e[t] = (void*)r[6 + t];
      F.psi_h[t] = (void*)r[8 + t];
      F.psi_e_stride[t] = r[10 + t];
      F.psi_h_stride[t] = r[12 + t];
      F.air_e[t] = (const unsigned char*)r[14 + t];
      F.air_h[t] = (const unsigned char*)r[16 + t];
      MGH_REQUIRE(F.psi_e[t] && F.psi_h[t], "CPML face %d: null psi pointer", f);
    }
    MGH_REQUIRE(F.b && F.c && F.ck, "CPML face %d: null coefficient pointer", f);
    (void)r[18];
  }
  return MGH_OK;
}

}  // namespace mgh
