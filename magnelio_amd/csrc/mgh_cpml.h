// CPML face descriptors shared by the E/H kernels.
//
// The host passes a flat int64 descriptor (plain words, no structs across the
// C-ABI):
//   desc[0] = number of faces F (0..6), then F records of MGH_CPML_WORDS words:
//     0 axis      normal axis (0/1/2)
//     1 k0, 2 n   first layer and number of layers along the normal
//     3 coef      device pointer: b[n] | c[n] | ck[n] in the field dtype
//     4 psi_e     device pointer: term 0 of every field set, then term 1
//                 ([K x S0 | K x S1], each slab x-fastest like the fields)
//     5 psi_h     same for the H side
//     6 air_e     device uint8 0/1 mask [S0 | S1] of the E terms, or 0 (all air)
//     7 air_h     same for the H side
//     8 K         field sets the psi buffers hold
// S_t is the slab size of term t's component: n times the component's
// extents along the two other axes.  Terms per normal axis (reference
// src/magnelio/boundaries/cpml.py:343-516):
//   x: E (Ey, Ez)  H (Hy, Hz)     y: E (Ex, Ez)  H (Hx, Hz)
//   z: E (Ex, Ey)  H (Hx, Hy)
// Faces apply in descriptor order (the reference's boundary dict order).
#pragma once

#include "mgh_common.h"

#define MGH_CPML_WORDS 9
#define MGH_CPML_MAX_FACES 6

namespace mgh {

struct CpmlFace {
  int axis, k0, n, nsets;
  const void* coef;
  void* psi_e;
  void* psi_h;
  const unsigned char* air_e;
  const unsigned char* air_h;
};

struct CpmlSet {
  int nfaces;
  CpmlFace f[MGH_CPML_MAX_FACES];
};

inline int load_cpml_set(const void* desc_v, int nx, int ny, int nz, int nsets, CpmlSet* out) {
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
    F.coef = (const void*)r[3];
    F.psi_e = (void*)r[4];
    F.psi_h = (void*)r[5];
    F.air_e = (const unsigned char*)r[6];
    F.air_h = (const unsigned char*)r[7];
    F.nsets = (int)r[8];
    MGH_REQUIRE(F.coef && F.psi_e && F.psi_h, "CPML face %d: null pointer", f);
    MGH_REQUIRE(F.nsets >= nsets, "CPML face %d holds %d field sets, launch needs %d", f, F.nsets,
                nsets);
    for (int g = 0; g < f; ++g) {
      const CpmlFace& G = out->f[g];
      MGH_REQUIRE(G.axis != F.axis || G.k0 + G.n <= F.k0 || F.k0 + F.n <= G.k0,
                  "CPML faces %d and %d overlap", g, f);
    }
  }
  int per_axis[3] = {0, 0, 0};
  for (int f = 0; f < nf; ++f) per_axis[out->f[f].axis]++;
  MGH_REQUIRE(per_axis[0] <= 2 && per_axis[1] <= 2 && per_axis[2] <= 2,
              "at most two CPML faces per axis");
  return MGH_OK;
}

}  // namespace mgh
