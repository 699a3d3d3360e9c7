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
// Work decomposition.  A thread owns one (i, j) node column of the node
// lattice i in [0, Nx], j in [0, Ny] and marches a chunk of z planes, keeping
// the z-neighbour values (H at k-1 in the E phase, E at k+1 in the H phase)
// in registers; it updates all three components at its node, so shared words
// are read once for the three updates.  Nodes are numbered x-fastest, 256 to
// a work-group (four 64-wide waves over consecutive nodes), which keeps every
// wave's loads on contiguous rows and leaves no partly empty tile at i = Nx.
// With two or more field sets a thread advances two sets (three when K is a
// multiple of 3), so the coefficient words serve all of them, and bands of
// consecutive units — the set groups of one tile first, then the next tiles —
// run on one XCD and share its L2 (BAND below).
//
// CPML, two forms with the same values.  With the curl (the
// mgh_update_*_pml entry points): every d of every face is a difference of
// neighbour words the curl already holds in registers, so an element in a
// slab applies its faces' terms, in descriptor order, to its updated value
// before the single store; only psi, the layer coefficients and the air mask
// add traffic.  One "node in some slab" test per element keeps the interior
// on the plain path, and the face loop is not unrolled (its per-face values
// are uniform scalar loads), which holds the register count of the plain
// kernel.  The x-slab elements would be a few lanes of half the waves of a
// row-tiled launch, and their recursion stalls the whole wave (measured +25
// us per phase on 256^3), so such a launch splits its planes (Plan) and the
// x-slab elements fill waves of their own; when the x slabs are not both at
// the domain ends the launch leaves them to a pass over the x-slab columns
// (k_cpml_xcols_e/h), which applies all their faces in order after the curl.
// Separate (mgh_cpml_e / mgh_cpml_h):
// one small launch per face over its slab, in descriptor order — the
// structure of the reference (cpml.update_E / update_H per face after the
// curl pass) — for listed launches and for dispersive meshes, whose ADE
// completion sits between the curl and the CPML pass.
#include "mgh_common.h"
#include "mgh_cpml.h"

#include <algorithm>

namespace {

constexpr int WG = 256;

// Units of WG nodes x kchunk planes x one set group, ordered set group
// fastest, then node tile (one row of the node plane, about), then z chunk.
//
// Split planes (launches that carry the CPML of two x slabs at the domain
// ends): tiles [0, tpa) cover the interior columns [a0, a1) of every row,
// tiles [tpa, ntp) the row seams — row j - 1's columns [a1, nx] followed by
// row j's columns [0, a0), adjacent in memory — so every wave holds only
// interior elements or only x-slab (and end-column) elements.
struct Plan {
  int nx, ny, nz;
  int kchunk, ngroups;
  int ntp, nkc, units;
  int split, a0, a1, tpa, wb;
  int band;         // units per XCD band
  const int* live;  // live spatial units (kc * ntp + tp), or null for all
};

// Work-groups are dealt to the 8 XCDs round robin (work-group b to XCD
// b mod 8).  Units go to XCDs in bands of B consecutive ones (B = 8 unless
// the launch shape asks otherwise): XCD x runs units [8 B s + B x,
// 8 B s + B x + B) for s = 0, 1, ..., so consecutive node tiles (and the set
// groups of one tile) share an XCD's L2 — the previous row a tile reads as
// its y neighbour was just fetched there — while all XCDs sweep the same z
// region together.  On 256^3 (K = 1) the curl launches ran 2-4 % faster with
// bands of 8 than with units dealt one by one, and faster than a contiguous
// unit range per XCD, which spreads the XCDs over eight z regions.
constexpr int BAND = 8;

__device__ __forceinline__ bool unit_of(const Plan& p, int& grp, int& i, int& j, int& k0,
                                        int& k1) {
  const int blk = blockIdx.x;
  const int sq = blk >> 3;  // this work-group's place in its XCD's sequence
  const int u = (sq / p.band) * (8 * p.band) + (blk & 7) * p.band + (sq % p.band);
  if (u >= p.units) return false;
  grp = u % p.ngroups;
  const int t0 = u / p.ngroups;
  const int t = p.live ? p.live[t0] : t0;
  if (t < 0 || t >= p.ntp * p.nkc) return false;
  const int tp = t % p.ntp;
  const int kc = t / p.ntp;
  const int nx1 = p.nx + 1;
  if (!p.split) {
    const int q = tp * WG + (int)threadIdx.x;
    if (q >= nx1 * (p.ny + 1)) return false;
    j = q / nx1;
    i = q - j * nx1;
  } else if (tp < p.tpa) {  // interior columns
    const int wa = p.a1 - p.a0;
    const int q = tp * WG + (int)threadIdx.x;
    j = q / wa;
    if (j > p.ny) return false;
    i = p.a0 + (q - j * wa);
  } else {  // row seams: [a1, nx] of row jj - 1, then [0, a0) of row jj
    const int q = (tp - p.tpa) * WG + (int)threadIdx.x;
    const int jj = q / p.wb;
    const int l = q - jj * p.wb;
    const int wh = nx1 - p.a1;
    if (l < wh) {
      if (jj == 0 || jj > p.ny + 1) return false;
      j = jj - 1;
      i = p.a1 + l;
    } else {
      if (jj > p.ny) return false;
      j = jj;
      i = l - wh;
    }
  }
  k0 = kc * p.kchunk;
  k1 = min(k0 + p.kchunk, p.nz + 1);
  return true;
}

// ---------------------------------------------------------------------------
// CPML slab corrections, one launch per face.
//
// Threads cover the face's slab nodes in the order of the psi layout (the
// in-plane index fastest: x for y/z faces, the layer for x faces) and every
// field set; each thread corrects the two tangential components at its node.
// Terms per normal axis (term 0, term 1), d and sign:
//   E  z: Ex  d = Hy(k) - Hy(k-1)   +      Ey  d = Hx(k) - Hx(k-1)   -
//      y: Ex  d = Hz(j) - Hz(j-1)   -      Ez  d = Hx(j) - Hx(j-1)   +
//      x: Ey  d = Hz(i) - Hz(i-1)   +      Ez  d = Hy(i) - Hy(i-1)   -
//   H  z: Hx  d = Ey(k+1) - Ey(k)   -      Hy  d = Ex(k+1) - Ex(k)   +
//      y: Hx  d = Ez(j+1) - Ez(j)   +      Hz  d = Ex(j+1) - Ex(j)   -
//      x: Hy  d = Ez(i+1) - Ez(i)   -      Hz  d = Ey(i+1) - Ey(i)   +
// (a neighbour at index -1 is zero, cpml.py:375-379).

template <typename T>
struct CFace {
  int k0, n, ksets;
  const T* coef;  // b[n] | c[n] | ck[n]
  T* psi;         // [K x S0 | K x S1]
  const unsigned char* air;  // [S0 | S1] or null
};

// psi-layout index of (i, j, k) in the slab of a face normal to AX, for a
// component with extents (n0, n1, n2)
template <int AX>
__device__ __forceinline__ int64_t slab_index(int i, int j, int k, int k0, int n, int n0, int n1) {
  if constexpr (AX == 0) return ((int64_t)k * n1 + j) * n + (i - k0);
  else if constexpr (AX == 1) return ((int64_t)k * n + (j - k0)) * n0 + i;
  else return ((int64_t)(k - k0) * n1 + j) * n0 + i;
}

template <typename T>
__device__ __forceinline__ void cpml_term(const CFace<T>& F, int term, int64_t size0, int64_t size,
                                          int64_t sidx, int s, int layer, T d, T beta, T* val,
                                          int sign) {
  T* psi = F.psi + (term ? (int64_t)F.ksets * size0 : 0) + (int64_t)s * size + sidx;
  const T b = F.coef[layer], c = F.coef[F.n + layer], ck = F.coef[2 * F.n + layer];
  T p = *psi;
  p = p * b;
  p = p + (c * d);
  if (F.air != nullptr && F.air[(term ? size0 : 0) + sidx] == 0) p = p * T(0);
  *psi = p;
  const T corr = beta * ((ck * d) + p);
  *val = sign > 0 ? (*val + corr) : (*val - corr);
}

// A CPML face of a fused launch (curl and CPML in one pass): its slab, the
// slab sizes of its two terms on the launch's side, its coefficient, psi and
// air buffers.
template <typename T>
struct PFace {
  int axis, k0, n, ksets;
  int64_t sz0, sz1;
  const T* coef;
  T* psi;
  const unsigned char* air;
};

// Every face in descriptor order, and per axis the (up to two) slab
// intervals of the node index, [lo, lo + len).  xin: the launch owns the
// x-slab elements (split planes: they fill whole waves of their own); when
// it does not, it leaves them, all their faces in order, to the x-column
// pass (k_cpml_xcols_e/h) and applies the y and z faces elsewhere.
template <typename T>
struct PmlArgs {
  int nf, xin;
  int lo[3][2], len[3][2];
  PFace<T> f[MGH_CPML_MAX_FACES];
};

template <typename T>
__device__ __forceinline__ bool pml_hit(const PmlArgs<T>& pa, int i, int j, int k) {
  const bool xs = (unsigned)(i - pa.lo[0][0]) < (unsigned)pa.len[0][0] ||
                  (unsigned)(i - pa.lo[0][1]) < (unsigned)pa.len[0][1];
  const bool yz = (unsigned)(j - pa.lo[1][0]) < (unsigned)pa.len[1][0] ||
                  (unsigned)(j - pa.lo[1][1]) < (unsigned)pa.len[1][1] ||
                  (unsigned)(k - pa.lo[2][0]) < (unsigned)pa.len[2][0] ||
                  (unsigned)(k - pa.lo[2][1]) < (unsigned)pa.len[2][1];
  return pa.xin ? (xs || yz) : (yz && !xs);
}

// Term `term` of face F on a component whose psi slab has x/y extents
// (n0, n1), at node (i, j, k) of set s, if the node lies in the slab: the
// recursion of cpml_term with the same operation order.
template <typename T>
__device__ __forceinline__ void pml_term(const PFace<T>& F, int term, int n0, int n1, int i, int j,
                                         int k, int s, T d, T beta, T* v, int sign) {
  const int c = F.axis == 0 ? i : (F.axis == 1 ? j : k);
  const int layer = c - F.k0;
  if ((unsigned)layer >= (unsigned)F.n) return;
  int64_t sidx;
  if (F.axis == 0) sidx = ((int64_t)k * n1 + j) * F.n + layer;
  else if (F.axis == 1) sidx = ((int64_t)k * F.n + layer) * n0 + i;
  else sidx = ((int64_t)layer * n1 + j) * n0 + i;
  T* psi = F.psi + (term ? (int64_t)F.ksets * F.sz0 : 0) + (int64_t)s * (term ? F.sz1 : F.sz0) +
           sidx;
  const T b = F.coef[layer], cc = F.coef[F.n + layer], ck = F.coef[2 * F.n + layer];
  T p = *psi;
  p = p * b;
  p = p + (cc * d);
  if (F.air != nullptr && F.air[(term ? F.sz0 : 0) + sidx] == 0) p = p * T(0);
  *psi = p;
  const T corr = beta * ((ck * d) + p);
  *v = sign > 0 ? (*v + corr) : (*v - corr);
}

// The CPML terms of one component, faces in descriptor order.  AXA/AXB: the
// component's two normal axes; TA/TB: its term index on those faces; SA/SB:
// the signs; dA/dB: the one-sided differences along them.  The loop is kept
// rolled, its face records read as uniform scalar loads: unrolled, the
// compiler holds every face's record in scalar registers, spills them and
// costs the whole launch two waves of occupancy per SIMD.
template <typename T, int AXA, int TA, int SA, int AXB, int TB, int SB>
__device__ __forceinline__ void pml_comp(const PmlArgs<T>& pa, int n0, int n1, int i, int j, int k,
                                         int s, T dA, T dB, T beta, T* v) {
#pragma unroll 1
  for (int fi = 0; fi < pa.nf; ++fi) {
    const int ax = pa.f[fi].axis;
    if (ax == AXA) pml_term(pa.f[fi], TA, n0, n1, i, j, k, s, dA, beta, v, SA);
    else if (ax == AXB) pml_term(pa.f[fi], TB, n0, n1, i, j, k, s, dB, beta, v, SB);
  }
}

// The E phase of one launch unit (a device function shared by the single
// launch and the packed launch of several grids).
template <typename T, int NS, bool CURL64, bool PML>
__device__ __forceinline__ void curl_e(T* __restrict__ e, const T* __restrict__ h,
                                       const T* __restrict__ ae, const T* __restrict__ be,
                                       Plan pl, int nsets, int64_t e_set, int64_t h_set,
                                       const PmlArgs<T>& pa) {
  int grp, i, j, k0, k1;
  if (!unit_of(pl, grp, i, j, k0, k1)) return;
  const int nx = pl.nx, ny = pl.ny, nz = pl.nz;
  const int s0 = grp * NS;

  const mgh::Dims d{nx, ny, nz};
  const int64_t oEy = d.n_ex(), oEz = oEy + d.n_ey();
  const int64_t oHy = d.n_hx(), oHz = oHy + d.n_hy();
  // per-plane sizes (x-fastest: plane = k)
  const int64_t pEx = (int64_t)(ny + 1) * nx, pEy = (int64_t)ny * (nx + 1),
                pEz = (int64_t)(ny + 1) * (nx + 1);
  const int64_t pHx = (int64_t)ny * (nx + 1), pHy = (int64_t)(ny + 1) * nx,
                pHz = (int64_t)ny * nx;
  // in-plane offsets: rows of nx (Ex, Hy, Hz) or nx + 1 (Ey, Ez, Hx) words
  const int64_t rN = (int64_t)j * nx + i;
  const int64_t rN1 = (int64_t)j * (nx + 1) + i;
  const bool inx = i < nx, iny = j < ny;

  T hy_prev[NS], hx_prev[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    hy_prev[q] = T(0);
    hx_prev[q] = T(0);
    if (s0 + q < nsets && k0 > 0) {
      const T* hs = h + (int64_t)(s0 + q) * h_set;
      if (inx) hy_prev[q] = hs[oHy + (int64_t)(k0 - 1) * pHy + rN];
      if (iny) hx_prev[q] = hs[(int64_t)(k0 - 1) * pHx + rN1];
    }
  }

  for (int k = k0; k < k1; ++k) {
    const bool knz = k < nz;
    bool hit = false;
    if constexpr (PML) hit = pml_hit(pa, i, j, k);
    const int64_t iEx = (int64_t)k * pEx + rN;
    const int64_t iEy = oEy + (int64_t)k * pEy + rN1;
    const int64_t iEz = oEz + (int64_t)k * pEz + rN1;
    T a_x = T(0), b_x = T(0), a_y = T(0), b_y = T(0), a_z = T(0), b_z = T(0);
    if (inx) { a_x = ae[iEx]; b_x = be[iEx]; }
    if (iny) { a_y = ae[iEy]; b_y = be[iEy]; }
    if (knz) { a_z = ae[iEz]; b_z = be[iEz]; }

#pragma unroll
    for (int q = 0; q < NS; ++q) {
      if (s0 + q >= nsets) break;
      const T* hs = h + (int64_t)(s0 + q) * h_set;
      T* es = e + (int64_t)(s0 + q) * e_set;
      // Every word of this set at this plane is loaded before its first store.
      // The three E components share one base pointer, and the loads sit in
      // the branches of components that may not exist, so a load placed after
      // an earlier component's store would stay there: each component would
      // cost the wave a memory round trip of its own.
      T hy = T(0), hx = T(0), hz = T(0), hz_jm = T(0), hz_im = T(0), hy_im = T(0), hx_jm = T(0);
      T ex = T(0), ey = T(0), ez = T(0);
      if (inx && knz) hy = hs[oHy + (int64_t)k * pHy + rN];
      if (iny && knz) hx = hs[(int64_t)k * pHx + rN1];
      if (inx && iny) hz = hs[oHz + (int64_t)k * pHz + rN];
      if (inx && j > 0) hz_jm = hs[oHz + (int64_t)k * pHz + rN - nx];
      if (iny && i > 0) hz_im = hs[oHz + (int64_t)k * pHz + rN - 1];
      if (knz && i > 0) hy_im = hs[oHy + (int64_t)k * pHy + rN - 1];
      if (knz && j > 0) hx_jm = hs[(int64_t)k * pHx + rN1 - (nx + 1)];
      if (inx) ex = es[iEx];
      if (iny) ey = es[iEy];
      if (knz) ez = es[iEz];

      // All three components are computed before the first store (each
      // store ahead of a later component's arithmetic held the wave on it).
      T vx = T(0), vy = T(0), vz = T(0);
      if (inx) {  // Ex(i,j,k): dHz/dy - dHy/dz
        if constexpr (CURL64) {
          double c = 0.0;
          if (iny) c += (double)hz;
          if (j > 0) c -= (double)hz_jm;
          if (knz) c -= (double)hy;
          if (k > 0) c += (double)hy_prev[q];
          vx = (T)((double)(a_x * ex) + (double)b_x * c);
        } else {
          T c = T(0);
          if (iny) c += hz;
          if (j > 0) c -= hz_jm;
          if (knz) c -= hy;
          if (k > 0) c += hy_prev[q];
          vx = (a_x * ex) + (b_x * c);
        }
      }
      if (iny) {  // Ey(i,j,k): dHx/dz - dHz/dx
        if constexpr (CURL64) {
          double c = 0.0;
          if (knz) c += (double)hx;
          if (k > 0) c -= (double)hx_prev[q];
          if (inx) c -= (double)hz;
          if (i > 0) c += (double)hz_im;
          vy = (T)((double)(a_y * ey) + (double)b_y * c);
        } else {
          T c = T(0);
          if (knz) c += hx;
          if (k > 0) c -= hx_prev[q];
          if (inx) c -= hz;
          if (i > 0) c += hz_im;
          vy = (a_y * ey) + (b_y * c);
        }
      }
      if (knz) {  // Ez(i,j,k): dHy/dx - dHx/dy
        if constexpr (CURL64) {
          double c = 0.0;
          if (inx) c += (double)hy;
          if (i > 0) c -= (double)hy_im;
          if (iny) c -= (double)hx;
          if (j > 0) c += (double)hx_jm;
          vz = (T)((double)(a_z * ez) + (double)b_z * c);
        } else {
          T c = T(0);
          if (inx) c += hy;
          if (i > 0) c -= hy_im;
          if (iny) c -= hx;
          if (j > 0) c += hx_jm;
          vz = (a_z * ez) + (b_z * c);
        }
      }
      if constexpr (PML) {
        if (hit) {
          // Ex: z faces d = Hy(k) - Hy(k-1) (+), y faces Hz(j) - Hz(j-1) (-)
          if (inx) pml_comp<T, 2, 0, +1, 1, 0, -1>(pa, nx, ny + 1, i, j, k, s0 + q,
                                                   hy - hy_prev[q], hz - hz_jm, b_x, &vx);
          // Ey: z faces d = Hx(k) - Hx(k-1) (-), x faces Hz(i) - Hz(i-1) (+)
          if (iny) pml_comp<T, 2, 1, -1, 0, 0, +1>(pa, nx + 1, ny, i, j, k, s0 + q,
                                                   hx - hx_prev[q], hz - hz_im, b_y, &vy);
          // Ez: y faces d = Hx(j) - Hx(j-1) (+), x faces Hy(i) - Hy(i-1) (-)
          if (knz) pml_comp<T, 1, 1, +1, 0, 1, -1>(pa, nx + 1, ny + 1, i, j, k, s0 + q,
                                                   hx - hx_jm, hy - hy_im, b_z, &vz);
        }
      }
      if (inx) es[iEx] = vx;
      if (iny) es[iEy] = vy;
      if (knz) es[iEz] = vz;
      hy_prev[q] = hy;
      hx_prev[q] = hx;
    }
  }
}

template <typename T, int NS, bool CURL64, bool PML>
__global__ __launch_bounds__(WG) void k_update_e(T* __restrict__ e, const T* __restrict__ h,
                                                 const T* __restrict__ ae,
                                                 const T* __restrict__ be, Plan pl, int nsets,
                                                 int64_t e_set, int64_t h_set, PmlArgs<T> pa) {
  curl_e<T, NS, CURL64, PML>(e, h, ae, be, pl, nsets, e_set, h_set, pa);
}

template <typename T, int NS, bool PML>
__global__ __launch_bounds__(WG) void k_update_h(const T* __restrict__ e, T* __restrict__ h,
                                                 const T* __restrict__ ah,
                                                 const T* __restrict__ bh, Plan pl, int nsets,
                                                 int64_t e_set, int64_t h_set, PmlArgs<T> pa) {
#include "mgh_curl_h_body.inc"
}

// Packed launches: the curl phase of several grids (jobs of a serving loop)
// in one launch, work-group row y for job y.  Each job keeps its own plan,
// buffers and XCD bands; one set per thread; no CPML.
constexpr int PACK_MAX = 16;

template <typename T>
struct PackJob {
  T* e;
  T* h;
  const T* a;  // alpha of the phase's side
  const T* b;  // beta
  int64_t e_set, h_set;
  int nsets, pad;
  Plan pl;
};

template <typename T>
struct Pack {
  int n, pad;
  PackJob<T> j[PACK_MAX];
};

template <typename T>
__global__ __launch_bounds__(WG) void k_update_e_pack(Pack<T> pk) {
  const PackJob<T>& J = pk.j[blockIdx.y];
  const PmlArgs<T> none{};
  curl_e<T, 1, false, false>(J.e, J.h, J.a, J.b, J.pl, J.nsets, J.e_set, J.h_set, none);
}

template <typename T>
__global__ __launch_bounds__(WG) void k_update_h_pack(Pack<T> pk) {
  constexpr int NS = 1;
  constexpr bool PML = false;
  const PackJob<T>& J = pk.j[blockIdx.y];
  const T* __restrict__ e = J.e;
  T* __restrict__ h = J.h;
  const T* __restrict__ ah = J.a;
  const T* __restrict__ bh = J.b;
  const Plan pl = J.pl;
  const int nsets = J.nsets;
  const int64_t e_set = J.e_set, h_set = J.h_set;
  const PmlArgs<T> pa{};
#include "mgh_curl_h_body.inc"
}

// ---------------------------------------------------------------------------
// CPML slab passes (separate launches: listed launches, dispersive meshes).

// slab node (i, j, k) of thread t, in psi-layout order
template <int AX>
__device__ __forceinline__ bool slab_node(int64_t t, int nx, int ny, int nz, int k0, int n, int& i,
                                          int& j, int& k) {
  if constexpr (AX == 2) {
    const int64_t per = (int64_t)(nx + 1) * (ny + 1);
    if (t >= per * n) return false;
    const int l = (int)(t / per);
    const int r = (int)(t - (int64_t)l * per);
    j = r / (nx + 1);
    i = r - j * (nx + 1);
    k = k0 + l;
  } else if constexpr (AX == 1) {
    const int64_t per = (int64_t)(nx + 1) * n;
    if (t >= per * (nz + 1)) return false;
    k = (int)(t / per);
    const int r = (int)(t - (int64_t)k * per);
    const int l = r / (nx + 1);
    i = r - l * (nx + 1);
    j = k0 + l;
  } else {
    const int64_t per = (int64_t)n * (ny + 1);
    if (t >= per * (nz + 1)) return false;
    k = (int)(t / per);
    const int r = (int)(t - (int64_t)k * per);
    j = r / n;
    i = k0 + (r - j * n);
  }
  return true;
}

template <typename T, int AX>
__global__ __launch_bounds__(WG) void k_cpml_e(T* __restrict__ e, const T* __restrict__ h,
                                               const T* __restrict__ be, int nx, int ny, int nz,
                                               int64_t e_set, int64_t h_set, CFace<T> F) {
  int i, j, k;
  if (!slab_node<AX>((int64_t)blockIdx.x * WG + threadIdx.x, nx, ny, nz, F.k0, F.n, i, j, k))
    return;
  const int s = blockIdx.y;
  const mgh::Dims d{nx, ny, nz};
  T* es = e + (int64_t)s * e_set;
  const T* hs = h + (int64_t)s * h_set;
  const int64_t oEy = d.n_ex(), oEz = oEy + d.n_ey();
  const int64_t oHy = d.n_hx(), oHz = oHy + d.n_hy();
  const int64_t pEx = (int64_t)(ny + 1) * nx, pEy = (int64_t)ny * (nx + 1),
                pEz = (int64_t)(ny + 1) * (nx + 1);
  const int64_t pHx = (int64_t)ny * (nx + 1), pHy = (int64_t)(ny + 1) * nx,
                pHz = (int64_t)ny * nx;
  const int64_t rN = (int64_t)j * nx + i, rN1 = (int64_t)j * (nx + 1) + i;
  auto Hx = [&](int di, int dj, int dk) { return hs[(int64_t)(k + dk) * pHx + rN1 + dj * (nx + 1) + di]; };
  auto Hy = [&](int di, int dj, int dk) { return hs[oHy + (int64_t)(k + dk) * pHy + rN + dj * nx + di]; };
  auto Hz = [&](int di, int dj, int dk) { return hs[oHz + (int64_t)(k + dk) * pHz + rN + dj * nx + di]; };
  if constexpr (AX == 2) {
    const int layer = k - F.k0;
    const int64_t s0 = (int64_t)F.n * nx * (ny + 1);  // Ex slab
    if (i < nx) {  // Ex
      const int64_t ie = (int64_t)k * pEx + rN;
      const T dd = Hy(0, 0, 0) - (k > 0 ? Hy(0, 0, -1) : T(0));
      T v = es[ie];
      cpml_term(F, 0, s0, s0, slab_index<2>(i, j, k, F.k0, F.n, nx, ny + 1), s, layer, dd, be[ie],
                &v, +1);
      es[ie] = v;
    }
    if (j < ny) {  // Ey
      const int64_t ie = oEy + (int64_t)k * pEy + rN1;
      const T dd = Hx(0, 0, 0) - (k > 0 ? Hx(0, 0, -1) : T(0));
      T v = es[ie];
      cpml_term(F, 1, s0, (int64_t)F.n * (nx + 1) * ny,
                slab_index<2>(i, j, k, F.k0, F.n, nx + 1, ny), s, layer, dd, be[ie], &v, -1);
      es[ie] = v;
    }
  } else if constexpr (AX == 1) {
    const int layer = j - F.k0;
    const int64_t s0 = (int64_t)F.n * nx * (nz + 1);  // Ex slab
    if (i < nx) {  // Ex
      const int64_t ie = (int64_t)k * pEx + rN;
      const T dd = Hz(0, 0, 0) - (j > 0 ? Hz(0, -1, 0) : T(0));
      T v = es[ie];
      cpml_term(F, 0, s0, s0, slab_index<1>(i, j, k, F.k0, F.n, nx, ny + 1), s, layer, dd, be[ie],
                &v, -1);
      es[ie] = v;
    }
    if (k < nz) {  // Ez
      const int64_t ie = oEz + (int64_t)k * pEz + rN1;
      const T dd = Hx(0, 0, 0) - (j > 0 ? Hx(0, -1, 0) : T(0));
      T v = es[ie];
      cpml_term(F, 1, s0, (int64_t)F.n * (nx + 1) * nz,
                slab_index<1>(i, j, k, F.k0, F.n, nx + 1, ny + 1), s, layer, dd, be[ie], &v, +1);
      es[ie] = v;
    }
  } else {
    const int layer = i - F.k0;
    const int64_t s0 = (int64_t)F.n * ny * (nz + 1);  // Ey slab
    if (j < ny) {  // Ey
      const int64_t ie = oEy + (int64_t)k * pEy + rN1;
      const T dd = Hz(0, 0, 0) - (i > 0 ? Hz(-1, 0, 0) : T(0));
      T v = es[ie];
      cpml_term(F, 0, s0, s0, slab_index<0>(i, j, k, F.k0, F.n, nx + 1, ny), s, layer, dd, be[ie],
                &v, +1);
      es[ie] = v;
    }
    if (k < nz) {  // Ez
      const int64_t ie = oEz + (int64_t)k * pEz + rN1;
      const T dd = Hy(0, 0, 0) - (i > 0 ? Hy(-1, 0, 0) : T(0));
      T v = es[ie];
      cpml_term(F, 1, s0, (int64_t)F.n * (ny + 1) * nz,
                slab_index<0>(i, j, k, F.k0, F.n, nx + 1, ny + 1), s, layer, dd, be[ie], &v, -1);
      es[ie] = v;
    }
  }
}

template <typename T, int AX>
__global__ __launch_bounds__(WG) void k_cpml_h(const T* __restrict__ e, T* __restrict__ h,
                                               const T* __restrict__ bh, int nx, int ny, int nz,
                                               int64_t e_set, int64_t h_set, CFace<T> F) {
  int i, j, k;
  if (!slab_node<AX>((int64_t)blockIdx.x * WG + threadIdx.x, nx, ny, nz, F.k0, F.n, i, j, k))
    return;
  const int s = blockIdx.y;
  const mgh::Dims d{nx, ny, nz};
  const T* es = e + (int64_t)s * e_set;
  T* hs = h + (int64_t)s * h_set;
  const int64_t oEy = d.n_ex(), oEz = oEy + d.n_ey();
  const int64_t oHy = d.n_hx(), oHz = oHy + d.n_hy();
  const int64_t pEx = (int64_t)(ny + 1) * nx, pEy = (int64_t)ny * (nx + 1),
                pEz = (int64_t)(ny + 1) * (nx + 1);
  const int64_t pHx = (int64_t)ny * (nx + 1), pHy = (int64_t)(ny + 1) * nx,
                pHz = (int64_t)ny * nx;
  const int64_t rN = (int64_t)j * nx + i, rN1 = (int64_t)j * (nx + 1) + i;
  auto Ex = [&](int di, int dj, int dk) { return es[(int64_t)(k + dk) * pEx + rN + dj * nx + di]; };
  auto Ey = [&](int di, int dj, int dk) { return es[oEy + (int64_t)(k + dk) * pEy + rN1 + dj * (nx + 1) + di]; };
  auto Ez = [&](int di, int dj, int dk) { return es[oEz + (int64_t)(k + dk) * pEz + rN1 + dj * (nx + 1) + di]; };
  // H components exist for: Hx (j < ny, k < nz), Hy (i < nx, k < nz), Hz (i < nx, j < ny)
  if constexpr (AX == 2) {
    const int layer = k - F.k0;
    const int64_t s0 = (int64_t)F.n * (nx + 1) * ny;  // Hx slab
    if (j < ny) {  // Hx
      const int64_t ih = (int64_t)k * pHx + rN1;
      const T dd = Ey(0, 0, 1) - Ey(0, 0, 0);
      T v = hs[ih];
      cpml_term(F, 0, s0, s0, slab_index<2>(i, j, k, F.k0, F.n, nx + 1, ny), s, layer, dd, bh[ih],
                &v, -1);
      hs[ih] = v;
    }
    if (i < nx) {  // Hy
      const int64_t ih = oHy + (int64_t)k * pHy + rN;
      const T dd = Ex(0, 0, 1) - Ex(0, 0, 0);
      T v = hs[ih];
      cpml_term(F, 1, s0, (int64_t)F.n * nx * (ny + 1),
                slab_index<2>(i, j, k, F.k0, F.n, nx, ny + 1), s, layer, dd, bh[ih], &v, +1);
      hs[ih] = v;
    }
  } else if constexpr (AX == 1) {
    const int layer = j - F.k0;
    const int64_t s0 = (int64_t)F.n * (nx + 1) * nz;  // Hx slab
    if (k < nz) {  // Hx
      const int64_t ih = (int64_t)k * pHx + rN1;
      const T dd = Ez(0, 1, 0) - Ez(0, 0, 0);
      T v = hs[ih];
      cpml_term(F, 0, s0, s0, slab_index<1>(i, j, k, F.k0, F.n, nx + 1, ny), s, layer, dd, bh[ih],
                &v, +1);
      hs[ih] = v;
    }
    if (i < nx) {  // Hz
      const int64_t ih = oHz + (int64_t)k * pHz + rN;
      const T dd = Ex(0, 1, 0) - Ex(0, 0, 0);
      T v = hs[ih];
      cpml_term(F, 1, s0, (int64_t)F.n * nx * (nz + 1),
                slab_index<1>(i, j, k, F.k0, F.n, nx, ny), s, layer, dd, bh[ih], &v, -1);
      hs[ih] = v;
    }
  } else {
    const int layer = i - F.k0;
    const int64_t s0 = (int64_t)F.n * (ny + 1) * nz;  // Hy slab
    if (k < nz) {  // Hy
      const int64_t ih = oHy + (int64_t)k * pHy + rN;
      const T dd = Ez(1, 0, 0) - Ez(0, 0, 0);
      T v = hs[ih];
      cpml_term(F, 0, s0, s0, slab_index<0>(i, j, k, F.k0, F.n, nx, ny + 1), s, layer, dd, bh[ih],
                &v, -1);
      hs[ih] = v;
    }
    if (j < ny) {  // Hz
      const int64_t ih = oHz + (int64_t)k * pHz + rN;
      const T dd = Ey(1, 0, 0) - Ey(0, 0, 0);
      T v = hs[ih];
      cpml_term(F, 1, s0, (int64_t)F.n * ny * (nz + 1),
                slab_index<0>(i, j, k, F.k0, F.n, nx, ny), s, layer, dd, bh[ih], &v, +1);
      hs[ih] = v;
    }
  }
}

// CPML of the x-slab columns: the elements of every x-normal face's slab,
// which the curl launch leaves alone when it carries the CPML.  Each element
// applies the terms of every face whose slab holds it (x, y and z faces), in
// descriptor order, to its freshly updated value — the d of each term
// computed as the separate slab passes above compute it — so the result is
// theirs bit for bit.  Threads cover each x face's slab layer-fastest (the
// psi order) by seams: the upper x slab of row j - 1 and the lower x slab of
// row j are adjacent in memory (a row's end meets the next row's start), so
// the threads of one seam, upper layers first, touch one or two cache lines
// per field array where the two slabs walked apart touch each line twice.
template <typename T>
struct XCols {
  int nf;  // every face, descriptor order
  int axis[MGH_CPML_MAX_FACES];
  CFace<T> f[MGH_CPML_MAX_FACES];
  int lo_k0, lo_n, hi_k0, hi_n;  // the lower / upper x slab (n = 0: absent)
};

template <typename T>
__device__ __forceinline__ bool xcols_node(const XCols<T>& X, int64_t t, int ny, int nz, int& i,
                                           int& j, int& k) {
  const int w = X.hi_n + X.lo_n;  // threads per seam
  const int64_t seam = t / w;
  const int l = (int)(t - seam * w);
  k = (int)(seam / (ny + 2));
  const int jj = (int)(seam - (int64_t)k * (ny + 2));  // seam between rows jj - 1 and jj
  if (k > nz) return false;
  if (l < X.hi_n) {  // upper slab of row jj - 1
    if (jj == 0) return false;
    j = jj - 1;
    i = X.hi_k0 + l;
  } else {  // lower slab of row jj
    if (jj > ny) return false;
    j = jj;
    i = X.lo_k0 + (l - X.hi_n);
  }
  return true;
}

// layer of coordinate c in face F's slab, or -1
template <typename T>
__device__ __forceinline__ int in_slab(const CFace<T>& F, int c) {
  const int l = c - F.k0;
  return (unsigned)l < (unsigned)F.n ? l : -1;
}

template <typename T>
__global__ __launch_bounds__(WG) void k_cpml_xcols_e(T* __restrict__ e, const T* __restrict__ h,
                                                     const T* __restrict__ be, int nx, int ny,
                                                     int nz, int64_t e_set, int64_t h_set,
                                                     XCols<T> X) {
  int i, j, k;
  if (!xcols_node(X, (int64_t)blockIdx.x * WG + threadIdx.x, ny, nz, i, j, k)) return;
  const int s = blockIdx.y;
  const mgh::Dims d{nx, ny, nz};
  T* es = e + (int64_t)s * e_set;
  const T* hs = h + (int64_t)s * h_set;
  const int64_t oEy = d.n_ex(), oEz = oEy + d.n_ey();
  const int64_t oHy = d.n_hx(), oHz = oHy + d.n_hy();
  const int64_t pEx = (int64_t)(ny + 1) * nx, pEy = (int64_t)ny * (nx + 1),
                pEz = (int64_t)(ny + 1) * (nx + 1);
  const int64_t pHx = (int64_t)ny * (nx + 1), pHy = (int64_t)(ny + 1) * nx,
                pHz = (int64_t)ny * nx;
  const int64_t rN = (int64_t)j * nx + i, rN1 = (int64_t)j * (nx + 1) + i;
  auto Hx = [&](int di, int dj, int dk) { return hs[(int64_t)(k + dk) * pHx + rN1 + dj * (nx + 1) + di]; };
  auto Hy = [&](int di, int dj, int dk) { return hs[oHy + (int64_t)(k + dk) * pHy + rN + dj * nx + di]; };
  auto Hz = [&](int di, int dj, int dk) { return hs[oHz + (int64_t)(k + dk) * pHz + rN + dj * nx + di]; };
  if (i < nx) {  // Ex: z faces (term 0, +), y faces (term 0, -)
    const int64_t ie = (int64_t)k * pEx + rN;
    T v = es[ie];
    bool any = false;
#pragma unroll 1
    for (int f = 0; f < X.nf; ++f) {
      const CFace<T>& F = X.f[f];
      if (X.axis[f] == 2) {
        const int l = in_slab(F, k);
        if (l < 0) continue;
        const int64_t s0 = (int64_t)F.n * nx * (ny + 1);
        const T dd = Hy(0, 0, 0) - (k > 0 ? Hy(0, 0, -1) : T(0));
        cpml_term(F, 0, s0, s0, slab_index<2>(i, j, k, F.k0, F.n, nx, ny + 1), s, l, dd, be[ie], &v, +1);
        any = true;
      } else if (X.axis[f] == 1) {
        const int l = in_slab(F, j);
        if (l < 0) continue;
        const int64_t s0 = (int64_t)F.n * nx * (nz + 1);
        const T dd = Hz(0, 0, 0) - (j > 0 ? Hz(0, -1, 0) : T(0));
        cpml_term(F, 0, s0, s0, slab_index<1>(i, j, k, F.k0, F.n, nx, ny + 1), s, l, dd, be[ie], &v, -1);
        any = true;
      }
    }
    if (any) es[ie] = v;
  }
  if (j < ny) {  // Ey: z faces (term 1, -), x faces (term 0, +)
    const int64_t ie = oEy + (int64_t)k * pEy + rN1;
    T v = es[ie];
    bool any = false;
#pragma unroll 1
    for (int f = 0; f < X.nf; ++f) {
      const CFace<T>& F = X.f[f];
      if (X.axis[f] == 2) {
        const int l = in_slab(F, k);
        if (l < 0) continue;
        const T dd = Hx(0, 0, 0) - (k > 0 ? Hx(0, 0, -1) : T(0));
        cpml_term(F, 1, (int64_t)F.n * nx * (ny + 1), (int64_t)F.n * (nx + 1) * ny,
                  slab_index<2>(i, j, k, F.k0, F.n, nx + 1, ny), s, l, dd, be[ie], &v, -1);
        any = true;
      } else if (X.axis[f] == 0) {
        const int l = in_slab(F, i);
        if (l < 0) continue;
        const int64_t s0 = (int64_t)F.n * ny * (nz + 1);
        const T dd = Hz(0, 0, 0) - (i > 0 ? Hz(-1, 0, 0) : T(0));
        cpml_term(F, 0, s0, s0, slab_index<0>(i, j, k, F.k0, F.n, nx + 1, ny), s, l, dd, be[ie], &v, +1);
        any = true;
      }
    }
    if (any) es[ie] = v;
  }
  if (k < nz) {  // Ez: y faces (term 1, +), x faces (term 1, -)
    const int64_t ie = oEz + (int64_t)k * pEz + rN1;
    T v = es[ie];
    bool any = false;
#pragma unroll 1
    for (int f = 0; f < X.nf; ++f) {
      const CFace<T>& F = X.f[f];
      if (X.axis[f] == 1) {
        const int l = in_slab(F, j);
        if (l < 0) continue;
        const T dd = Hx(0, 0, 0) - (j > 0 ? Hx(0, -1, 0) : T(0));
        cpml_term(F, 1, (int64_t)F.n * nx * (nz + 1), (int64_t)F.n * (nx + 1) * nz,
                  slab_index<1>(i, j, k, F.k0, F.n, nx + 1, ny + 1), s, l, dd, be[ie], &v, +1);
        any = true;
      } else if (X.axis[f] == 0) {
        const int l = in_slab(F, i);
        if (l < 0) continue;
        const T dd = Hy(0, 0, 0) - (i > 0 ? Hy(-1, 0, 0) : T(0));
        cpml_term(F, 1, (int64_t)F.n * ny * (nz + 1), (int64_t)F.n * (ny + 1) * nz,
                  slab_index<0>(i, j, k, F.k0, F.n, nx + 1, ny + 1), s, l, dd, be[ie], &v, -1);
        any = true;
      }
    }
    if (any) es[ie] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(WG) void k_cpml_xcols_h(const T* __restrict__ e, T* __restrict__ h,
                                                     const T* __restrict__ bh, int nx, int ny,
                                                     int nz, int64_t e_set, int64_t h_set,
                                                     XCols<T> X) {
  int i, j, k;
  if (!xcols_node(X, (int64_t)blockIdx.x * WG + threadIdx.x, ny, nz, i, j, k)) return;
  const int s = blockIdx.y;
  const mgh::Dims d{nx, ny, nz};
  const T* es = e + (int64_t)s * e_set;
  T* hs = h + (int64_t)s * h_set;
  const int64_t oEy = d.n_ex(), oEz = oEy + d.n_ey();
  const int64_t oHy = d.n_hx(), oHz = oHy + d.n_hy();
  const int64_t pEx = (int64_t)(ny + 1) * nx, pEy = (int64_t)ny * (nx + 1),
                pEz = (int64_t)(ny + 1) * (nx + 1);
  const int64_t pHx = (int64_t)ny * (nx + 1), pHy = (int64_t)(ny + 1) * nx,
                pHz = (int64_t)ny * nx;
  const int64_t rN = (int64_t)j * nx + i, rN1 = (int64_t)j * (nx + 1) + i;
  auto Ex = [&](int di, int dj, int dk) { return es[(int64_t)(k + dk) * pEx + rN + dj * nx + di]; };
  auto Ey = [&](int di, int dj, int dk) { return es[oEy + (int64_t)(k + dk) * pEy + rN1 + dj * (nx + 1) + di]; };
  auto Ez = [&](int di, int dj, int dk) { return es[oEz + (int64_t)(k + dk) * pEz + rN1 + dj * (nx + 1) + di]; };
  if (j < ny && k < nz) {  // Hx: z faces (term 0, -), y faces (term 0, +)
    const int64_t ih = (int64_t)k * pHx + rN1;
    T v = hs[ih];
    bool any = false;
#pragma unroll 1
    for (int f = 0; f < X.nf; ++f) {
      const CFace<T>& F = X.f[f];
      if (X.axis[f] == 2) {
        const int l = in_slab(F, k);
        if (l < 0) continue;
        const int64_t s0 = (int64_t)F.n * (nx + 1) * ny;
        const T dd = Ey(0, 0, 1) - Ey(0, 0, 0);
        cpml_term(F, 0, s0, s0, slab_index<2>(i, j, k, F.k0, F.n, nx + 1, ny), s, l, dd, bh[ih], &v, -1);
        any = true;
      } else if (X.axis[f] == 1) {
        const int l = in_slab(F, j);
        if (l < 0) continue;
        const int64_t s0 = (int64_t)F.n * (nx + 1) * nz;
        const T dd = Ez(0, 1, 0) - Ez(0, 0, 0);
        cpml_term(F, 0, s0, s0, slab_index<1>(i, j, k, F.k0, F.n, nx + 1, ny), s, l, dd, bh[ih], &v, +1);
        any = true;
      }
    }
    if (any) hs[ih] = v;
  }
  if (k < nz) {  // Hy: z faces (term 1, +), x faces (term 0, -)
    const int64_t ih = oHy + (int64_t)k * pHy + rN;
    T v = hs[ih];
    bool any = false;
#pragma unroll 1
    for (int f = 0; f < X.nf; ++f) {
      const CFace<T>& F = X.f[f];
      if (X.axis[f] == 2) {
        const int l = in_slab(F, k);
        if (l < 0) continue;
        const T dd = Ex(0, 0, 1) - Ex(0, 0, 0);
        cpml_term(F, 1, (int64_t)F.n * (nx + 1) * ny, (int64_t)F.n * nx * (ny + 1),
                  slab_index<2>(i, j, k, F.k0, F.n, nx, ny + 1), s, l, dd, bh[ih], &v, +1);
        any = true;
      } else if (X.axis[f] == 0) {
        const int l = in_slab(F, i);
        if (l < 0) continue;
        const int64_t s0 = (int64_t)F.n * (ny + 1) * nz;
        const T dd = Ez(1, 0, 0) - Ez(0, 0, 0);
        cpml_term(F, 0, s0, s0, slab_index<0>(i, j, k, F.k0, F.n, nx, ny + 1), s, l, dd, bh[ih], &v, -1);
        any = true;
      }
    }
    if (any) hs[ih] = v;
  }
  if (j < ny) {  // Hz: y faces (term 1, -), x faces (term 1, +)
    const int64_t ih = oHz + (int64_t)k * pHz + rN;
    T v = hs[ih];
    bool any = false;
#pragma unroll 1
    for (int f = 0; f < X.nf; ++f) {
      const CFace<T>& F = X.f[f];
      if (X.axis[f] == 1) {
        const int l = in_slab(F, j);
        if (l < 0) continue;
        const T dd = Ex(0, 1, 0) - Ex(0, 0, 0);
        cpml_term(F, 1, (int64_t)F.n * (nx + 1) * nz, (int64_t)F.n * nx * (nz + 1),
                  slab_index<1>(i, j, k, F.k0, F.n, nx, ny), s, l, dd, bh[ih], &v, -1);
        any = true;
      } else if (X.axis[f] == 0) {
        const int l = in_slab(F, i);
        if (l < 0) continue;
        const T dd = Ey(1, 0, 0) - Ey(0, 0, 0);
        cpml_term(F, 1, (int64_t)F.n * (ny + 1) * nz, (int64_t)F.n * ny * (nz + 1),
                  slab_index<0>(i, j, k, F.k0, F.n, nx, ny), s, l, dd, bh[ih], &v, +1);
        any = true;
      }
    }
    if (any) hs[ih] = v;
  }
}

// ---------------------------------------------------------------------------
// host side

// field sets per thread: each coefficient word a thread loads serves them
// all.  Three when K is a multiple of 3 (K = 3 on 192x96x288: 158 against
// 181 us per E launch with pairs, which leave one group half empty; K = 6:
// E / H 266-275 / 277-281 us against 268-277 / 273-288 with pairs, each
// component computed before the first store), else two.  Six at K = 6:
// 372 / 321 us (140-153 VGPRs, 3 waves per SIMD).
constexpr int kSetsPerThread = 2;

inline int sets_per_thread(int nsets) {
  if (nsets < 2) return 1;
  return nsets % 3 == 0 ? 3 : kSetsPerThread;
}

// a0 < a1: split the planes at the interior columns [a0, a1) (see Plan)
Plan make_plan(int nx, int ny, int nz, int nsets, int kchunk, int ns, int band,
               const int* live = nullptr, int nlive = 0, int a0 = 0, int a1 = 0) {
  Plan p;
  std::memset(&p, 0, sizeof(p));
  p.band = band;
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.kchunk = kchunk;
  p.ngroups = (nsets + ns - 1) / ns;
  p.ntp = (int)(((int64_t)(nx + 1) * (ny + 1) + WG - 1) / WG);
  if (!live && a0 < a1) {
    p.split = 1;
    p.a0 = a0;
    p.a1 = a1;
    p.wb = (nx + 1 - a1) + a0;
    p.tpa = (int)(((int64_t)(a1 - a0) * (ny + 1) + WG - 1) / WG);
    p.ntp = p.tpa + (int)(((int64_t)p.wb * (ny + 2) + WG - 1) / WG);
  }
  p.nkc = (nz + 1 + kchunk - 1) / kchunk;
  p.live = live;
  p.units = (live ? nlive : p.ntp * p.nkc) * p.ngroups;
  return p;
}

// work-groups of a launch: units rounded up to whole bands on every XCD
int plan_blocks(const Plan& p) {
  const int64_t w = 8 * (int64_t)p.band;
  return (int)((p.units + w - 1) / w * w);
}

// The launch shape word of the curl entry points: z planes marched per
// thread in bits 0-15, field sets per thread in bits 16-19 (0: the
// library's choice, sets_per_thread), the XCD band in launch units in bits
// 20-27 (0: BAND).  Results do not depend on it.
struct Shape {
  int kchunk, ns, band;
};

inline Shape decode_shape(int word, int nsets) {
  Shape s;
  s.kchunk = word & 0xffff;
  const int ns = (word >> 16) & 0xf, band = (word >> 20) & 0xff;
  s.ns = ns ? ns : sets_per_thread(nsets);
  s.band = band ? band : BAND;
  return s;
}

template <typename T>
CFace<T> face_of(const mgh::CpmlFace& F, bool e_side) {
  CFace<T> c;
  c.k0 = F.k0;
  c.n = F.n;
  c.ksets = F.nsets;
  c.coef = (const T*)F.coef;
  c.psi = (T*)(e_side ? F.psi_e : F.psi_h);
  c.air = e_side ? F.air_e : F.air_h;
  return c;
}

int64_t slab_threads(int axis, int nx, int ny, int nz, int n) {
  if (axis == 2) return (int64_t)(nx + 1) * (ny + 1) * n;
  if (axis == 1) return (int64_t)(nx + 1) * n * (nz + 1);
  return (int64_t)n * (ny + 1) * (nz + 1);
}

// every face of a CPML set for a fused launch (E or H side), descriptor
// order, and every slab interval; xin is set by the caller
template <typename T>
PmlArgs<T> pml_args(const mgh::CpmlSet* pml, bool e_side, int nx, int ny, int nz) {
  PmlArgs<T> a;
  std::memset(&a, 0, sizeof(a));  // empty intervals: len 0
  if (!pml) return a;
  int used[3] = {0, 0, 0};
  for (int f = 0; f < pml->nfaces; ++f) {
    const mgh::CpmlFace& F = pml->f[f];
    const int u = used[F.axis]++;  // at most two faces per axis (load_cpml_set)
    a.lo[F.axis][u] = F.k0;
    a.len[F.axis][u] = F.n;
    PFace<T>& P = a.f[a.nf++];
    P.axis = F.axis;
    P.k0 = F.k0;
    P.n = F.n;
    P.ksets = F.nsets;
    P.coef = (const T*)F.coef;
    P.psi = (T*)(e_side ? F.psi_e : F.psi_h);
    P.air = e_side ? F.air_e : F.air_h;
    const int64_t n = F.n;
    // term slab sizes: n times the component's extents along the other axes
    if (e_side) {
      if (F.axis == 2) { P.sz0 = n * nx * (ny + 1); P.sz1 = n * (nx + 1) * ny; }
      else if (F.axis == 1) { P.sz0 = n * nx * (nz + 1); P.sz1 = n * (nx + 1) * nz; }
      else { P.sz0 = n * ny * (nz + 1); P.sz1 = n * (ny + 1) * nz; }
    } else {
      if (F.axis == 2) { P.sz0 = n * (nx + 1) * ny; P.sz1 = n * nx * (ny + 1); }
      else if (F.axis == 1) { P.sz0 = n * (nx + 1) * nz; P.sz1 = n * nx * (nz + 1); }
      else { P.sz0 = n * (ny + 1) * nz; P.sz1 = n * ny * (nz + 1); }
    }
  }
  return a;
}

// Split planes for a CPML set: both x slabs at the domain ends ([0, a0) and
// [a1, nx)) with interior columns between them.
bool x_split(const mgh::CpmlSet* pml, int nx, int* a0, int* a1) {
  if (!pml) return false;
  int lo = -1, hi = -1;
  for (int f = 0; f < pml->nfaces; ++f) {
    const mgh::CpmlFace& F = pml->f[f];
    if (F.axis != 0) continue;
    if (F.k0 == 0) lo = f;
    else if (F.k0 + F.n == nx) hi = f;
  }
  if (lo < 0 || hi < 0) return false;
  *a0 = pml->f[lo].n;
  *a1 = pml->f[hi].k0;
  return *a0 < *a1;
}

// every face of a CPML set and its x faces' slabs, for the x-column pass
template <typename T>
XCols<T> xcols_args(const mgh::CpmlSet& pml, bool e_side, int ny, int nz, int64_t* threads) {
  XCols<T> x;
  std::memset(&x, 0, sizeof(x));
  x.nf = pml.nfaces;
  int xk0[2] = {0, 0}, xn[2] = {0, 0}, nxf = 0;
  for (int f = 0; f < pml.nfaces; ++f) {
    x.axis[f] = pml.f[f].axis;
    x.f[f] = face_of<T>(pml.f[f], e_side);
    if (pml.f[f].axis == 0) {
      xk0[nxf] = pml.f[f].k0;
      xn[nxf] = pml.f[f].n;
      nxf++;
    }
  }
  if (nxf == 2 && xk0[1] < xk0[0]) {
    std::swap(xk0[0], xk0[1]);
    std::swap(xn[0], xn[1]);
  }
  if (nxf == 1) {  // a lone x slab: lower slot, no upper
    x.lo_k0 = xk0[0];
    x.lo_n = xn[0];
  } else if (nxf == 2) {
    x.lo_k0 = xk0[0];
    x.lo_n = xn[0];
    x.hi_k0 = xk0[1];
    x.hi_n = xn[1];
  }
  *threads = nxf ? (int64_t)(x.lo_n + x.hi_n) * (ny + 2) * (nz + 1) : 0;
  return x;
}

// pml: CPML faces fused into the curl launch (null: curl only)
template <typename T, bool C64>
int run_e(void* e, const void* h, const void* ae, const void* be, int nx, int ny, int nz,
          int nsets, int kchunk, int64_t es, int64_t hs, hipStream_t st, const int* live = nullptr,
          int nlive = 0, const mgh::CpmlSet* pml = nullptr) {
  const Shape sh = decode_shape(kchunk, nsets);
  const int ns = sh.ns;
  int a0 = 0, a1 = 0;
  const bool split = !live && x_split(pml, nx, &a0, &a1);
  const Plan pl = make_plan(nx, ny, nz, nsets, sh.kchunk, ns, sh.band, live, nlive, a0, a1);
  if (pl.units == 0) return MGH_OK;
  PmlArgs<T> pa = pml_args<T>(pml, true, nx, ny, nz);
  pa.xin = split;
  const bool fused = pml && pa.nf > 0;
  const dim3 grd(plan_blocks(pl));
#define MGH_LAUNCH_E(NS_, PML_)                                                               \
  hipLaunchKernelGGL((k_update_e<T, NS_, C64, PML_>), grd, dim3(WG), 0, st, (T*)e, (const T*)h, \
                     (const T*)ae, (const T*)be, pl, nsets, es, hs, pa)
  if (ns == 3) {
    if (fused) MGH_LAUNCH_E(3, true);
    else MGH_LAUNCH_E(3, false);
  } else if (ns == kSetsPerThread) {
    if (fused) MGH_LAUNCH_E(kSetsPerThread, true);
    else MGH_LAUNCH_E(kSetsPerThread, false);
  } else {
    if (fused) MGH_LAUNCH_E(1, true);
    else MGH_LAUNCH_E(1, false);
  }
#undef MGH_LAUNCH_E
  MGH_CHECK_LAUNCH("k_update_e");
  return MGH_OK;
}

template <typename T>
int run_cpml_e(void* e, const void* h, const void* be, int nx, int ny, int nz, int nsets,
               int64_t es, int64_t hs, const mgh::CpmlSet& pml, hipStream_t st) {
  for (int f = 0; f < pml.nfaces; ++f) {
    const mgh::CpmlFace& F = pml.f[f];
    const int64_t nt = slab_threads(F.axis, nx, ny, nz, F.n);
    const dim3 grd((unsigned)((nt + WG - 1) / WG), (unsigned)nsets, 1);
    const CFace<T> c = face_of<T>(F, true);
    if (F.axis == 0)
      hipLaunchKernelGGL((k_cpml_e<T, 0>), grd, dim3(WG), 0, st, (T*)e, (const T*)h, (const T*)be,
                         nx, ny, nz, es, hs, c);
    else if (F.axis == 1)
      hipLaunchKernelGGL((k_cpml_e<T, 1>), grd, dim3(WG), 0, st, (T*)e, (const T*)h, (const T*)be,
                         nx, ny, nz, es, hs, c);
    else
      hipLaunchKernelGGL((k_cpml_e<T, 2>), grd, dim3(WG), 0, st, (T*)e, (const T*)h, (const T*)be,
                         nx, ny, nz, es, hs, c);
    MGH_CHECK_LAUNCH("k_cpml_e");
  }
  return MGH_OK;
}

template <typename T>
int run_h(const void* e, void* h, const void* ah, const void* bh, int nx, int ny, int nz,
          int nsets, int kchunk, int64_t es, int64_t hs, hipStream_t st, const int* live = nullptr,
          int nlive = 0, const mgh::CpmlSet* pml = nullptr) {
  const Shape sh = decode_shape(kchunk, nsets);
  const int ns = sh.ns;
  int a0 = 0, a1 = 0;
  const bool split = !live && x_split(pml, nx, &a0, &a1);
  const Plan pl = make_plan(nx, ny, nz, nsets, sh.kchunk, ns, sh.band, live, nlive, a0, a1);
  if (pl.units == 0) return MGH_OK;
  PmlArgs<T> pa = pml_args<T>(pml, false, nx, ny, nz);
  pa.xin = split;
  const bool fused = pml && pa.nf > 0;
  const dim3 grd(plan_blocks(pl));
#define MGH_LAUNCH_H(NS_, PML_)                                                              \
  hipLaunchKernelGGL((k_update_h<T, NS_, PML_>), grd, dim3(WG), 0, st, (const T*)e, (T*)h,     \
                     (const T*)ah, (const T*)bh, pl, nsets, es, hs, pa)
  if (ns == 3) {
    if (fused) MGH_LAUNCH_H(3, true);
    else MGH_LAUNCH_H(3, false);
  } else if (ns == kSetsPerThread) {
    if (fused) MGH_LAUNCH_H(kSetsPerThread, true);
    else MGH_LAUNCH_H(kSetsPerThread, false);
  } else {
    if (fused) MGH_LAUNCH_H(1, true);
    else MGH_LAUNCH_H(1, false);
  }
#undef MGH_LAUNCH_H
  MGH_CHECK_LAUNCH("k_update_h");
  return MGH_OK;
}

template <typename T>
int run_cpml_h(const void* e, void* h, const void* bh, int nx, int ny, int nz, int nsets,
               int64_t es, int64_t hs, const mgh::CpmlSet& pml, hipStream_t st) {
  for (int f = 0; f < pml.nfaces; ++f) {
    const mgh::CpmlFace& F = pml.f[f];
    const int64_t nt = slab_threads(F.axis, nx, ny, nz, F.n);
    const dim3 grd((unsigned)((nt + WG - 1) / WG), (unsigned)nsets, 1);
    const CFace<T> c = face_of<T>(F, false);
    if (F.axis == 0)
      hipLaunchKernelGGL((k_cpml_h<T, 0>), grd, dim3(WG), 0, st, (const T*)e, (T*)h, (const T*)bh,
                         nx, ny, nz, es, hs, c);
    else if (F.axis == 1)
      hipLaunchKernelGGL((k_cpml_h<T, 1>), grd, dim3(WG), 0, st, (const T*)e, (T*)h, (const T*)bh,
                         nx, ny, nz, es, hs, c);
    else
      hipLaunchKernelGGL((k_cpml_h<T, 2>), grd, dim3(WG), 0, st, (const T*)e, (T*)h, (const T*)bh,
                         nx, ny, nz, es, hs, c);
    MGH_CHECK_LAUNCH("k_cpml_h");
  }
  return MGH_OK;
}

// the x-slab columns of a CPML set: every face, in order (after the curl
// launch that carries the rest of the CPML)
template <typename T>
int run_xcols_e(void* e, const void* h, const void* be, int nx, int ny, int nz, int nsets,
                int64_t es, int64_t hs, const mgh::CpmlSet& pml, hipStream_t st) {
  int a0, a1;
  if (x_split(&pml, nx, &a0, &a1)) return MGH_OK;  // the curl launch took the x slabs
  int64_t nt = 0;
  const XCols<T> X = xcols_args<T>(pml, true, ny, nz, &nt);
  if (nt == 0) return MGH_OK;
  hipLaunchKernelGGL((k_cpml_xcols_e<T>), dim3((unsigned)((nt + WG - 1) / WG), nsets), dim3(WG), 0,
                     st, (T*)e, (const T*)h, (const T*)be, nx, ny, nz, es, hs, X);
  MGH_CHECK_LAUNCH("k_cpml_xcols_e");
  return MGH_OK;
}

template <typename T>
int run_xcols_h(const void* e, void* h, const void* bh, int nx, int ny, int nz, int nsets,
                int64_t es, int64_t hs, const mgh::CpmlSet& pml, hipStream_t st) {
  int a0, a1;
  if (x_split(&pml, nx, &a0, &a1)) return MGH_OK;  // the curl launch took the x slabs
  int64_t nt = 0;
  const XCols<T> X = xcols_args<T>(pml, false, ny, nz, &nt);
  if (nt == 0) return MGH_OK;
  hipLaunchKernelGGL((k_cpml_xcols_h<T>), dim3((unsigned)((nt + WG - 1) / WG), nsets), dim3(WG), 0,
                     st, (const T*)e, (T*)h, (const T*)bh, nx, ny, nz, es, hs, X);
  MGH_CHECK_LAUNCH("k_cpml_xcols_h");
  return MGH_OK;
}

int check_common(int dtype, int nx, int ny, int nz, int nsets, int shape) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 (float32) or 8 (float64), got %d",
              dtype);
  MGH_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "grid must have at least one cell per axis");
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "nsets must be in 1..65535, got %d", nsets);
  MGH_REQUIRE(shape >= 0 && (shape >> 28) == 0, "bad launch shape word %d", shape);
  const int kchunk = shape & 0xffff, ns = (shape >> 16) & 0xf;
  MGH_REQUIRE(kchunk >= 1, "kchunk must be >= 1, got %d", kchunk);
  MGH_REQUIRE(ns <= 3, "sets per thread must be 1, 2 or 3 (0: default), got %d", ns);
  // node numbers, unit counts and slab thread blocks are 32-bit
  const int64_t npl = (int64_t)(nx + 1) * (ny + 1);
  const int64_t units =
      ((npl + WG - 1) / WG) * ((int64_t)(nz + 1 + kchunk - 1) / kchunk) * (int64_t)nsets;
  MGH_REQUIRE(npl < (1ll << 31) && units + 8 * 256 < (1ll << 31),
              "grid too large for the launch plan");
  return MGH_OK;
}

// The packed launch of one phase (E: side 0, H: side 1) over n jobs given as
// host words per job: nx, ny, nz, nsets, E, E set stride, H, H set stride,
// alpha and beta of the phase's side.
template <typename T>
int run_packed(int side, int n, const long long* w, int shape, hipStream_t st) {
  Pack<T> pk;
  std::memset(&pk, 0, sizeof(pk));
  pk.n = n;
  int blocks = 0;
  for (int q = 0; q < n; ++q) {
    const long long* r = w + (int64_t)q * 10;
    const int nx = (int)r[0], ny = (int)r[1], nz = (int)r[2], nsets = (int)r[3];
    int rc = check_common(sizeof(T), nx, ny, nz, nsets, shape);
    if (rc) return rc;
    PackJob<T>& J = pk.j[q];
    J.e = (T*)r[4];
    J.h = (T*)r[6];
    J.a = (const T*)r[8];
    J.b = (const T*)r[9];
    MGH_REQUIRE(J.e && J.h && J.a && J.b, "packed job %d: null device pointer", q);
    J.e_set = r[5];
    J.h_set = r[7];
    J.nsets = nsets;
    const Shape sh = decode_shape(shape, nsets);
    J.pl = make_plan(nx, ny, nz, nsets, sh.kchunk, 1, sh.band);
    blocks = std::max(blocks, plan_blocks(J.pl));
  }
  if (blocks == 0) return MGH_OK;
  if (side == 0)
    hipLaunchKernelGGL((k_update_e_pack<T>), dim3(blocks, n), dim3(WG), 0, st, pk);
  else
    hipLaunchKernelGGL((k_update_h_pack<T>), dim3(blocks, n), dim3(WG), 0, st, pk);
  MGH_CHECK_LAUNCH("k_update_pack");
  return MGH_OK;
}

}  // namespace

extern "C" {

int mgh_update_e(int dtype, int nx, int ny, int nz, int nsets, void* e, long long e_set_stride,
                 const void* h, long long h_set_stride, const void* alpha_e, const void* beta_e,
                 int curl64, int kchunk, void* stream) {
  int rc = check_common(dtype, nx, ny, nz, nsets, kchunk);
  if (rc) return rc;
  MGH_REQUIRE(e && h && alpha_e && beta_e, "null device pointer");
  MGH_REQUIRE(!(curl64 && dtype == 8), "curl64 applies to float32 fields only");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    return curl64 ? run_e<float, true>(e, h, alpha_e, beta_e, nx, ny, nz, nsets, kchunk,
                                        e_set_stride, h_set_stride, st)
                  : run_e<float, false>(e, h, alpha_e, beta_e, nx, ny, nz, nsets, kchunk,
                                         e_set_stride, h_set_stride, st);
  return run_e<double, false>(e, h, alpha_e, beta_e, nx, ny, nz, nsets, kchunk, e_set_stride,
                              h_set_stride, st);
}

int mgh_update_h(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                 long long e_set_stride, void* h, long long h_set_stride, const void* alpha_h,
                 const void* beta_h, int kchunk, void* stream) {
  int rc = check_common(dtype, nx, ny, nz, nsets, kchunk);
  if (rc) return rc;
  MGH_REQUIRE(e && h && alpha_h && beta_h, "null device pointer");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    return run_h<float>(e, h, alpha_h, beta_h, nx, ny, nz, nsets, kchunk, e_set_stride,
                        h_set_stride, st);
  return run_h<double>(e, h, alpha_h, beta_h, nx, ny, nz, nsets, kchunk, e_set_stride,
                       h_set_stride, st);
}

int mgh_update_e_tiles(int dtype, int nx, int ny, int nz, int nsets, void* e,
                       long long e_set_stride, const void* h, long long h_set_stride,
                       const void* alpha_e, const void* beta_e, int curl64, int kchunk,
                       const int* live, int n_live, void* stream) {
  int rc = check_common(dtype, nx, ny, nz, nsets, kchunk);
  if (rc) return rc;
  MGH_REQUIRE(e && h && alpha_e && beta_e && (live || n_live == 0), "null device pointer");
  MGH_REQUIRE(!(curl64 && dtype == 8), "curl64 applies to float32 fields only");
  const int kc = kchunk & 0xffff;
  const int64_t all = (((int64_t)(nx + 1) * (ny + 1) + WG - 1) / WG) * ((nz + kc) / kc);
  MGH_REQUIRE(n_live >= 0 && n_live <= all, "n_live %d outside 0..%lld", n_live, (long long)all);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    return curl64 ? run_e<float, true>(e, h, alpha_e, beta_e, nx, ny, nz, nsets, kchunk,
                                        e_set_stride, h_set_stride, st, live, n_live)
                  : run_e<float, false>(e, h, alpha_e, beta_e, nx, ny, nz, nsets, kchunk,
                                         e_set_stride, h_set_stride, st, live, n_live);
  return run_e<double, false>(e, h, alpha_e, beta_e, nx, ny, nz, nsets, kchunk, e_set_stride,
                              h_set_stride, st, live, n_live);
}

int mgh_update_h_tiles(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                       long long e_set_stride, void* h, long long h_set_stride,
                       const void* alpha_h, const void* beta_h, int kchunk, const int* live,
                       int n_live, void* stream) {
  int rc = check_common(dtype, nx, ny, nz, nsets, kchunk);
  if (rc) return rc;
  MGH_REQUIRE(e && h && alpha_h && beta_h && (live || n_live == 0), "null device pointer");
  const int kc = kchunk & 0xffff;
  const int64_t all = (((int64_t)(nx + 1) * (ny + 1) + WG - 1) / WG) * ((nz + kc) / kc);
  MGH_REQUIRE(n_live >= 0 && n_live <= all, "n_live %d outside 0..%lld", n_live, (long long)all);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    return run_h<float>(e, h, alpha_h, beta_h, nx, ny, nz, nsets, kchunk, e_set_stride,
                        h_set_stride, st, live, n_live);
  return run_h<double>(e, h, alpha_h, beta_h, nx, ny, nz, nsets, kchunk, e_set_stride,
                       h_set_stride, st, live, n_live);
}

int mgh_cpml_e(int dtype, int nx, int ny, int nz, int nsets, void* e, long long e_set_stride,
               const void* h, long long h_set_stride, const void* beta_e, const void* cpml,
               void* stream) {
  int rc = check_common(dtype, nx, ny, nz, nsets, 1);
  if (rc) return rc;
  MGH_REQUIRE(e && h && beta_e && cpml, "null pointer");
  mgh::CpmlSet pml{};
  rc = mgh::load_cpml_set(cpml, nx, ny, nz, nsets, &pml);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    return run_cpml_e<float>(e, h, beta_e, nx, ny, nz, nsets, e_set_stride, h_set_stride, pml, st);
  return run_cpml_e<double>(e, h, beta_e, nx, ny, nz, nsets, e_set_stride, h_set_stride, pml, st);
}

int mgh_cpml_h(int dtype, int nx, int ny, int nz, int nsets, const void* e,
               long long e_set_stride, void* h, long long h_set_stride, const void* beta_h,
               const void* cpml, void* stream) {
  int rc = check_common(dtype, nx, ny, nz, nsets, 1);
  if (rc) return rc;
  MGH_REQUIRE(e && h && beta_h && cpml, "null pointer");
  mgh::CpmlSet pml{};
  rc = mgh::load_cpml_set(cpml, nx, ny, nz, nsets, &pml);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    return run_cpml_h<float>(e, h, beta_h, nx, ny, nz, nsets, e_set_stride, h_set_stride, pml, st);
  return run_cpml_h<double>(e, h, beta_h, nx, ny, nz, nsets, e_set_stride, h_set_stride, pml, st);
}

int mgh_update_e_pml_core(int dtype, int nx, int ny, int nz, int nsets, void* e,
                          long long e_set_stride, const void* h, long long h_set_stride,
                          const void* alpha_e, const void* beta_e, int curl64, int kchunk,
                          const void* cpml, void* stream) {
  if (!cpml)
    return mgh_update_e(dtype, nx, ny, nz, nsets, e, e_set_stride, h, h_set_stride, alpha_e,
                        beta_e, curl64, kchunk, stream);
  int rc = check_common(dtype, nx, ny, nz, nsets, kchunk);
  if (rc) return rc;
  MGH_REQUIRE(e && h && alpha_e && beta_e, "null device pointer");
  MGH_REQUIRE(!(curl64 && dtype == 8), "curl64 applies to float32 fields only");
  mgh::CpmlSet pml{};
  rc = mgh::load_cpml_set(cpml, nx, ny, nz, nsets, &pml);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    return curl64 ? run_e<float, true>(e, h, alpha_e, beta_e, nx, ny, nz, nsets, kchunk,
                                        e_set_stride, h_set_stride, st, nullptr, 0, &pml)
                  : run_e<float, false>(e, h, alpha_e, beta_e, nx, ny, nz, nsets, kchunk,
                                         e_set_stride, h_set_stride, st, nullptr, 0, &pml);
  return run_e<double, false>(e, h, alpha_e, beta_e, nx, ny, nz, nsets, kchunk, e_set_stride,
                              h_set_stride, st, nullptr, 0, &pml);
}

int mgh_cpml_xcols_e(int dtype, int nx, int ny, int nz, int nsets, void* e,
                     long long e_set_stride, const void* h, long long h_set_stride,
                     const void* beta_e, const void* cpml, void* stream) {
  int rc = check_common(dtype, nx, ny, nz, nsets, 1);
  if (rc) return rc;
  MGH_REQUIRE(e && h && beta_e && cpml, "null pointer");
  mgh::CpmlSet pml{};
  rc = mgh::load_cpml_set(cpml, nx, ny, nz, nsets, &pml);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    return run_xcols_e<float>(e, h, beta_e, nx, ny, nz, nsets, e_set_stride, h_set_stride, pml, st);
  return run_xcols_e<double>(e, h, beta_e, nx, ny, nz, nsets, e_set_stride, h_set_stride, pml, st);
}

int mgh_update_e_pml(int dtype, int nx, int ny, int nz, int nsets, void* e,
                     long long e_set_stride, const void* h, long long h_set_stride,
                     const void* alpha_e, const void* beta_e, int curl64, int kchunk,
                     const void* cpml, void* stream) {
  int rc = mgh_update_e_pml_core(dtype, nx, ny, nz, nsets, e, e_set_stride, h, h_set_stride,
                                 alpha_e, beta_e, curl64, kchunk, cpml, stream);
  if (rc || !cpml) return rc;
  return mgh_cpml_xcols_e(dtype, nx, ny, nz, nsets, e, e_set_stride, h, h_set_stride, beta_e,
                          cpml, stream);
}

int mgh_update_h_pml_core(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                          long long e_set_stride, void* h, long long h_set_stride,
                          const void* alpha_h, const void* beta_h, int kchunk, const void* cpml,
                          void* stream) {
  if (!cpml)
    return mgh_update_h(dtype, nx, ny, nz, nsets, e, e_set_stride, h, h_set_stride, alpha_h,
                        beta_h, kchunk, stream);
  int rc = check_common(dtype, nx, ny, nz, nsets, kchunk);
  if (rc) return rc;
  MGH_REQUIRE(e && h && alpha_h && beta_h, "null device pointer");
  mgh::CpmlSet pml{};
  rc = mgh::load_cpml_set(cpml, nx, ny, nz, nsets, &pml);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    return run_h<float>(e, h, alpha_h, beta_h, nx, ny, nz, nsets, kchunk, e_set_stride,
                        h_set_stride, st, nullptr, 0, &pml);
  return run_h<double>(e, h, alpha_h, beta_h, nx, ny, nz, nsets, kchunk, e_set_stride,
                       h_set_stride, st, nullptr, 0, &pml);
}

int mgh_cpml_xcols_h(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                     long long e_set_stride, void* h, long long h_set_stride, const void* beta_h,
                     const void* cpml, void* stream) {
  int rc = check_common(dtype, nx, ny, nz, nsets, 1);
  if (rc) return rc;
  MGH_REQUIRE(e && h && beta_h && cpml, "null pointer");
  mgh::CpmlSet pml{};
  rc = mgh::load_cpml_set(cpml, nx, ny, nz, nsets, &pml);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    return run_xcols_h<float>(e, h, beta_h, nx, ny, nz, nsets, e_set_stride, h_set_stride, pml, st);
  return run_xcols_h<double>(e, h, beta_h, nx, ny, nz, nsets, e_set_stride, h_set_stride, pml, st);
}

int mgh_update_h_pml(int dtype, int nx, int ny, int nz, int nsets, const void* e,
                     long long e_set_stride, void* h, long long h_set_stride,
                     const void* alpha_h, const void* beta_h, int kchunk, const void* cpml,
                     void* stream) {
  int rc = mgh_update_h_pml_core(dtype, nx, ny, nz, nsets, e, e_set_stride, h, h_set_stride,
                                 alpha_h, beta_h, kchunk, cpml, stream);
  if (rc || !cpml) return rc;
  return mgh_cpml_xcols_h(dtype, nx, ny, nz, nsets, e, e_set_stride, h, h_set_stride, beta_h,
                          cpml, stream);
}

// Packed E / H phases of several grids in one launch each (a serving loop's
// small jobs): jobs = n_jobs x 10 words (nx, ny, nz, nsets, E, E set stride,
// H, H set stride, alpha and beta of the phase's side), shape = the launch
// shape word (its sets-per-thread part must be 0 or 1: one set per thread).
int mgh_update_e_packed(int dtype, int n_jobs, const long long* jobs, int shape, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 (float32) or 8 (float64), got %d",
              dtype);
  MGH_REQUIRE(n_jobs >= 1 && n_jobs <= PACK_MAX, "n_jobs must be 1..%d, got %d", PACK_MAX,
              n_jobs);
  MGH_REQUIRE(jobs, "null job table");
  MGH_REQUIRE(((shape >> 16) & 0xf) <= 1, "packed launches take one set per thread");
  hipStream_t st = (hipStream_t)stream;
  return dtype == 4 ? run_packed<float>(0, n_jobs, jobs, shape, st)
                    : run_packed<double>(0, n_jobs, jobs, shape, st);
}

int mgh_update_h_packed(int dtype, int n_jobs, const long long* jobs, int shape, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 (float32) or 8 (float64), got %d",
              dtype);
  MGH_REQUIRE(n_jobs >= 1 && n_jobs <= PACK_MAX, "n_jobs must be 1..%d, got %d", PACK_MAX,
              n_jobs);
  MGH_REQUIRE(jobs, "null job table");
  MGH_REQUIRE(((shape >> 16) & 0xf) <= 1, "packed launches take one set per thread");
  hipStream_t st = (hipStream_t)stream;
  return dtype == 4 ? run_packed<float>(1, n_jobs, jobs, shape, st)
                    : run_packed<double>(1, n_jobs, jobs, shape, st);
}

}  // extern "C"
