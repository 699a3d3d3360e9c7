// Modal-port descriptors shared by the port-step kernel.
//
// The host passes MGH_PORT_WORDS int64 words per port (plain words across the
// C-ABI; device pointers as integers):
//    0 n_u, 1 n_v, 2 n_modes M (<= 8)
//    3 idx_port, 4 idx_int, 5 idx_h      int64 device-layout indices, n = n_u+n_v
//                                         ([u edges | v edges]; h: [h_v | h_u])
//    6 me_port, 7 me_int, 8 mh           float64 masses, n each
//    9 prof_e, 10 prof_h                 float64 (M, n) profiles
//   11 params                            float64 (M, 8): kind (1 = DTBC, 0 = Mur),
//                                         r^2, q^2, mur_r, record_scale,
//                                         2 - 2r^2 - q^2, 0, 0
//   12 kern, 13 kern_stride, 14 k_init   float64 packed kernel versions per mode
//   15 state                             float64 (nsets, M, 4): u, u_prev,
//                                         V_port_prev, V_int_prev
//   16 w_hist, 17 s_hist, 18 cap         float64 (nsets, M, cap) histories
//   19 exc_mode                          int32 (nsets): excited mode or -1; 0 = none
//   20 src_tab, 21 tab_len               float64 (nsets, tab_len) DTBC ghost source
//   22 mur_tab                           float64 (nsets, tab_len, 4) or 0
//   23 V_rec, 24 I_rec, 25 rec_cap       float64 (nsets, M, rec_cap) or 0
//   26 absmax                            float64 (nsets) |V| envelope or 0
//   27 work, 28 work_len                 float64 scratch of the split step, at least
//                                         mgh::port_work_len(n, M, cap, nsets) words
//   29 comp_r, 30 comp_live              float64 (n) complement absorber: per-edge Mur
//                                         coefficient and 0/1 live mask, or 0 (off)
//   31 comp_state                        float64 (nsets, 2, n): interior remainder and
//                                         written remainder of the previous step
#pragma once

#include "mgh_common.h"

#define MGH_PORT_WORDS 32
#define MGH_PORT_MAX_PORTS 8
#define MGH_PORT_MAX_MODES 8
// points of a port plane (and DTBC history terms) per work-group chunk
#define MGH_PORT_CHUNK 1024
// partial words per (set, chunk): 2 per mode (u/v part) then 2 per mode at +16
#define MGH_PORT_PQ (4 * MGH_PORT_MAX_MODES)

namespace mgh {

struct PortDev {
  int n_u, n_v, M, pad;
  const long long* idx_port;
  const long long* idx_int;
  const long long* idx_h;
  const double* me_port;
  const double* me_int;
  const double* mh;
  const double* prof_e;
  const double* prof_h;
  const double* params;
  const double* kern;
  long long kern_stride, k_init;
  double* state;
  double* w_hist;
  double* s_hist;
  long long cap;
  const int* exc_mode;
  const double* src_tab;
  long long tab_len;
  const double* mur_tab;
  double* V_rec;
  double* I_rec;
  long long rec_cap;
  double* absmax;
  double* work;
  long long work_len;
  const double* comp_r;
  const double* comp_live;
  double* comp_state;
  // the field sets this port acts on (set by the step call, not the
  // descriptor): E and H of set 0 and their set strides, the set count and
  // the step counter of the solver the port belongs to
  void* e;
  const void* h;
  long long e_set, h_set;
  const long long* ctr;
  int nsets, pad2;
};

// chunks of the projection pass: the plane and the history, whichever is longer
__host__ __device__ inline int port_chunks(int n_pts, long long cap) {
  const long long a = (n_pts + MGH_PORT_CHUNK - 1) / MGH_PORT_CHUNK;
  const long long b = (cap + MGH_PORT_CHUNK - 1) / MGH_PORT_CHUNK;
  const long long c = a > b ? a : b;
  return (int)(c > 0 ? c : 1);
}
__host__ __device__ inline int port_plane_chunks(int n_pts) {
  return (n_pts + MGH_PORT_CHUNK - 1) / MGH_PORT_CHUNK;
}
// scratch words: projection partials, record partials, corrected voltages
__host__ __device__ inline long long port_work_len(int n_pts, long long cap, int nsets) {
  return (long long)nsets * (2LL * port_chunks(n_pts, cap) * MGH_PORT_PQ + MGH_PORT_MAX_MODES);
}

struct PortSet {
  int n;
  int pad;
  PortDev p[MGH_PORT_MAX_PORTS];
};

inline int load_port(const long long* w, PortDev* P, int nsets) {
  P->n_u = (int)w[0];
  P->n_v = (int)w[1];
  P->M = (int)w[2];
  MGH_REQUIRE(P->n_u >= 0 && P->n_v >= 0 && P->n_u + P->n_v >= 1, "port: empty plane");
  MGH_REQUIRE(P->M >= 1 && P->M <= MGH_PORT_MAX_MODES, "port: %d modes (1..%d)", P->M,
              MGH_PORT_MAX_MODES);
  P->idx_port = (const long long*)w[3];
  P->idx_int = (const long long*)w[4];
  P->idx_h = (const long long*)w[5];
  P->me_port = (const double*)w[6];
  P->me_int = (const double*)w[7];
  P->mh = (const double*)w[8];
  P->prof_e = (const double*)w[9];
  P->prof_h = (const double*)w[10];
  P->params = (const double*)w[11];
  P->kern = (const double*)w[12];
  P->kern_stride = w[13];
  P->k_init = w[14];
  P->state = (double*)w[15];
  P->w_hist = (double*)w[16];
  P->s_hist = (double*)w[17];
  P->cap = w[18];
  P->exc_mode = (const int*)w[19];
  P->src_tab = (const double*)w[20];
  P->tab_len = w[21];
  P->mur_tab = (const double*)w[22];
  P->V_rec = (double*)w[23];
  P->I_rec = (double*)w[24];
  P->rec_cap = w[25];
  P->absmax = (double*)w[26];
  P->work = (double*)w[27];
  P->work_len = w[28];
  P->comp_r = (const double*)w[29];
  P->comp_live = (const double*)w[30];
  P->comp_state = (double*)w[31];
  MGH_REQUIRE((P->comp_r == nullptr) == (P->comp_live == nullptr) &&
                  (P->comp_r == nullptr) == (P->comp_state == nullptr),
              "port: complement absorber needs r, live and state together");
  MGH_REQUIRE(P->idx_port && P->idx_int && P->idx_h && P->me_port && P->me_int && P->mh &&
                  P->prof_e && P->prof_h && P->params && P->state,
              "port: null array pointer");
  MGH_REQUIRE(P->k_init >= 2 || P->kern == nullptr, "port: bad kernel length");
  MGH_REQUIRE(P->cap >= 0 && P->tab_len >= 0 && P->rec_cap >= 0, "port: negative length");
  MGH_REQUIRE(!P->V_rec || P->I_rec, "port: V_rec without I_rec");
  MGH_REQUIRE(P->work && P->work_len >= port_work_len(P->n_u + P->n_v, P->cap, nsets),
              "port: work area of %lld words, %lld needed", P->work_len,
              port_work_len(P->n_u + P->n_v, P->cap, nsets));
  return MGH_OK;
}

}  // namespace mgh
