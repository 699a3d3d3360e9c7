// On-the-fly DFT monitors: per-step field sampling into a 64-step block, and
// the block contraction against a phasor matrix on the MFMA units with
// float64 (optionally compensated) bin accumulation.
//
// Reference: DFTAccumulator.accumulate (src/magnelio/monitors/_dft.py:37-52)
//   bins[f, s] += exp(+j w_f t_n) dt * x_n[s]
// with cell-centred physical samples (monitors/base.py:51-173) taken after
// the H update of step n, E stamped t = n dt and H stamped t + dt/2
// (field_frequency.py:249-294, far_field.py:230-248).
//
// Factorisation used here: for a block of B steps starting at n0,
//   sum_b exp(j w (n0 + b) dt) dt x_b = exp(j w t0) * sum_b P[f, b] x_b,
//   P[f, b] = exp(j w_f b dt) dt   (float32, host-built),
//   t0 = n0 dt (E samples) or n0 dt + dt/2 (H samples).
// The inner sum is a (2F x B) x (B x S) product done with
// v_mfma_f32_16x16x4f32 (16 frequencies x 16 samples per wave, K = B in B/4
// MFMA steps, B up to 512); the outer factor exp(j w t0) is float64 and the
// update of the float64 bins is fused into the MFMA epilogue, so the partial
// products never reach HBM.  Each wave keeps its 16-sample x B-step operand
// in registers and sweeps every frequency, so x is read once per block and
// the bins are read and written once per launch: the kernel is bound by the
// bin traffic (16 B complex128 per (f, s), + 8 B of float32 compensation), which a
// longer block divides — or several consecutive blocks per launch, each its
// own contraction and phase, summed before the single bin update.
#include "mgh_common.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int DFT_NT = 8;      // terms per sample (cell-centre averaging)
constexpr int DFT_BMAX = 512;  // longest block (MFMA K)

// phasor-table row length for a block of B steps: a whole number of 64-step
// register tiles up to 256, 512 beyond
inline int dft_bpad(int B) { return B <= 256 ? (B + 63) / 64 * 64 : 512; }

template <typename T>
__global__ __launch_bounds__(256) void k_dft_sample(const T* __restrict__ e, int64_t e_set,
                                                    const T* __restrict__ h, int64_t h_set,
                                                    const int* __restrict__ idx,
                                                    const float* __restrict__ w, int S,
                                                    float* __restrict__ X, int S_pad, int B,
                                                    const long long* __restrict__ ctr,
                                                    long long n_start) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int set = blockIdx.y;
  if (s >= S) return;
  const long long n = *ctr;
  const int b = (int)((n - n_start) % B);
  const T* es = e + (int64_t)set * e_set;
  const T* hs = h + (int64_t)set * h_set;
  float acc = 0.0f;
#pragma unroll
  for (int t = 0; t < DFT_NT; ++t) {
    const int id = idx[(int64_t)s * DFT_NT + t];
    const float wt = w[(int64_t)s * DFT_NT + t];
    if (wt != 0.0f) {
      const float v = id >= 0 ? (float)es[id] : (float)hs[-(id + 1)];
      acc += wt * v;
    }
  }
  X[((int64_t)set * B + b) * S_pad + s] = acc;
}

// Structured sampling: the samples of one monitor component on one face or
// box form a group whose terms are affine in the sample's natural
// coordinates (c0 fastest): term t reads field[base_t + c0 s0 + c1 s1 +
// c2 s2] with weight wtab[woff_t + c_{waxis_t}] (one coordinate, or a
// constant when waxis_t < 0).  The weights and indices are those of the
// generic map (the host derives the groups from it and checks them), so the
// sum — taps in map order, zero weights skipped — is the same float32 value;
// only the 64 B of per-sample index/weight words per step are gone.
//
// Each group owns a range of work-groups.  A group whose c0 is unit-stride
// in memory (faces normal to y or z, boxes) maps 256 consecutive samples to
// a work-group, so every tap load is one contiguous row.  A group whose
// fastest coordinate is strided (faces normal to x: c0 = y, c1 = z) maps a
// 16 x 16 patch of (c0, c1) to a work-group instead: each sample's taps sit
// in the rows of its own and the next c0 / c1 neighbours, which then belong
// to the same work-group and are fetched once, not once per neighbour.
constexpr int DG_W = 40;  // int32 words per group descriptor

template <typename T>
__global__ __launch_bounds__(256) void k_dft_sample_groups(
    const T* __restrict__ e, int64_t e_set, const T* __restrict__ h, int64_t h_set,
    const int* __restrict__ grp, int G, const float* __restrict__ wtab, float* __restrict__ X,
    int S_pad, int B, const long long* __restrict__ ctr, long long n_start) {
  const int blk = blockIdx.x;
  const int set = blockIdx.y;
  // group of this work-group: the last whose first block is <= blk (uniform)
  int g = 0;
  while (g + 1 < G && grp[(g + 1) * DG_W + 33] <= blk) ++g;
  const int* d = grp + g * DG_W;
  const int n0 = d[2], n1 = d[3];
  const int n2 = d[1] / (n0 * n1);
  const int wb = blk - d[33];  // work-group within the group
  int c0, c1, c2;
  if (d[34] == 0) {  // linear: 256 consecutive samples
    const int loc = wb * 256 + (int)threadIdx.x;
    if (loc >= d[1]) return;
    c0 = loc % n0;
    const int r = loc / n0;
    c1 = r % n1;
    c2 = r / n1;
  } else {  // 16 x 16 patches of (c0, c1)
    const int p0 = (n0 + 15) >> 4, p1 = (n1 + 15) >> 4;
    const int tc0 = wb % p0;
    const int r = wb / p0;
    const int tc1 = r % p1;
    c2 = r / p1;
    c0 = tc0 * 16 + (int)(threadIdx.x & 15);
    c1 = tc1 * 16 + (int)(threadIdx.x >> 4);
    if (c0 >= n0 || c1 >= n1 || c2 >= n2) return;
  }
  const long long n = *ctr;
  const int b = (int)((n - n_start) % B);
  const int64_t off = (int64_t)c0 * d[6] + (int64_t)c1 * d[7] + (int64_t)c2 * d[8];
  const T* f = d[4] ? h + (int64_t)set * h_set : e + (int64_t)set * e_set;
  const int nt = d[5];
  // all tap weights, then all live taps' field words, then the sum in tap
  // order: two memory round trips per sample instead of one per tap
  float wv[DFT_NT], fv[DFT_NT];
#pragma unroll
  for (int t = 0; t < DFT_NT; ++t) {
    wv[t] = 0.0f;
    if (t < nt) {
      const int ax = d[17 + t];
      const int c = ax == 0 ? c0 : (ax == 1 ? c1 : (ax == 2 ? c2 : 0));
      wv[t] = wtab[d[25 + t] + c];
    }
  }
#pragma unroll
  for (int t = 0; t < DFT_NT; ++t) {
    fv[t] = 0.0f;
    if (wv[t] != 0.0f) fv[t] = (float)f[(int64_t)d[9 + t] + off];
  }
  float acc = 0.0f;
#pragma unroll
  for (int t = 0; t < DFT_NT; ++t)
    if (wv[t] != 0.0f) acc += wv[t] * fv[t];
  const int s = d[0] + c0 + n0 * (c1 + n1 * c2);
  X[((int64_t)set * B + b) * S_pad + s] = acc;
}

// Block phase factors exp(j w t0) for E and H stamps of each of the nsub
// consecutive B-step blocks a contraction covers: blk[(q*F + f)*4 + 0..3] =
// cos(w t0E), sin(w t0E), cos(w t0H), sin(w t0H) of block q.
__global__ void k_dft_phase(const double* __restrict__ omega, int F, double dt,
                            const long long* __restrict__ ctr, long long n_start, int B,
                            int nsub, double* __restrict__ blk) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (f >= F) return;
  const long long n = *ctr;
  const long long n0 = n - ((n - n_start) % ((long long)B * nsub)) + (long long)q * B;
  const double te = (double)n0 * dt;
  const double th = te + 0.5 * dt;
  double s, c;
  double* o = blk + ((int64_t)q * F + f) * 4;
  sincos(omega[f] * te, &s, &c);
  o[0] = c;
  o[1] = s;
  sincos(omega[f] * th, &s, &c);
  o[2] = c;
  o[3] = s;
}

// TwoSum of the float64 bin and the block's contribution; the rounding error
// (at most half an ulp of the bin) joins the float32 compensation word.
// Float32 holds it with 24 bits to spare over what the float32 block partials
// carry, at 4 of the 8 bytes a float64 word would add to the bin traffic.
__device__ __forceinline__ void two_sum_acc(double* hi, float* lo, double v) {
  const double a = *hi;
  const double s = a + v;
  const double bp = s - a;
  const double err = (a - (s - bp)) + (v - bp);
  *hi = s;
  *lo = (float)((double)*lo + err);
}

// NB: MFMA k-steps per block (the padded block length divided by 4).
// COMP: TwoSum-compensated bins.
// NSUB: consecutive blocks per launch (rows q*B .. q*B + B - 1 of X are block
// q).  Each block is its own MFMA contraction against the same phasor tile
// with its own phase factor; their float64 contributions are summed in
// registers and the bins are read and written once for all of them, which
// divides the bin traffic — the bound of short blocks — by NSUB.
//
// The four waves of a work-group take 64 consecutive samples and share each
// frequency tile's phasor operands: the work-group copies the tile (cos and
// sin, NB x 64 lanes) from L2 into LDS, at most 256 steps at a time, and
// every wave reads its lanes' A operands from there (16-byte reads,
// lane-contiguous, no bank conflicts) — a quarter of the L2 traffic of
// per-wave loads.  A 512-step block takes the tile in two passes into the
// same float32 accumulators; it halves the bin traffic per step.  The bins of
// the tile are loaded before the MFMA chain so their HBM latency hides
// behind the matrix work, and are read and written as complex pairs.  Up to
// 256 steps of samples per lane (NB x NSUB <= 64) the kernel is held to 128
// VGPRs (4 waves per SIMD, no spills; 3 without the bound); a 512-step block
// needs ~200 and runs at 2.
template <int NB, bool COMP, int NSUB>
__global__ __launch_bounds__(256, (NSUB > 1 ? 3 : (NB <= 64 ? 4 : 1))) void k_dft_block(
    const float* __restrict__ X, int S, int S_pad, int B, int nb_valid,
    const float* __restrict__ Pc, const float* __restrict__ Ps, int F,
    const double* __restrict__ blk, int s_h_start, double* __restrict__ bins,
    float* __restrict__ lo) {
  constexpr int NQ = NB / 4;                // float4 groups per lane
  constexpr int NQP = NQ < 16 ? NQ : 16;    // groups per LDS pass (256 steps)
  constexpr int NP = NQ / NQP;              // passes
  static_assert(NQ % NQP == 0, "block tile must be whole passes");
  static_assert(NSUB == 1 || NP == 1, "several blocks per launch need a one-pass tile");
  __shared__ float4 sc[NQP * 64];
  __shared__ float4 ss[NQP * 64];
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int set = blockIdx.y;
  const int s0 = wave * 16;
  const bool active = s0 < S;  // idle waves still take part in the tile copies
  const int col = lane & 15;   // sample within the tile (B operand column / D column)
  const int krow = lane >> 4;  // 0..3: k within an MFMA step (A/B), row group of D
  const float* Xs = X + (int64_t)set * B * NSUB * S_pad;
  float xb[NB * NSUB];
#pragma unroll
  for (int q = 0; q < NSUB; ++q) {
#pragma unroll
    for (int kk = 0; kk < NB; ++kk) {
      const int b = 4 * kk + krow;
      const int row = q * B + b;
      // (one block: nb_valid <= B already bounds b)
      const bool in = NSUB == 1 ? b < nb_valid : (b < B && row < nb_valid);
      xb[q * NB + kk] = (active && in)
                            ? Xs[(int64_t)row * S_pad + s0 + col]
                            : 0.0f;
    }
  }
  const int ph = (s0 >= s_h_start) ? 2 : 0;
  const int s = s0 + col;
  const bool own = active && s < S;
  double2* bset = (double2*)(bins + (int64_t)set * F * S * 2);
  float2* lset = COMP ? (float2*)(lo + (int64_t)set * F * S * 2) : nullptr;
  for (int f0 = 0; f0 < F; f0 += 16) {
    double2 bv[4];
    float2 lv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int f = f0 + 4 * krow + r;
      if (own && f < F) {
        const int64_t o = (int64_t)f * S + s;
        bv[r] = bset[o];
        if (COMP) lv[r] = lset[o];
      }
    }
    // this tile's phasor operands, [q][lane] float4s, into LDS (pass by pass)
    const float4* gc = (const float4*)Pc + (int64_t)(f0 >> 4) * NQ * 64;
    const float4* gs = (const float4*)Ps + (int64_t)(f0 >> 4) * NQ * 64;
    double2 dv[4];  // running sum of the earlier blocks' contributions
#pragma unroll
    for (int q = 0; q < NSUB; ++q) {
      v4f cc = {0.f, 0.f, 0.f, 0.f};
      v4f cs = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int pass = 0; pass < NP; ++pass) {
        if (q == 0) {  // one tile serves every block of the launch
          __syncthreads();
          for (int g = threadIdx.x; g < NQP * 64; g += 256) {
            sc[g] = gc[pass * NQP * 64 + g];
            ss[g] = gs[pass * NQP * 64 + g];
          }
          __syncthreads();
        }
        if (active) {
          const float* xq = xb + q * NB;
#pragma unroll
          for (int u = 0; u < NQP; ++u) {
            const int qq = pass * NQP + u;
            const float4 ac = sc[u * 64 + lane];
            const float4 as = ss[u * 64 + lane];
            cc = __builtin_amdgcn_mfma_f32_16x16x4f32(ac.x, xq[4 * qq + 0], cc, 0, 0, 0);
            cs = __builtin_amdgcn_mfma_f32_16x16x4f32(as.x, xq[4 * qq + 0], cs, 0, 0, 0);
            cc = __builtin_amdgcn_mfma_f32_16x16x4f32(ac.y, xq[4 * qq + 1], cc, 0, 0, 0);
            cs = __builtin_amdgcn_mfma_f32_16x16x4f32(as.y, xq[4 * qq + 1], cs, 0, 0, 0);
            cc = __builtin_amdgcn_mfma_f32_16x16x4f32(ac.z, xq[4 * qq + 2], cc, 0, 0, 0);
            cs = __builtin_amdgcn_mfma_f32_16x16x4f32(as.z, xq[4 * qq + 2], cs, 0, 0, 0);
            cc = __builtin_amdgcn_mfma_f32_16x16x4f32(ac.w, xq[4 * qq + 3], cc, 0, 0, 0);
            cs = __builtin_amdgcn_mfma_f32_16x16x4f32(as.w, xq[4 * qq + 3], cs, 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int f = f0 + 4 * krow + r;
        if (own && f < F) {
          const double pre = (double)cc[r];
          const double pim = (double)cs[r];
          const double c = blk[(q * F + f) * 4 + ph];
          const double sn = blk[(q * F + f) * 4 + ph + 1];
          const double dre = c * pre - sn * pim;
          const double dim = sn * pre + c * pim;
          if (q + 1 < NSUB) {  // an earlier block: into the running sum
            dv[r].x = q == 0 ? dre : dv[r].x + dre;
            dv[r].y = q == 0 ? dim : dv[r].y + dim;
          } else {  // the last block: the launch's contribution into the bins
            const double tre = q == 0 ? dre : dv[r].x + dre;
            const double tim = q == 0 ? dim : dv[r].y + dim;
            const int64_t o = (int64_t)f * S + s;
            if (COMP) {
              two_sum_acc(&bv[r].x, &lv[r].x, tre);
              two_sum_acc(&bv[r].y, &lv[r].y, tim);
              lset[o] = lv[r];
            } else {
              bv[r].x += tre;
              bv[r].y += tim;
            }
            bset[o] = bv[r];
          }
        }
      }
    }
  }
}

}  // namespace

extern "C" {

// Sample one step into row ((n - n_start) mod B) of the block buffer
// X (nsets x B x S_pad float32).  idx/w: S x 8 terms; idx >= 0 indexes the
// E vector, idx < 0 the H vector at -(idx + 1) (device layout); weight 0
// marks an unused term.
int mgh_dft_sample(int dtype, int nsets, const void* e, long long e_set_stride, const void* h,
                   long long h_set_stride, const int* idx, const float* w, int S, float* X,
                   int S_pad, int B, const long long* step_ctr, long long n_start, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8");
  MGH_REQUIRE(S >= 1 && S_pad >= S && S_pad % 16 == 0, "bad sample counts S=%d S_pad=%d", S, S_pad);
  MGH_REQUIRE(B >= 1 && B <= DFT_BMAX, "block length must be 1..512, got %d", B);
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "bad nsets");
  MGH_REQUIRE(e && h && idx && w && X && step_ctr, "null pointer");
  dim3 grd((S + 255) / 256, nsets, 1);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    hipLaunchKernelGGL(k_dft_sample<float>, grd, dim3(256), 0, st, (const float*)e, e_set_stride,
                       (const float*)h, h_set_stride, idx, w, S, X, S_pad, B, step_ctr, n_start);
  else
    hipLaunchKernelGGL(k_dft_sample<double>, grd, dim3(256), 0, st, (const double*)e, e_set_stride,
                       (const double*)h, h_set_stride, idx, w, S, X, S_pad, B, step_ctr, n_start);
  MGH_CHECK_LAUNCH("k_dft_sample");
  return MGH_OK;
}

// Structured form of mgh_dft_sample (same X row, same values): grp holds G
// descriptors of DG_W int32 words, sorted by start:
//   [0] start  [1] count  [2] n0  [3] n1  [4] field (0 E, 1 H)  [5] terms
//   [6..8] index strides of c0, c1, c2  [9..16] base index per term
//   [17..24] weight coordinate per term (0..2, or -1: constant)
//   [25..32] offset of the term's weight table in wtab
//   [33] first work-group of the group  [34] mapping (0 linear, 1 16 x 16
//   patches of (c0, c1))  [35] work-groups of the group
// n_blocks: the work-groups of all groups.  Samples outside every group are
// not written (the X buffer holds zeros there from its allocation).
int mgh_dft_sample_groups(int dtype, int nsets, const void* e, long long e_set_stride,
                          const void* h, long long h_set_stride, const int* grp, int G,
                          int n_blocks, const float* wtab, int S, float* X, int S_pad, int B,
                          const long long* step_ctr, long long n_start, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8");
  MGH_REQUIRE(S >= 1 && S_pad >= S && S_pad % 16 == 0, "bad sample counts S=%d S_pad=%d", S, S_pad);
  MGH_REQUIRE(B >= 1 && B <= DFT_BMAX, "block length must be 1..512, got %d", B);
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "bad nsets");
  MGH_REQUIRE(G >= 1 && G <= 4096, "bad group count %d", G);
  MGH_REQUIRE(n_blocks >= 1, "bad work-group count %d", n_blocks);
  MGH_REQUIRE(e && h && grp && wtab && X && step_ctr, "null pointer");
  dim3 grd((unsigned)n_blocks, nsets, 1);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    hipLaunchKernelGGL(k_dft_sample_groups<float>, grd, dim3(256), 0, st, (const float*)e,
                       e_set_stride, (const float*)h, h_set_stride, grp, G, wtab, X, S_pad, B,
                       step_ctr, n_start);
  else
    hipLaunchKernelGGL(k_dft_sample_groups<double>, grd, dim3(256), 0, st, (const double*)e,
                       e_set_stride, (const double*)h, h_set_stride, grp, G, wtab, X, S_pad, B,
                       step_ctr, n_start);
  MGH_CHECK_LAUNCH("k_dft_sample_groups");
  return MGH_OK;
}

// Contract nsub consecutive B-step blocks (X: nsets x nsub*B x S_pad, rows
// q*B .. q*B + B - 1 are block q; the first nb_valid rows hold samples) and
// accumulate into the float64 bins (nsets x F x S complex, interleaved
// re/im) once; lo (same shape or NULL; float32) holds the compensation
// terms.  nsub is 1, 2 or 4, with B rounded up to a multiple of 64 times nsub
// at most 256.  Pc/Ps: the float32 phasor tables
// P[f][b] = cos/sin(w_f b dt) dt (f < F_pad, b < Bp = B rounded up to a
// multiple of 64, zero beyond B and F) stored in MFMA lane order, in float4
// groups: element [t][q][l][c] (t < F_pad/16, q < Bp/16, l < 64, c < 4) =
// P[16t + l%16][4(4q + c) + l/16] — a tile is a lane-contiguous block that
// a work-group copies straight into LDS.  Samples with index >= s_h_start carry the H time stamp.
// blk: nsub x F x 4 float64 scratch for the block phases, filled here from
// *step_ctr (n = last step of the last block, or one past it for a partial one).
int mgh_dft_blocks(int nsets, const float* X, int S, int S_pad, int B, int nsub, int nb_valid,
                   const float* Pc, const float* Ps, const double* omega, int F, int F_pad,
                   double dt, double* blk, int s_h_start, double* bins, float* lo,
                   const long long* step_ctr, long long n_start, void* stream) {
  MGH_REQUIRE(S >= 1 && S_pad >= S && S_pad % 16 == 0, "bad sample counts");
  MGH_REQUIRE(F >= 1 && F_pad >= F && F_pad % 16 == 0, "bad frequency counts F=%d F_pad=%d", F,
              F_pad);
  MGH_REQUIRE(B >= 1 && B <= DFT_BMAX, "bad block length %d", B);
  MGH_REQUIRE(nsub == 1 || nsub == 2 || nsub == 4, "blocks per launch must be 1, 2 or 4, got %d",
              nsub);
  MGH_REQUIRE(nsub == 1 || dft_bpad(B) * nsub <= 256,
              "%d blocks of %d steps exceed 256 sample rows per lane", nsub, B);
  MGH_REQUIRE(nb_valid >= 0 && nb_valid <= B * nsub, "bad valid row count %d", nb_valid);
  MGH_REQUIRE(X && Pc && Ps && omega && blk && bins && step_ctr, "null pointer");
  MGH_REQUIRE(s_h_start % 16 == 0, "s_h_start must be a multiple of 16");
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "bad nsets");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dft_phase, dim3((F + 255) / 256, nsub), dim3(256), 0, st, omega, F, dt,
                     step_ctr, n_start, B, nsub, blk);
  MGH_CHECK_LAUNCH("k_dft_phase");
  const int waves = (S + 15) / 16;
  dim3 grd((waves + 3) / 4, nsets, 1);
  const int bp = dft_bpad(B);
#define MGH_DFT_LAUNCH(NB, NSUB)                                                             \
  do {                                                                                       \
    if (lo)                                                                                  \
      hipLaunchKernelGGL((k_dft_block<NB, true, NSUB>), grd, dim3(256), 0, st, X, S, S_pad, \
                         B, nb_valid, Pc, Ps, F, (const double*)blk, s_h_start, bins, lo);   \
    else                                                                                     \
      hipLaunchKernelGGL((k_dft_block<NB, false, NSUB>), grd, dim3(256), 0, st, X, S,       \
                         S_pad, B, nb_valid, Pc, Ps, F, (const double*)blk, s_h_start, bins, \
                         lo);                                                                \
  } while (0)
  if (nsub == 4) MGH_DFT_LAUNCH(16, 4);
  else if (nsub == 2 && bp == 64) MGH_DFT_LAUNCH(16, 2);
  else if (nsub == 2) MGH_DFT_LAUNCH(32, 2);
  else if (bp == 64) MGH_DFT_LAUNCH(16, 1);
  else if (bp == 128) MGH_DFT_LAUNCH(32, 1);
  else if (bp == 192) MGH_DFT_LAUNCH(48, 1);
  else if (bp == 256) MGH_DFT_LAUNCH(64, 1);
  else MGH_DFT_LAUNCH(128, 1);
#undef MGH_DFT_LAUNCH
  MGH_CHECK_LAUNCH("k_dft_block");
  return MGH_OK;
}

int mgh_dft_block(int nsets, const float* X, int S, int S_pad, int B, int nb_valid,
                  const float* Pc, const float* Ps, const double* omega, int F, int F_pad,
                  double dt, double* blk, int s_h_start, double* bins, float* lo,
                  const long long* step_ctr, long long n_start, void* stream) {
  return mgh_dft_blocks(nsets, X, S, S_pad, B, 1, nb_valid, Pc, Ps, omega, F, F_pad, dt, blk,
                        s_h_start, bins, lo, step_ctr, n_start, stream);
}

}  // extern "C"
