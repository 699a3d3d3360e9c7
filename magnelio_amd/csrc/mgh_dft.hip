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
// the bins are read and written once per block: the kernel is bound by the
// bin traffic (16 B complex128 per (f, s), + 8 B of float32 compensation), which a
// longer block divides.
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

// Block phase factors exp(j w t0) for E and H stamps: blk[f*4 + 0..3] =
// cos(w t0E), sin(w t0E), cos(w t0H), sin(w t0H).
__global__ void k_dft_phase(const double* __restrict__ omega, int F, double dt,
                            const long long* __restrict__ ctr, long long n_start, int B,
                            double* __restrict__ blk) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const long long n = *ctr;
  const long long n0 = n - ((n - n_start) % B);
  const double te = (double)n0 * dt;
  const double th = te + 0.5 * dt;
  double s, c;
  sincos(omega[f] * te, &s, &c);
  blk[f * 4 + 0] = c;
  blk[f * 4 + 1] = s;
  sincos(omega[f] * th, &s, &c);
  blk[f * 4 + 2] = c;
  blk[f * 4 + 3] = s;
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
//
// The four waves of a work-group take 64 consecutive samples and share each
// frequency tile's phasor operands: the work-group copies the tile (cos and
// sin, NB x 64 lanes) from L2 into LDS, at most 256 steps at a time, and
// every wave reads its lanes' A operands from there (16-byte reads,
// lane-contiguous, no bank conflicts) — a quarter of the L2 traffic of
// per-wave loads.  A 512-step block takes the tile in two passes into the
// same float32 accumulators; it halves the bin traffic per step.  The bins of
// the tile are loaded before the MFMA chain so their HBM latency hides
// behind the matrix work, and are read and written as complex pairs.  Blocks
// of up to 256 steps are held to 128 VGPRs (4 waves per SIMD, no spills; 3
// without the bound); a 512-step block needs ~200 and runs at 2.
template <int NB, bool COMP>
__global__ __launch_bounds__(256, (NB <= 64 ? 4 : 1)) void k_dft_block(const float* __restrict__ X, int S, int S_pad,
                                                   int B, int nb_valid,
                                                   const float* __restrict__ Pc,
                                                   const float* __restrict__ Ps, int F,
                                                   const double* __restrict__ blk, int s_h_start,
                                                   double* __restrict__ bins,
                                                   float* __restrict__ lo) {
  constexpr int NQ = NB / 4;                // float4 groups per lane
  constexpr int NQP = NQ < 16 ? NQ : 16;    // groups per LDS pass (256 steps)
  constexpr int NP = NQ / NQP;              // passes
  static_assert(NQ % NQP == 0, "block tile must be whole passes");
  __shared__ float4 sc[NQP * 64];
  __shared__ float4 ss[NQP * 64];
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int set = blockIdx.y;
  const int s0 = wave * 16;
  const bool active = s0 < S;  // idle waves still take part in the tile copies
  const int col = lane & 15;   // sample within the tile (B operand column / D column)
  const int krow = lane >> 4;  // 0..3: k within an MFMA step (A/B), row group of D
  const float* Xs = X + (int64_t)set * B * S_pad;
  float xb[NB];
#pragma unroll
  for (int kk = 0; kk < NB; ++kk) {
    const int b = 4 * kk + krow;
    xb[kk] = (active && b < nb_valid) ? Xs[(int64_t)b * S_pad + s0 + col] : 0.0f;
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
    v4f cc = {0.f, 0.f, 0.f, 0.f};
    v4f cs = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int pass = 0; pass < NP; ++pass) {
      __syncthreads();
      for (int g = threadIdx.x; g < NQP * 64; g += 256) {
        sc[g] = gc[pass * NQP * 64 + g];
        ss[g] = gs[pass * NQP * 64 + g];
      }
      __syncthreads();
      if (active) {
#pragma unroll
        for (int q = 0; q < NQP; ++q) {
          const int qq = pass * NQP + q;
          const float4 ac = sc[q * 64 + lane];
          const float4 as = ss[q * 64 + lane];
          cc = __builtin_amdgcn_mfma_f32_16x16x4f32(ac.x, xb[4 * qq + 0], cc, 0, 0, 0);
          cs = __builtin_amdgcn_mfma_f32_16x16x4f32(as.x, xb[4 * qq + 0], cs, 0, 0, 0);
          cc = __builtin_amdgcn_mfma_f32_16x16x4f32(ac.y, xb[4 * qq + 1], cc, 0, 0, 0);
          cs = __builtin_amdgcn_mfma_f32_16x16x4f32(as.y, xb[4 * qq + 1], cs, 0, 0, 0);
          cc = __builtin_amdgcn_mfma_f32_16x16x4f32(ac.z, xb[4 * qq + 2], cc, 0, 0, 0);
          cs = __builtin_amdgcn_mfma_f32_16x16x4f32(as.z, xb[4 * qq + 2], cs, 0, 0, 0);
          cc = __builtin_amdgcn_mfma_f32_16x16x4f32(ac.w, xb[4 * qq + 3], cc, 0, 0, 0);
          cs = __builtin_amdgcn_mfma_f32_16x16x4f32(as.w, xb[4 * qq + 3], cs, 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int f = f0 + 4 * krow + r;
      if (own && f < F) {
        const double pre = (double)cc[r];
        const double pim = (double)cs[r];
        const double c = blk[f * 4 + ph];
        const double sn = blk[f * 4 + ph + 1];
        const double dre = c * pre - sn * pim;
        const double dim = sn * pre + c * pim;
        const int64_t o = (int64_t)f * S + s;
        if (COMP) {
          two_sum_acc(&bv[r].x, &lv[r].x, dre);
          two_sum_acc(&bv[r].y, &lv[r].y, dim);
          lset[o] = lv[r];
        } else {
          bv[r].x += dre;
          bv[r].y += dim;
        }
        bset[o] = bv[r];
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

// Contract the block (first nb_valid rows of X) and accumulate into the
// float64 bins (nsets x F x S complex, interleaved re/im); lo (same shape or
// NULL; float32) holds the compensation terms.  Pc/Ps: the float32 phasor tables
// P[f][b] = cos/sin(w_f b dt) dt (f < F_pad, b < Bp = B rounded up to a
// multiple of 64, zero beyond B and F) stored in MFMA lane order, in float4
// groups: element [t][q][l][c] (t < F_pad/16, q < Bp/16, l < 64, c < 4) =
// P[16t + l%16][4(4q + c) + l/16] — a tile is a lane-contiguous block that
// a work-group copies straight into LDS.  Samples with index >= s_h_start carry the H time stamp.
// blk: F x 4 float64 scratch for the block phase, filled here from
// *step_ctr (n = last step of the block, or one past it for a partial one).
int mgh_dft_block(int nsets, const float* X, int S, int S_pad, int B, int nb_valid,
                  const float* Pc, const float* Ps, const double* omega, int F, int F_pad,
                  double dt, double* blk, int s_h_start, double* bins, float* lo,
                  const long long* step_ctr, long long n_start, void* stream) {
  MGH_REQUIRE(S >= 1 && S_pad >= S && S_pad % 16 == 0, "bad sample counts");
  MGH_REQUIRE(F >= 1 && F_pad >= F && F_pad % 16 == 0, "bad frequency counts F=%d F_pad=%d", F,
              F_pad);
  MGH_REQUIRE(B >= 1 && B <= DFT_BMAX && nb_valid >= 0 && nb_valid <= B, "bad block length");
  MGH_REQUIRE(X && Pc && Ps && omega && blk && bins && step_ctr, "null pointer");
  MGH_REQUIRE(s_h_start % 16 == 0, "s_h_start must be a multiple of 16");
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "bad nsets");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dft_phase, dim3((F + 255) / 256), dim3(256), 0, st, omega, F, dt, step_ctr,
                     n_start, B, blk);
  MGH_CHECK_LAUNCH("k_dft_phase");
  const int waves = (S + 15) / 16;
  dim3 grd((waves + 3) / 4, nsets, 1);
  const int bp = dft_bpad(B);
#define MGH_DFT_LAUNCH(NB)                                                                  \
  do {                                                                                      \
    if (lo)                                                                                 \
      hipLaunchKernelGGL((k_dft_block<NB, true>), grd, dim3(256), 0, st, X, S, S_pad, B,    \
                         nb_valid, Pc, Ps, F, (const double*)blk, s_h_start, bins, lo);     \
    else                                                                                    \
      hipLaunchKernelGGL((k_dft_block<NB, false>), grd, dim3(256), 0, st, X, S, S_pad, B,   \
                         nb_valid, Pc, Ps, F, (const double*)blk, s_h_start, bins, lo);     \
  } while (0)
  if (bp == 64) MGH_DFT_LAUNCH(16);
  else if (bp == 128) MGH_DFT_LAUNCH(32);
  else if (bp == 192) MGH_DFT_LAUNCH(48);
  else if (bp == 256) MGH_DFT_LAUNCH(64);
  else MGH_DFT_LAUNCH(128);
#undef MGH_DFT_LAUNCH
  MGH_CHECK_LAUNCH("k_dft_block");
  (void)F_pad;
  return MGH_OK;
}

}  // extern "C"
