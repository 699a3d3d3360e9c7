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
                                                    long long x_buf,
                                                    const long long* __restrict__ ctr,
                                                    long long n_start) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int set = blockIdx.y;
  if (s >= S) return;
  const long long rel = *ctr - n_start;
  const int b = (int)(rel % B);
  X += ((rel / B) & 1) * x_buf;  // the period's buffer (x_buf 0: one buffer)
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
    int S_pad, int B, long long x_buf, const long long* __restrict__ ctr, long long n_start) {
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
  const long long rel = *ctr - n_start;
  const int b = (int)(rel % B);
  X += ((rel / B) & 1) * x_buf;  // the period's buffer (x_buf 0: one buffer)
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
// cos(w t0E), sin(w t0E), cos(w t0H), sin(w t0H) of block q.  The first step
// of block 0 is n0_fixed + z P for work-group z (P = nsub B: a table of
// consecutive periods) when n0_fixed >= 0, else the first step of the period
// holding *ctr.
__global__ void k_dft_phase(const double* __restrict__ omega, int F, double dt,
                            const long long* __restrict__ ctr, long long n_start, int B,
                            int nsub, long long n0_fixed, double* __restrict__ blk) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (f >= F) return;
  const long long P = (long long)B * nsub;
  long long base;
  if (n0_fixed >= 0) {
    base = n0_fixed + (long long)blockIdx.z * P;
  } else {
    const long long n = *ctr;
    base = n - ((n - n_start) % P);
  }
  const long long n0 = base + (long long)q * B;
  const double te = (double)n0 * dt;
  const double th = te + 0.5 * dt;
  double s, c;
  double* o = blk + (((int64_t)blockIdx.z * nsub + q) * F + f) * 4;
  sincos(omega[f] * te, &s, &c);
  o[0] = c;
  o[1] = s;
  sincos(omega[f] * th, &s, &c);
  o[2] = c;
  o[3] = s;
}

// Which samples and frequencies a contraction launch covers.  Samples go in
// waves of 16; work-group x of the launch takes waves wave_lo + 4x .. +3 and
// work-group z the frequency tiles (16 frequencies each) z*tiles_per ..
// (z+1)*tiles_per - 1.  With ctr set the launch is one slice of a
// pipelined contraction: the period p holding *ctr (P = nsub*B steps from
// n_start) is being sampled into buffer p mod 2 while row r = (*ctr -
// n_start) mod P of it is written; the launch, made at the steps with
// (r + 1) mod cad = 0, contracts slice j = r div cad of the P / cad slices
// of the finished period p - 1 — waves [j W / (P/cad), (j + 1) W / (P/cad))
// (W = all waves), all P rows, buffer (p - 1) mod 2 — with that period's
// phases from the table blk (one nsub x F x 4 entry per period from
// n_start).  Period 0 has nothing to contract.  The P steps of period p
// contract every wave of period p - 1 once, before its buffer is
// overwritten by period p + 1.
struct DftRange {
  const long long* ctr;  // pipelined slice: the step counter; NULL: fixed range
  long long n_start;
  long long x_buf;       // floats per sample buffer (pipelined)
  int wave_lo, wave_hi;  // fixed range of sample waves
  int tiles_per;         // frequency tiles per work-group z
  int cad;               // pipelined: steps between slices (divides P)
  int n_periods;         // pipelined: periods in the phase table
};

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
    float* __restrict__ lo, DftRange rg) {
  constexpr int NQ = NB / 4;                // float4 groups per lane
  constexpr int NQP = NQ < 16 ? NQ : 16;    // groups per LDS pass (256 steps)
  constexpr int NP = NQ / NQP;              // passes
  static_assert(NQ % NQP == 0, "block tile must be whole passes");
  static_assert(NSUB == 1 || NP == 1, "several blocks per launch need a one-pass tile");
  __shared__ float4 sc[NQP * 64];
  __shared__ float4 ss[NQP * 64];
  int wlo = rg.wave_lo, whi = rg.wave_hi;
  if (rg.ctr) {  // pipelined slice (uniform over the launch)
    const long long P = (long long)B * NSUB;
    const long long rel = *rg.ctr - rg.n_start;
    const long long p = rel / P;
    // no finished period yet; a period past the phase table (never for a
    // run within its horizon) is left alone rather than read out of bounds
    if (p <= 0 || p > rg.n_periods) return;
    const long long ns = P / rg.cad;
    const long long j = (rel % P) / rg.cad;
    const long long W = (S + 15) / 16;
    wlo = (int)(j * W / ns);
    whi = (int)((j + 1) * W / ns);
    X += ((p - 1) & 1) * rg.x_buf;
    blk += (p - 1) * NSUB * F * 4;
    nb_valid = (int)P;
  }
  // a work-group wholly past the range leaves before any barrier
  if (wlo + (int)blockIdx.x * 4 >= whi) return;
  const int lane = threadIdx.x & 63;
  const int wave = wlo + (int)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int set = blockIdx.y;
  const int s0 = wave * 16;
  const bool active = wave < whi && s0 < S;  // idle waves still take part in the tile copies
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
  const int f_begin = (int)blockIdx.z * rg.tiles_per * 16;
  const int f_end = min(F, f_begin + rg.tiles_per * 16);
  for (int f0 = f_begin; f0 < f_end; f0 += 16) {
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

// Launch the contraction over a range (rg) with grid (gx, nsets, gz); the
// phases are in blk.  Variant by padded block length and blocks per launch.
static int launch_blocks(dim3 grd, hipStream_t st, const float* X, int S, int S_pad, int B,
                         int nsub, int nb_valid, const float* Pc, const float* Ps, int F,
                         const double* blk, int s_h_start, double* bins, float* lo,
                         const DftRange& rg) {
  const int bp = dft_bpad(B);
#define MGH_DFT_LAUNCH(NB, NSUB)                                                              \
  do {                                                                                        \
    if (lo)                                                                                   \
      hipLaunchKernelGGL((k_dft_block<NB, true, NSUB>), grd, dim3(256), 0, st, X, S, S_pad,  \
                         B, nb_valid, Pc, Ps, F, blk, s_h_start, bins, lo, rg);               \
    else                                                                                      \
      hipLaunchKernelGGL((k_dft_block<NB, false, NSUB>), grd, dim3(256), 0, st, X, S, S_pad, \
                         B, nb_valid, Pc, Ps, F, blk, s_h_start, bins, lo, rg);               \
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

// Checks shared by the contraction entry points.
#define MGH_DFT_CHECK_CONTRACTION()                                                           \
  do {                                                                                        \
    MGH_REQUIRE(S >= 1 && S_pad >= S && S_pad % 16 == 0, "bad sample counts");               \
    MGH_REQUIRE(F >= 1 && F_pad >= F && F_pad % 16 == 0, "bad frequency counts F=%d F_pad=%d", \
                F, F_pad);                                                                    \
    MGH_REQUIRE(B >= 1 && B <= DFT_BMAX, "bad block length %d", B);                           \
    MGH_REQUIRE(nsub == 1 || nsub == 2 || nsub == 4,                                          \
                "blocks per launch must be 1, 2 or 4, got %d", nsub);                         \
    MGH_REQUIRE(nsub == 1 || dft_bpad(B) * nsub <= 256,                                       \
                "%d blocks of %d steps exceed 256 sample rows per lane", nsub, B);            \
    MGH_REQUIRE(X && Pc && Ps && bins, "null pointer");                                        \
    MGH_REQUIRE(s_h_start % 16 == 0, "s_h_start must be a multiple of 16");                   \
    MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "bad nsets");                                   \
  } while (0)

}  // namespace

extern "C" {

// Sample one step into row ((n - n_start) mod B) of the block buffer
// X (nsets x B x S_pad float32).  idx/w: S x 8 terms; idx >= 0 indexes the
// E vector, idx < 0 the H vector at -(idx + 1) (device layout); weight 0
// marks an unused term.
int mgh_dft_sample(int dtype, int nsets, const void* e, long long e_set_stride, const void* h,
                   long long h_set_stride, const int* idx, const float* w, int S, float* X,
                   int S_pad, int B, long long x_buf_stride, const long long* step_ctr,
                   long long n_start, void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8");
  MGH_REQUIRE(S >= 1 && S_pad >= S && S_pad % 16 == 0, "bad sample counts S=%d S_pad=%d", S, S_pad);
  MGH_REQUIRE(B >= 1 && B <= DFT_BMAX, "block length must be 1..512, got %d", B);
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "bad nsets");
  MGH_REQUIRE(e && h && idx && w && X && step_ctr, "null pointer");
  MGH_REQUIRE(x_buf_stride == 0 || x_buf_stride >= (long long)nsets * B * S_pad,
              "sample buffers overlap (stride %lld)", x_buf_stride);
  dim3 grd((S + 255) / 256, nsets, 1);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    hipLaunchKernelGGL(k_dft_sample<float>, grd, dim3(256), 0, st, (const float*)e, e_set_stride,
                       (const float*)h, h_set_stride, idx, w, S, X, S_pad, B, x_buf_stride, step_ctr,
                       n_start);
  else
    hipLaunchKernelGGL(k_dft_sample<double>, grd, dim3(256), 0, st, (const double*)e, e_set_stride,
                       (const double*)h, h_set_stride, idx, w, S, X, S_pad, B, x_buf_stride, step_ctr,
                       n_start);
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
                          long long x_buf_stride, const long long* step_ctr, long long n_start,
                          void* stream) {
  MGH_REQUIRE(dtype == 4 || dtype == 8, "dtype must be 4 or 8");
  MGH_REQUIRE(S >= 1 && S_pad >= S && S_pad % 16 == 0, "bad sample counts S=%d S_pad=%d", S, S_pad);
  MGH_REQUIRE(B >= 1 && B <= DFT_BMAX, "block length must be 1..512, got %d", B);
  MGH_REQUIRE(nsets >= 1 && nsets <= 65535, "bad nsets");
  MGH_REQUIRE(G >= 1 && G <= 4096, "bad group count %d", G);
  MGH_REQUIRE(n_blocks >= 1, "bad work-group count %d", n_blocks);
  MGH_REQUIRE(e && h && grp && wtab && X && step_ctr, "null pointer");
  MGH_REQUIRE(x_buf_stride == 0 || x_buf_stride >= (long long)nsets * B * S_pad,
              "sample buffers overlap (stride %lld)", x_buf_stride);
  dim3 grd((unsigned)n_blocks, nsets, 1);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 4)
    hipLaunchKernelGGL(k_dft_sample_groups<float>, grd, dim3(256), 0, st, (const float*)e,
                       e_set_stride, (const float*)h, h_set_stride, grp, G, wtab, X, S_pad, B,
                       x_buf_stride, step_ctr, n_start);
  else
    hipLaunchKernelGGL(k_dft_sample_groups<double>, grd, dim3(256), 0, st, (const double*)e,
                       e_set_stride, (const double*)h, h_set_stride, grp, G, wtab, X, S_pad, B,
                       x_buf_stride, step_ctr, n_start);
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
  MGH_DFT_CHECK_CONTRACTION();
  MGH_REQUIRE(step_ctr && omega && blk, "null pointer");
  MGH_REQUIRE(nb_valid >= 0 && nb_valid <= B * nsub, "bad valid row count %d", nb_valid);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dft_phase, dim3((F + 255) / 256, nsub), dim3(256), 0, st, omega, F, dt,
                     step_ctr, n_start, B, nsub, -1LL, blk);
  MGH_CHECK_LAUNCH("k_dft_phase");
  const int waves = (S + 15) / 16;
  DftRange rg{nullptr, 0, 0, 0, waves, (F + 15) / 16, 1, 0};
  return launch_blocks(dim3((waves + 3) / 4, nsets, 1), st, X, S, S_pad, B, nsub, nb_valid, Pc,
                       Ps, F, blk, s_h_start, bins, lo, rg);
}

// mgh_dft_blocks over the sample waves [wave_lo, wave_hi) (16 samples each)
// with the first step of block 0 given (n0), frequencies split over fsplit
// groups of work-groups.
int mgh_dft_blocks_range(int nsets, const float* X, int S, int S_pad, int B, int nsub,
                         int nb_valid, int wave_lo, int wave_hi, long long n0, const float* Pc,
                         const float* Ps, const double* omega, int F, int F_pad, double dt,
                         double* blk, int s_h_start, double* bins, float* lo, int fsplit,
                         void* stream) {
  MGH_DFT_CHECK_CONTRACTION();
  MGH_REQUIRE(omega && blk, "null pointer");
  MGH_REQUIRE(nb_valid >= 0 && nb_valid <= B * nsub, "bad valid row count %d", nb_valid);
  const int waves = (S + 15) / 16;
  MGH_REQUIRE(wave_lo >= 0 && wave_lo <= wave_hi && wave_hi <= waves,
              "bad wave range [%d, %d) of %d", wave_lo, wave_hi, waves);
  MGH_REQUIRE(n0 >= 0, "bad first step %lld", n0);
  MGH_REQUIRE(fsplit >= 1, "bad frequency split %d", fsplit);
  if (wave_hi == wave_lo || nb_valid == 0) return MGH_OK;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dft_phase, dim3((F + 255) / 256, nsub), dim3(256), 0, st, omega, F, dt,
                     (const long long*)nullptr, 0LL, B, nsub, n0, blk);
  MGH_CHECK_LAUNCH("k_dft_phase");
  const int tiles = (F + 15) / 16;
  const int per = (tiles + fsplit - 1) / fsplit;
  DftRange rg{nullptr, 0, 0, wave_lo, wave_hi, per, 1, 0};
  return launch_blocks(dim3((wave_hi - wave_lo + 3) / 4, nsets, (tiles + per - 1) / per), st, X,
                       S, S_pad, B, nsub, nb_valid, Pc, Ps, F, blk, s_h_start, bins, lo, rg);
}

// Phase table of n_periods consecutive periods (nsub B steps each) from
// step n0: blk[((z nsub + q) F + f) 4 + 0..3] for period z.
int mgh_dft_phase_table(const double* omega, int F, double dt, long long n0, int B, int nsub,
                        int n_periods, double* blk, void* stream) {
  MGH_REQUIRE(omega && blk, "null pointer");
  MGH_REQUIRE(F >= 1 && B >= 1 && B <= DFT_BMAX && nsub >= 1 && nsub <= 4, "bad table shape");
  MGH_REQUIRE(n0 >= 0 && n_periods >= 1 && n_periods <= 65535, "bad periods n0=%lld n=%d", n0,
              n_periods);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dft_phase, dim3((F + 255) / 256, nsub, n_periods), dim3(256), 0, st,
                     omega, F, dt, (const long long*)nullptr, 0LL, B, nsub, n0, blk);
  MGH_CHECK_LAUNCH("k_dft_phase");
  return MGH_OK;
}

// One pipelined slice (DftRange): X holds two sample buffers x_buf_stride
// floats apart, each nsets x nsub*B x S_pad; the period holding *step_ctr is
// sampled into one while this launch contracts slice ((n - n_start) mod P)
// div cad of the P / cad slices of the other — the previous period — with
// its phases from the table (mgh_dft_phase_table from n_start, n_periods
// periods).  Made at the steps with ((n - n_start) mod P + 1) mod cad = 0.
int mgh_dft_slice(int nsets, const float* X, long long x_buf_stride, int S, int S_pad, int B,
                  int nsub, const float* Pc, const float* Ps, int F, int F_pad,
                  const double* phase_table, int n_periods, int s_h_start, double* bins, float* lo,
                  const long long* step_ctr, long long n_start, int cad, int fsplit,
                  void* stream) {
  MGH_DFT_CHECK_CONTRACTION();
  MGH_REQUIRE(step_ctr && phase_table, "null pointer");
  MGH_REQUIRE(x_buf_stride >= (long long)nsets * B * nsub * S_pad,
              "sample buffers overlap (stride %lld)", x_buf_stride);
  MGH_REQUIRE(fsplit >= 1, "bad frequency split %d", fsplit);
  MGH_REQUIRE(n_periods >= 1, "bad phase table length %d", n_periods);
  const long long P = (long long)B * nsub;
  MGH_REQUIRE(cad >= 1 && P % cad == 0, "slice cadence %d does not divide the period %lld", cad,
              P);
  hipStream_t st = (hipStream_t)stream;
  const long long waves = (S + 15) / 16;
  const long long ns = P / cad;
  const int slice = (int)((waves + ns - 1) / ns);  // the largest slice
  if (slice == 0) return MGH_OK;
  const int tiles = (F + 15) / 16;
  const int per = (tiles + fsplit - 1) / fsplit;
  DftRange rg{step_ctr, n_start, x_buf_stride, 0, 0, per, cad, n_periods};
  return launch_blocks(dim3((slice + 3) / 4, nsets, (tiles + per - 1) / per), st, X, S, S_pad, B,
                       nsub, (int)P, Pc, Ps, F, phase_table, s_h_start, bins, lo, rg);
}

int mgh_dft_block(int nsets, const float* X, int S, int S_pad, int B, int nb_valid,
                  const float* Pc, const float* Ps, const double* omega, int F, int F_pad,
                  double dt, double* blk, int s_h_start, double* bins, float* lo,
                  const long long* step_ctr, long long n_start, void* stream) {
  return mgh_dft_blocks(nsets, X, S, S_pad, B, 1, nb_valid, Pc, Ps, omega, F, F_pad, dt, blk,
                        s_h_start, bins, lo, step_ctr, n_start, stream);
}

}  // extern "C"
