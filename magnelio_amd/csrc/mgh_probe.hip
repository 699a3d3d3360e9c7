// HBM streaming probe: the rate at which this part moves a curl launch's
// access mix when nothing else limits it.
//
// A curl launch streams, per cell, a fixed set of equally long arrays at the
// same offset: on the E side the six alpha/beta words, then per field set
// three E words read, three H words read and three E words written (the H
// side mirrors it).  The probe streams NR input arrays and NW output arrays
// of n floats each in exactly that way — one element per lane, every load of
// a lane issued before its arithmetic, out[w] = sum of in[r] for r = w mod NW
// — with no neighbour reads, no index arithmetic and no CPML, so its time is
// the floor for the launch's algorithmic bytes on the device it runs on.
// The mixes are those of the curl launches: 12 reads + 3 writes (one field
// set) and 42 + 18 (six sets, coefficients shared).
#include "mgh_common.h"

namespace {

template <int NR, int NW>
__global__ __launch_bounds__(256) void k_stream_probe(const float* __restrict__ in,
                                                      float* __restrict__ out, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  float x[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) x[r] = in[(int64_t)r * n + t];
  float acc[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w) acc[w] = 0.0f;
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r % NW] += x[r];
#pragma unroll
  for (int w = 0; w < NW; ++w) out[(int64_t)w * n + t] = acc[w];
}

}  // namespace

extern "C" {

// in: n_read arrays of n floats back to back; out: n_write arrays.
// Supported mixes: (12, 3) and (42, 18).
int mgh_stream_probe(int n_read, int n_write, const void* in, void* out, long long n,
                     void* stream) {
  MGH_REQUIRE(in && out, "null device pointer");
  MGH_REQUIRE(n >= 1 && (n + 255) / 256 < (1ll << 31), "n = %lld out of range", n);
  const bool k1 = n_read == 12 && n_write == 3, k6 = n_read == 42 && n_write == 18;
  MGH_REQUIRE(k1 || k6, "unsupported mix %d reads + %d writes (12 + 3 or 42 + 18)", n_read,
              n_write);
  const dim3 grd((unsigned)((n + 255) / 256));
  hipStream_t st = (hipStream_t)stream;
  if (k1)
    hipLaunchKernelGGL((k_stream_probe<12, 3>), grd, dim3(256), 0, st, (const float*)in,
                       (float*)out, (int64_t)n);
  else
    hipLaunchKernelGGL((k_stream_probe<42, 18>), grd, dim3(256), 0, st, (const float*)in,
                       (float*)out, (int64_t)n);
  MGH_CHECK_LAUNCH("k_stream_probe");
  return MGH_OK;
}

}  // extern "C"
