// Macroscopic fields from scattered moments: mass and momentum binned onto a square grid, in FP64.
//
// Reference: bin_fields, direbm/fields.py:17-41 (numpy histogram2d of per-moment mass sum_k f_k
// and momentum sum_k f_k c_k, FP64).  Bin of a coordinate: the number of edges <= x, minus one,
// and the last edge itself belongs to the last bin (numpy.histogramdd's searchsorted rule).
// Grid layout [iy][ix] (the reference transposes histogram2d's [ix][iy]).
//
// Deterministic without atomics: moments are sorted by bin with the stable radix sort, and each
// bin sums its members in ascending moment order -- the order numpy.bincount adds the weights in.
// Per-moment mass is the sequential FP64 sum over k (numpy's sum over a length-7 row).
#include "engine.h"

namespace drb {

__constant__ static const double kCx64[7] = {0.0, 1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
__constant__ static const double kCy64[7] = {0.0, 0.0, 0.8660254037844386, 0.8660254037844386, 0.0,
                                             -0.8660254037844386, -0.8660254037844386};

// numpy.histogramdd bin of x over edges[0..nb]: -1 when outside (or NaN).
__device__ __forceinline__ int bin_of(double x, const double* __restrict__ e, int nb) {
  if (!(x >= e[0]) || !(x <= e[nb])) return -1;
  if (x == e[nb]) return nb - 1;
  int lo = 0, hi = nb + 1;  // count of edges <= x lies in [1, nb]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (e[mid - 1] <= x) lo = mid;
    else hi = mid;
  }
  return lo - 1;
}

__global__ __launch_bounds__(TB) void k_bin_keys(const float* __restrict__ pos, int m, const double* __restrict__ e,
                                                 int nb, uint32_t* __restrict__ keys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int ix = bin_of((double)pos[2 * (size_t)i], e, nb), iy = bin_of((double)pos[2 * (size_t)i + 1], e, nb);
  keys[i] = (ix < 0 || iy < 0) ? (uint32_t)nb * (uint32_t)nb : (uint32_t)iy * (uint32_t)nb + (uint32_t)ix;
}

// First index of the sorted keys >= v.
__device__ __forceinline__ int lower_bound(const uint32_t* __restrict__ k, int n, uint32_t v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (k[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(TB) void k_bin_sum(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ idx,
                                                int m, int nbins, const float* __restrict__ f,
                                                double* __restrict__ mass, double* __restrict__ mom) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  const int s0 = lower_bound(keys, m, (uint32_t)b), s1 = lower_bound(keys, m, (uint32_t)b + 1u);
  double am = 0.0, ax = 0.0, ay = 0.0;
  for (int s = s0; s < s1; ++s) {
    const float* row = f + 7 * (size_t)idx[s];
    double rm = 0.0, rx = 0.0, ry = 0.0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const double v = (double)row[k];
      rm += v;
      rx += v * kCx64[k];
      ry += v * kCy64[k];
    }
    am += rm;
    ax += rx;
    ay += ry;
  }
  mass[b] = am;
  mom[2 * (size_t)b] = ax;
  mom[2 * (size_t)b + 1] = ay;
}

// gf[i, k] = gmass[bin] + gmom[bin] . c_k (0 outside the grid).
__global__ __launch_bounds__(TB) void k_bin_vjp(const float* __restrict__ pos, int m, const double* __restrict__ e,
                                                int nb, const double* __restrict__ gmass,
                                                const double* __restrict__ gmom, float* __restrict__ gf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int ix = bin_of((double)pos[2 * (size_t)i], e, nb), iy = bin_of((double)pos[2 * (size_t)i + 1], e, nb);
  if (ix < 0 || iy < 0) {
#pragma unroll
    for (int k = 0; k < 7; ++k) gf[7 * (size_t)i + k] = 0.f;
    return;
  }
  const size_t b = (size_t)iy * nb + ix;
  const double gm = gmass ? gmass[b] : 0.0, gx = gmom ? gmom[2 * b] : 0.0, gy = gmom ? gmom[2 * b + 1] : 0.0;
#pragma unroll
  for (int k = 0; k < 7; ++k) gf[7 * (size_t)i + k] = (float)(gm + gx * kCx64[k] + gy * kCy64[k]);
}

}  // namespace drb

using namespace drb;

extern "C" int drb_bin_fields(drb_stage* st, const float* pos, const float* f, int64_t m, const double* edges,
                              int32_t nb, double* mass, double* mom, drb_stream_t stream) {
  if (!st || m < 0 || nb < 1 || !edges || !mass || !mom || (m > 0 && (!pos || !f)) || m > INT_MAX / 2 ||
      (int64_t)nb * nb >= (1LL << 31))
    return fail(DRB_ERR_INVALID_ARG, "drb_bin_fields: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const long long nbins = (long long)nb * nb;
  const size_t mz = (size_t)std::max<long long>(m, 1);
  DRB_CHECK(st->keys.ensure(mz * 4));
  DRB_CHECK(st->vals.ensure(mz * 4));
  if (m > 0) {
    hipLaunchKernelGGL(k_bin_keys, dim3(nblk(m)), dim3(TB), 0, s, pos, (int)m, edges, (int)nb, st->keys.as<uint32_t>());
    int bits = 1;
    while (bits < 32 && (1ULL << bits) <= (unsigned long long)nbins) ++bits;
    DRB_CHECK(st->sorter.sort(st->keys.as<uint32_t>(), st->vals.as<uint32_t>(), (int)m, bits, true, s));
  }
  hipLaunchKernelGGL(k_bin_sum, dim3(nblk(nbins)), dim3(TB), 0, s, st->keys.as<uint32_t>(), st->vals.as<uint32_t>(),
                     (int)m, (int)nbins, f, mass, mom);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

extern "C" int drb_bin_fields_vjp(const float* pos, int64_t m, const double* edges, int32_t nb, const double* gmass,
                                  const double* gmom, float* gf, drb_stream_t stream) {
  if (m < 0 || nb < 1 || !edges || (m > 0 && (!pos || !gf)) || m > INT_MAX / 2)
    return fail(DRB_ERR_INVALID_ARG, "drb_bin_fields_vjp: bad arguments");
  if (m == 0) return DRB_OK;
  hipLaunchKernelGGL(k_bin_vjp, dim3(nblk(m)), dim3(TB), 0, (hipStream_t)stream, pos, (int)m, edges, (int)nb, gmass,
                     gmom, gf);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}
