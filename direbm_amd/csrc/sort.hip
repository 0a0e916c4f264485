// Device-wide primitives of the engine: stable LSD radix sort of (key, value) pairs, an
// exclusive scan of int2 counts and a deterministic FP64 reduction.
//
// These replace warp.utils.radix_sort_pairs / runlength_encode / array_scan (warp-lang 1.14.0,
// called at direbm/warp/propagation.py:138,143,149).  The sort is the classic three-kernel
// reduce-then-scan LSD form, redesigned for wave64: per-block digit histograms in LDS, one
// block per digit for the column scan, and a downsweep that ranks keys stably with wave64
// ballots (8 ballots per key give the mask of lanes holding the same digit), reorders a
// 4096-key tile in LDS and writes each digit run contiguously.  Stability (the index tie-break
// of the reference) follows from ranking keys in input order: earlier waves, earlier rounds and
// lower lanes rank first.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace drb {

// ----------------------------------------------------------------------------- errors
static thread_local Error g_err = {DRB_OK, ""};

void set_error(int code, const std::string& msg) {
  g_err.code = code;
  g_err.msg = msg;
}
int fail(int code, const std::string& msg) {
  set_error(code, msg);
  return code;
}

float sqrt_threshold(float r) {
  // Host sqrtf is IEEE correctly rounded; walk from fp32(r * r) to the last s with sqrtf(s) <= r.
  float s = r * r;
  while (std::sqrt(std::nextafter(s, INFINITY)) <= r) s = std::nextafter(s, INFINITY);
  while (s > 0.f && std::sqrt(s) > r) s = std::nextafter(s, 0.f);
  return s;
}

int Buf::ensure(size_t want) {
  if (view) return fail(DRB_ERR_STATE, "internal: a buffer view cannot grow");
  if (want <= bytes) return DRB_OK;
  // Geometric growth: 1.5x the first request, at least twice the old size after that.  A regrowth
  // frees (which synchronises the device) and maps new memory, a few ms at the sizes of a 1M-point
  // step, so a cloud that keeps growing should meet it only a few times.
  size_t b = std::max<size_t>(want + want / 2, 256);
  if (bytes) b = std::max(b, 2 * bytes);
  if (p) {
    DRB_HIP(hipFree(p));
    p = nullptr;
    bytes = 0;
  }
  DRB_HIP(hipMalloc(&p, b));
  bytes = b;
  return DRB_OK;
}
void Buf::release() {
  if (p && !view) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  view = false;
}

// ------------------------------------------------------------------------ radix sort
constexpr int kSortThreads = 256;
constexpr int kWaves = kSortThreads / 64;
#ifndef DRB_SORT_ITEMS
#define DRB_SORT_ITEMS 16
#endif
constexpr int kItems = DRB_SORT_ITEMS;
constexpr int kTile = kSortThreads * kItems;  // 4096 keys
constexpr int kRadix = 256;
#ifndef DRB_SORT_BLOCKS
#define DRB_SORT_BLOCKS 1024
#endif
// two 4096-key tiles per block at 7M keys, every block resident: sort 0.154-0.157 ms per step at
// 1M points (512 blocks: 0.173-0.176; 2048: 0.161-0.163)
constexpr int kMaxBlocks = DRB_SORT_BLOCKS;

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned popc_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Lanes of the wave holding the same 8-bit digit d (valid lanes only): eight ballots.
__device__ __forceinline__ unsigned long long digit_peers(bool valid, uint32_t d) {
  unsigned long long peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    unsigned long long bb = __ballot((d >> b) & 1u);
    peers &= ((d >> b) & 1u) ? bb : ~bb;
  }
  return peers;
}

// Tiles per block when the live count is known only on the device (graph mode launches the
// capacity's grid): the live tiles spread over every block of the grid, not packed into the first
// ones.  The up-sweep and the down-sweep of a pass derive the same split from the same count.
__device__ __forceinline__ int live_tiles_per_block(int n) {
  const int nt = (int)(((long long)n + kTile - 1) / kTile);
  return max(1, (nt + (int)gridDim.x - 1) / (int)gridDim.x);
}

// Per-block digit histogram, written digit-major: hist[d * nblocks + b].  Each thread loads its
// kItems keys of a tile before counting any (the tile's loads are all in flight at once), with the
// tile layout of the down-sweep, and counts them with LDS atomics into its wave's row.  (Counting a
// round of 64 keys with ballots instead, as the down-sweep ranks, measured slower.)
__global__ __launch_bounds__(kSortThreads) void k_sort_upsweep(const uint32_t* __restrict__ keys, int n_arg,
                                                               const int* __restrict__ ndev, int shift,
                                                               int tiles_per_block, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kWaves][kRadix];
  const int n = ndev ? min(*ndev, n_arg) : n_arg;
  if (tiles_per_block == 0) tiles_per_block = live_tiles_per_block(n);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int i = t; i < kWaves * kRadix; i += kSortThreads) (&h[0][0])[i] = 0;
  __syncthreads();
  const long long tile0 = (long long)blockIdx.x * tiles_per_block;
  for (int tt = 0; tt < tiles_per_block; ++tt) {
    const long long tbase = (tile0 + tt) * kTile;
    if (tbase >= n) break;
    const int tn = (int)min((long long)kTile, (long long)n - tbase);
    uint32_t key[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const int li = w * (64 * kItems) + j * 64 + lane;
      key[j] = li < tn ? keys[tbase + li] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const int li = w * (64 * kItems) + j * 64 + lane;
      const bool valid = li < tn;
      const uint32_t d = (key[j] >> shift) & (kRadix - 1);
      if (valid) atomicAdd(&h[w][d], 1u);
    }
  }
  __syncthreads();
  for (int d = t; d < kRadix; d += kSortThreads) {
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) s += h[q][d];
    hist[(size_t)d * gridDim.x + blockIdx.x] = s;
  }
}

// Block-wide exclusive scan of one value per thread (256 threads); returns the block total.
__device__ __forceinline__ uint32_t block_exscan(uint32_t v, uint32_t* sh /* >= kWaves+1 */, uint32_t& total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
  for (int q = 0; q < kWaves; ++q) {
    uint32_t s = sh[q];
    if (q < w) wbase += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return wbase + x - v;
}

// Column scan: block d turns hist[d * nb + b] into the exclusive prefix over b, and writes the
// column total to tot[d].
__global__ __launch_bounds__(kSortThreads) void k_sort_colscan(uint32_t* __restrict__ hist, int nb,
                                                               uint32_t* __restrict__ tot) {
  __shared__ uint32_t sh[kWaves + 1];
  uint32_t* row = hist + (size_t)blockIdx.x * nb;
  uint32_t carry = 0;
  for (int base = 0; base < nb; base += kSortThreads) {
    int i = base + threadIdx.x;
    uint32_t v = i < nb ? row[i] : 0u;
    uint32_t total;
    uint32_t ex = block_exscan(v, sh, total);
    if (i < nb) row[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// Stable scatter of one pass.  Block b owns tiles [b * tpb, (b+1) * tpb).
__global__ __launch_bounds__(kSortThreads) void k_sort_downsweep(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int n_arg, const int* __restrict__ ndev, int shift, int tiles_per_block,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ tot, int vals_iota) {
  __shared__ uint32_t run_off[kRadix];
  __shared__ uint32_t wcnt[kWaves][kRadix];
  __shared__ uint32_t lstart[kRadix];
  __shared__ uint32_t tcnt[kRadix];
  __shared__ uint32_t lkeys[kTile];
  __shared__ uint32_t lvals[kTile];
  __shared__ uint32_t sh[kWaves + 1];

  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int nb = gridDim.x;
  const int n = ndev ? min(*ndev, n_arg) : n_arg;
  if (tiles_per_block == 0) tiles_per_block = live_tiles_per_block(n);
  // run_off[d] = global start of digit d for this block = digit base + column prefix.
  {
    uint32_t total;
    uint32_t dbase = block_exscan(tot[t], sh, total);  // kSortThreads == kRadix
    run_off[t] = dbase + hist[(size_t)t * nb + blockIdx.x];
  }
  __syncthreads();

  const long long tile0 = (long long)blockIdx.x * tiles_per_block;
  for (int tt = 0; tt < tiles_per_block; ++tt) {
    const long long tbase = (tile0 + tt) * kTile;
    if (tbase >= n) break;
    const int tn = (int)min((long long)kTile, (long long)n - tbase);
    for (int i = t; i < kWaves * kRadix; i += kSortThreads) (&wcnt[0][0])[i] = 0;
    __syncthreads();

    uint32_t key[kItems], val[kItems], rank[kItems];
    // Every key and value of the tile is loaded before any is ranked, from an index clamped into the
    // tile (no branch around a load): the loads are in flight together.  A load inside each ranking
    // round waited its whole latency sixteen times per tile.
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const int li = min(w * (64 * kItems) + j * 64 + lane, tn - 1);
      key[j] = kin[tbase + li];
    }
    if (vals_iota) {
#pragma unroll
      for (int j = 0; j < kItems; ++j) val[j] = (uint32_t)(tbase + w * (64 * kItems) + j * 64 + lane);
    } else {
#pragma unroll
      for (int j = 0; j < kItems; ++j) val[j] = vin[tbase + min(w * (64 * kItems) + j * 64 + lane, tn - 1)];
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const int li = w * (64 * kItems) + j * 64 + lane;  // index within the tile
      const bool valid = li < tn;
      const uint32_t k = valid ? key[j] : 0u;
      const uint32_t v = valid ? val[j] : 0u;
      uint32_t d = (k >> shift) & (kRadix - 1);
      const unsigned long long peers = digit_peers(valid, d);
      uint32_t below = popc_below(peers);
      uint32_t base = valid ? wcnt[w][d] : 0u;
      __builtin_amdgcn_wave_barrier();
      if (valid && below == 0) wcnt[w][d] = base + (uint32_t)__popcll(peers);
      __builtin_amdgcn_wave_barrier();
      key[j] = k;
      val[j] = v;
      rank[j] = valid ? base + below : 0xFFFFFFFFu;
    }
    __syncthreads();
    // Per digit: exclusive prefix over waves, and the tile count.
    {
      const int d = t;
      uint32_t s = 0;
#pragma unroll
      for (int q = 0; q < kWaves; ++q) {
        uint32_t c = wcnt[q][d];
        wcnt[q][d] = s;
        s += c;
      }
      tcnt[d] = s;
      uint32_t total;
      lstart[d] = block_exscan(s, sh, total);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      if (rank[j] != 0xFFFFFFFFu) {
        uint32_t d = (key[j] >> shift) & (kRadix - 1);
        uint32_t lp = lstart[d] + wcnt[w][d] + rank[j];
        lkeys[lp] = key[j];
        lvals[lp] = val[j];
      }
    }
    __syncthreads();
    for (int i = t; i < tn; i += kSortThreads) {
      uint32_t k = lkeys[i];
      uint32_t d = (k >> shift) & (kRadix - 1);
      uint32_t g = run_off[d] + (uint32_t)i - lstart[d];
      kout[g] = k;
      vout[g] = lvals[i];
    }
    __syncthreads();
    run_off[t] += tcnt[t];
    __syncthreads();
  }
}

int RadixSort::reserve(int n) {
  const long long ntiles = ((long long)std::max(n, 1) + kTile - 1) / kTile;
  const int tpb = (int)((ntiles + kMaxBlocks - 1) / kMaxBlocks);
  const int nb = (int)((ntiles + tpb - 1) / tpb);
  DRB_CHECK(k1.ensure((size_t)std::max(n, 1) * 4));
  DRB_CHECK(v1.ensure((size_t)std::max(n, 1) * 4));
  DRB_CHECK(hist.ensure((size_t)kRadix * nb * 4));
  DRB_CHECK(tot.ensure((size_t)kRadix * 4));
  return DRB_OK;
}

size_t radix_hist_words(int n) {
  const long long ntiles = ((long long)std::max(n, 1) + kTile - 1) / kTile;
  const int tpb = (int)((ntiles + kMaxBlocks - 1) / kMaxBlocks);
  const int nb = (int)((ntiles + tpb - 1) / tpb);
  return (size_t)kRadix * nb;
}

int radix_sort_into(uint32_t* keys, uint32_t* vals, int n, int bits, bool vals_are_iota, uint32_t* k1, uint32_t* v1,
                    uint32_t* hist, uint32_t* tot, hipStream_t s, const int* ndev, uint32_t** kres,
                    uint32_t** vres) {
  if (n <= 0) return DRB_OK;
  if (bits < 1) bits = 1;
  const int passes = (bits + 7) / 8;
  const long long ntiles = ((long long)n + kTile - 1) / kTile;
  const int tpb_cap = (int)((ntiles + kMaxBlocks - 1) / kMaxBlocks);
  const int nb = (int)((ntiles + tpb_cap - 1) / tpb_cap);
  // with the count on the device, the kernels split the live tiles over the grid themselves
  const int tpb = ndev ? 0 : tpb_cap;
  uint32_t* ka = keys;
  uint32_t* va = vals;
  uint32_t* kb = k1;
  uint32_t* vb = v1;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(k_sort_upsweep, dim3(nb), dim3(kSortThreads), 0, s, ka, n, ndev, shift, tpb, hist);
    hipLaunchKernelGGL(k_sort_colscan, dim3(kRadix), dim3(kSortThreads), 0, s, hist, nb, tot);
    hipLaunchKernelGGL(k_sort_downsweep, dim3(nb), dim3(kSortThreads), 0, s, ka, va, kb, vb, n, ndev, shift, tpb,
                       hist, tot, (vals_are_iota && p == 0) ? 1 : 0);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (kres) {
    // the caller reads the result where the last pass wrote it
    *kres = ka;
    *vres = va;
  } else if (ka != keys) {
    DRB_HIP(hipMemcpyAsync(keys, ka, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    DRB_HIP(hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  }
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

int RadixSort::sort_swap(Buf& keys, Buf& vals, int n, int bits, bool vals_are_iota, hipStream_t s) {
  if (n <= 0) return DRB_OK;
  DRB_CHECK(reserve(n));
  if (bits < 1) bits = 1;
  const int passes = (bits + 7) / 8;
  if (passes % 2 == 0 || keys.view || vals.view)
    return radix_sort_into(keys.as<uint32_t>(), vals.as<uint32_t>(), n, bits, vals_are_iota, k1.as<uint32_t>(),
                           v1.as<uint32_t>(), hist.as<uint32_t>(), tot.as<uint32_t>(), s);
  // odd: sort from the alternate buffers into (keys, vals) after exchanging them, so the last pass
  // writes the caller's buffers.  The first pass reads what the caller filled (now k1 / v1).
  std::swap(keys, k1);
  std::swap(vals, v1);
  uint32_t* ka = k1.as<uint32_t>();  // input (the caller's filled buffers)
  uint32_t* va = v1.as<uint32_t>();
  uint32_t* kb = keys.as<uint32_t>();
  uint32_t* vb = vals.as<uint32_t>();
  // both pairs must hold n items: the exchanged-in scratch was reserved for n above
  const long long ntiles = ((long long)n + kTile - 1) / kTile;
  const int tpb = (int)((ntiles + kMaxBlocks - 1) / kMaxBlocks);
  const int nb = (int)((ntiles + tpb - 1) / tpb);
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(k_sort_upsweep, dim3(nb), dim3(kSortThreads), 0, s, ka, n, (const int*)nullptr, shift, tpb,
                       hist.as<uint32_t>());
    hipLaunchKernelGGL(k_sort_colscan, dim3(kRadix), dim3(kSortThreads), 0, s, hist.as<uint32_t>(), nb,
                       tot.as<uint32_t>());
    hipLaunchKernelGGL(k_sort_downsweep, dim3(nb), dim3(kSortThreads), 0, s, ka, va, kb, vb, n, (const int*)nullptr,
                       shift, tpb, hist.as<uint32_t>(), tot.as<uint32_t>(), (vals_are_iota && p == 0) ? 1 : 0);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

int RadixSort::sort_graph(uint32_t* keys, uint32_t* vals, int n, int bits, hipStream_t s, const int* ndev,
                          uint32_t** kres, uint32_t** vres) {
  *kres = keys;
  *vres = vals;
  if (n <= 0) return DRB_OK;
  DRB_CHECK(reserve(n));
  return radix_sort_into(keys, vals, n, bits, true, k1.as<uint32_t>(), v1.as<uint32_t>(), hist.as<uint32_t>(),
                         tot.as<uint32_t>(), s, ndev, kres, vres);
}

int RadixSort::sort(uint32_t* keys, uint32_t* vals, int n, int bits, bool vals_are_iota, hipStream_t s,
                    float* ms_out, const int* ndev) {
  if (n <= 0) return DRB_OK;
  DRB_CHECK(reserve(n));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms_out) {
    DRB_HIP(hipEventCreate(&e0));
    DRB_HIP(hipEventCreate(&e1));
    DRB_HIP(hipEventRecord(e0, s));
  }
  DRB_CHECK(radix_sort_into(keys, vals, n, bits, vals_are_iota, k1.as<uint32_t>(), v1.as<uint32_t>(),
                            hist.as<uint32_t>(), tot.as<uint32_t>(), s, ndev));
  if (ms_out) {
    DRB_HIP(hipEventRecord(e1, s));
    DRB_HIP(hipEventSynchronize(e1));
    DRB_HIP(hipEventElapsedTime(ms_out, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return DRB_OK;
}

// ------------------------------------------------------------------------------ scan
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kScanThreads * kScanItems;

__device__ __forceinline__ int2 add2(int2 a, int2 b) { return make_int2(a.x + b.x, a.y + b.y); }

// Block-wide exclusive scan of int2 (256 threads).
__device__ __forceinline__ int2 block_exscan2(int2 v, int2* sh, int2& total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int2 x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int yx = __shfl_up(x.x, o, 64), yy = __shfl_up(x.y, o, 64);
    if (lane >= o) { x.x += yx; x.y += yy; }
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  int2 wbase = make_int2(0, 0), tot = make_int2(0, 0);
  for (int q = 0; q < kScanThreads / 64; ++q) {
    int2 s = sh[q];
    if (q < w) wbase = add2(wbase, s);
    tot = add2(tot, s);
  }
  __syncthreads();
  total = tot;
  return make_int2(wbase.x + x.x - v.x, wbase.y + x.y - v.y);
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int2* __restrict__ in, long long n,
                                                              int2* __restrict__ part) {
  __shared__ int2 sh[kScanThreads / 64 + 1];
  long long beg = (long long)blockIdx.x * kScanChunk;
  // all loads in flight before any is used (index clamped into the array, value masked)
  int2 v[kScanItems];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) v[j] = in[min(beg + (long long)j * kScanThreads + threadIdx.x, n - 1)];
  int2 s = make_int2(0, 0);
#pragma unroll
  for (int j = 0; j < kScanItems; ++j)
    if (beg + (long long)j * kScanThreads + threadIdx.x < n) s = add2(s, v[j]);
  int2 total;
  (void)block_exscan2(s, sh, total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// Large scans: a single block turns the partials into their exclusive prefix in place.
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(int2* __restrict__ part, int nb) {
  __shared__ int2 sh[kScanThreads / 64 + 1];
  int2 carry = make_int2(0, 0);
  for (int base = 0; base < nb; base += kScanThreads) {
    int i = base + threadIdx.x;
    int2 v = i < nb ? part[i] : make_int2(0, 0);
    int2 total;
    int2 ex = block_exscan2(v, sh, total);
    if (i < nb) part[i] = add2(carry, ex);
    carry = add2(carry, total);
  }
}

// Scans up to kScanFold chunks skip k_scan_partials: each block adds the partials of the blocks
// before it itself (one wave, integer sums).
constexpr long long kScanFold = 1024;

// Down-sweep: the block's offset (prefixed partials, or the sum of the raw partials before it),
// then the chunk's scan; the last block writes the grand total out[n].
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const int2* __restrict__ in, long long n,
                                                            const int2* __restrict__ part, int nb, int prefixed,
                                                            int2* __restrict__ out) {
  __shared__ int2 sh[kScanThreads / 64 + 1];
  __shared__ int2 base;
  if (threadIdx.x < 64) {
    int2 c = make_int2(0, 0);
    if (prefixed) {
      if (threadIdx.x == 0) c = part[blockIdx.x];
    } else {
      for (int b = threadIdx.x; b < (int)blockIdx.x; b += 64) c = add2(c, part[b]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      c.x += __shfl_xor(c.x, o, 64);
      c.y += __shfl_xor(c.y, o, 64);
    }
    if (threadIdx.x == 0) base = c;
  }
  long long beg = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanItems;
  // all loads in flight before any is used (index clamped into the array, value masked)
  int2 v[kScanItems];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) v[j] = in[min(beg + j, n - 1)];
  int2 s = make_int2(0, 0);
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (beg + j >= n) v[j] = make_int2(0, 0);
    s = add2(s, v[j]);
  }
  int2 total;
  int2 ex = block_exscan2(s, sh, total);  // synchronises the block: base is set after it
  ex = add2(ex, base);
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    long long i = beg + j;
    if (i < n) out[i] = ex;
    ex = add2(ex, v[j]);
  }
  if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = add2(base, total);
}

__global__ void k_scan_empty(int2* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = make_int2(0, 0);
}

size_t scan2_part_elems(long long n) { return (size_t)((std::max<long long>(n, 1) + kScanChunk - 1) / kScanChunk + 1); }

int Scan2::reserve(long long n) { return part.ensure(scan2_part_elems(n) * sizeof(int2)); }

int scan2_into(const int2* in, int2* out, long long n, int2* part, hipStream_t s) {
  if (n <= 0) {
    // a kernel, not a memset node: callers may capture this in a graph
    hipLaunchKernelGGL(k_scan_empty, dim3(1), dim3(64), 0, s, out);
    DRB_HIP(hipGetLastError());
    return DRB_OK;
  }
  const long long nb = (n + kScanChunk - 1) / kScanChunk;
  if (nb > (1LL << 30)) return fail(DRB_ERR_CAPACITY, "scan too large");
  // The reduce kernel reads contiguous chunks, the down kernel assigns kScanItems consecutive
  // elements per thread; both cover [b * kScanChunk, (b + 1) * kScanChunk).
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, part);
  const int prefixed = nb > kScanFold ? 1 : 0;
  if (prefixed) hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, s, part, (int)nb);
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, part, (int)nb, prefixed, out);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

int Scan2::run(const int2* in, int2* out, long long n, hipStream_t s) {
  if (n > 0) DRB_CHECK(reserve(n));
  return scan2_into(in, out, n, part.as<int2>(), s);
}

// --------------------------------------------------------------------- FP64 reduction
__global__ void k_reduce_partials(const double* __restrict__ part, int nb, int nq, double* __restrict__ out) {
  // One thread per quantity, fixed summation order over blocks: deterministic.
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[(size_t)b * nq + q];
  out[q] = s;
}

int reduce_partials(const double* partials, int nblocks, int nq, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_partials, dim3((nq + 63) / 64), dim3(64), 0, s, partials, nblocks, nq, out);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

}  // namespace drb

extern "C" const char* drb_last_error(void) { return drb::g_err.msg.c_str(); }
extern "C" int drb_version(void) { return 100; }
