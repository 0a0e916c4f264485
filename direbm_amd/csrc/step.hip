// The DiReBM step engine: collide -> disperse -> cell-thinning -> refine -> resample, on one GPU.
//
// Reference: direbm/warp/simulator.py:50-73 (GpuSimulator.step) and the stage kernels of
// direbm/warp/propagation.py.  Design (MI355X-first, no atomics on values, no per-step allocation
// once capacities are reached, one host synchronisation per step):
//
//   components are implicit (c = 7 i + k, position recomputed from the parent moment);
//   a 32-bit key = dense query-cell index (x-major, per instance) << sub_bits | sub-cell is
//   radix-sorted stably with the component index, which reproduces the reference's
//   radix_sort_pairs order (ties in ascending component index);
//   the sorted components are gathered into 16-byte records and a dense per-cell exclusive
//   prefix (components, control points) is built with one device scan -- this is both the
//   run-length encoding of the reference (runs = thinning cells) and the cell list every
//   neighbourhood query walks (in place of wp.HashGrid);
//   refine, the scatter normaliser, the per-control-point scatter gather and the gap-fill each
//   walk the cells of a disk, column by column, over contiguous ranges of the sorted arrays;
//   the reference's atomic scatter (propagation.py:186) becomes a gather in a fixed order, so
//   results are bit-for-bit reproducible.
#include "engine.h"

#include <mutex>

namespace drb {

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup w runs on XCD w % 8), so
// neighbouring blocks, whose windows overlap, fill eight different L2s.  The default remap deals
// chunks of DRB_XCD_CHUNK consecutive blocks to the XCDs in turn: neighbours share an L2 within a
// chunk, and every XCD still draws chunks from the whole range, so regions of unequal work (the
// two-level fine and coarse regions) stay balanced across XCDs.  (DRB_XCD: the contiguous j-th
// eighth per XCD, which unbalanced them.)  The grid is a multiple of 8 chunks.
__device__ __forceinline__ int xcd_block(int nb) {
#if DRB_XCD_CHUNK > 0
  // chunks of C consecutive blocks, chunk k on XCD k mod 8: neighbouring blocks share an L2 while
  // every XCD still draws work from the whole range.  A permutation of the blocks when the grid is
  // a multiple of 8 C (xblk makes large grids so); other grids keep the hardware order.
  constexpr int C = DRB_XCD_CHUNK;
  const int b = (int)blockIdx.x;
  if (nb % (8 * C) != 0) return b;
  const int x = b & 7, j = b >> 3;
  return (((j / C) << 3) + x) * C + (j % C);
#elif DRB_XCD
  const int per = nb >> 3;
  return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
#else
  return (int)blockIdx.x;
#endif
}

// Normaliser tiles over one-level grids (k_norm_tiles): kTileX x kTileY query cells, numbered per
// instance with kTileY-row tiles stacked along y within each band of kTileX columns (tile_column).
constexpr int kTileX = 3, kTileY = 2;
constexpr long long kTileMin = 1LL << 20;  // components from which a step uses them
__host__ __device__ __forceinline__ long long grid_tiles(const Inst& p) {
  return (long long)((p.nqx + kTileX - 1) / kTileX) * ((p.nqy + kTileY - 1) / kTileY);
}

// ---------------------------------------------------------------------------- bbox helpers
__device__ __forceinline__ void point_box(const Inst& p, float x, float y, int& lox, int& hix, int& loy, int& hiy) {
  int qx, qy, sub;
  qcell_of(x, y, p, qx, qy, sub);
  // Components lie within dx of their moment: widen by ceil(dx / qcs) + 2 query cells.
  int e = (int)ceilf(p.dx / p.qcs) + 2;
  lox = min(lox, qx - e);
  hix = max(hix, qx + e);
  loy = min(loy, qy - e);
  hiy = max(hiy, qy + e);
}

// Widen box[b] to cover [lox, hix] x [loy, hiy]: a read first, and atomics only where it grows.
__device__ __forceinline__ void box_atomic(int b, int lox, int hix, int loy, int hiy, BBox* box) {
  BBox cur = box[b];
  if (lox < cur.qxmin) atomicMin(&box[b].qxmin, lox);
  if (hix > cur.qxmax) atomicMax(&box[b].qxmax, hix);
  if (loy < cur.qymin) atomicMin(&box[b].qymin, loy);
  if (hiy > cur.qymax) atomicMax(&box[b].qymax, hiy);
}

// Merge per-thread boxes into box[b] with few atomics: wave reduction when the wave holds one
// instance, per-lane atomics otherwise.  Only lanes with b >= 0 contribute.
__device__ __forceinline__ void box_flush(int b, int lox, int hix, int loy, int hiy, BBox* box) {
  int bmin = b < 0 ? INT_MAX : b, bmax = b;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    bmin = min(bmin, __shfl_xor(bmin, o, 64));
    bmax = max(bmax, __shfl_xor(bmax, o, 64));
  }
  if (bmax < 0) return;
  if (bmin == bmax) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lox = min(lox, __shfl_xor(lox, o, 64));
      hix = max(hix, __shfl_xor(hix, o, 64));
      loy = min(loy, __shfl_xor(loy, o, 64));
      hiy = max(hiy, __shfl_xor(hiy, o, 64));
    }
    if ((threadIdx.x & 63) != 0) return;
    b = bmin;
  } else if (b < 0) {
    return;
  }
  box_atomic(b, lox, hix, loy, hiy, box);
}

// box_flush for a whole block (every thread of the block calls it): the waves that hold one
// instance are merged in LDS and each instance is widened once per block.  Atomics on one address
// serialise in the L2: a launch over points in cell order widens qxmax from nearly every wave, so
// one atomic per wave costs more than the launch's own work.
__device__ __forceinline__ void box_flush_block(int b, int lox, int hix, int loy, int hiy, BBox* box) {
  __shared__ int sb[TB / 64];
  __shared__ int4 sq[TB / 64];
  int bmin = b < 0 ? INT_MAX : b, bmax = b;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    bmin = min(bmin, __shfl_xor(bmin, o, 64));
    bmax = max(bmax, __shfl_xor(bmax, o, 64));
  }
  const bool uni = bmax >= 0 && bmin == bmax;
  if (bmax >= 0 && !uni && b >= 0) box_atomic(b, lox, hix, loy, hiy, box);  // a mixed wave: per lane
  if (uni) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lox = min(lox, __shfl_xor(lox, o, 64));
      hix = max(hix, __shfl_xor(hix, o, 64));
      loy = min(loy, __shfl_xor(loy, o, 64));
      hiy = max(hiy, __shfl_xor(hiy, o, 64));
    }
  }
  const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sb[wv] = uni ? bmin : -1;
    sq[wv] = make_int4(lox, hix, loy, hiy);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < nw; ++i) {
      if (sb[i] < 0) continue;
      int4 a = sq[i];
      for (int j = i + 1; j < nw; ++j)
        if (sb[j] == sb[i]) {
          const int4 c = sq[j];
          a = make_int4(min(a.x, c.x), max(a.y, c.y), min(a.z, c.z), max(a.w, c.w));
          sb[j] = -1;
        }
      box_atomic(sb[i], a.x, a.y, a.z, a.w, box);
    }
  }
}

__global__ void k_box_init(BBox* box, int* disp, int n) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  box[b] = BBox{INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  disp[b] = 0;
}

// The segment table of a step passed by value (up to kInstArg instances), written to the device
// copy; with box, the instances' boxes and displacements are reset as k_box_init does.
constexpr int kInstArg = 8;
struct InstArg {
  Inst v[kInstArg];
};
__global__ void k_set_inst(InstArg a, int n, Inst* __restrict__ dst, BBox* box, int* disp) {
  const int b = threadIdx.x;
  if (b >= n) return;
  dst[b] = a.v[b];
  if (box) {
    box[b] = BBox{INT_MAX, INT_MIN, INT_MAX, INT_MIN};
    disp[b] = 0;
  }
}

// Keep a cell inside the instance's dense grid (the periodic box, or the bounding box of every
// finite point).  Only a non-finite coordinate can fall outside; it is flagged and the step fails.
__device__ __forceinline__ void clamp_cell(const Inst& p, int& qx, int& qy, unsigned* status) {
  const int cx = min(max(qx, p.qx0), p.qx0 + p.nqx - 1), cy = min(max(qy, p.qy0), p.qy0 + p.nqy - 1);
  if (cx != qx || cy != qy) {
    atomicOr(status, ST_NONFINITE);
    qx = cx;
    qy = cy;
  }
}

// ------------------------------------------------------------------------------ kernels
// Collide (SoA planes, pitch; in place, or into fo with pitch opitch) and/or accumulate the moment
// bounding box.
__global__ __launch_bounds__(TB) void k_collide_box(float* __restrict__ f, long long pitch, const float* __restrict__ mx,
                                                    const float* __restrict__ my, const int* __restrict__ minst, int m,
                                                    const int* __restrict__ mdev, const Inst* __restrict__ ins,
                                                    BBox* box, unsigned* status, int do_collide, int do_box,
                                                    float* __restrict__ fo, long long opitch) {
  if (mdev) m = min(m, *mdev);  // graph mode: the live count is on the device
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = -1, lox = INT_MAX, hix = INT_MIN, loy = INT_MAX, hiy = INT_MIN;
  unsigned st = 0;
  if (i < m) {
    b = minst[i];
    const Inst& p = ins[b];
    if (do_collide) {
      float v[7];
#pragma unroll
      for (int k = 0; k < 7; ++k) v[k] = f[k * pitch + i];
      collide7(v, p.tau, st);
      float* o = fo ? fo : f;
      const long long op = fo ? opitch : pitch;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        if (!isfinite(v[k])) st |= ST_NONFINITE;
        o[k * op + i] = v[k];
      }
    }
    float x = mx[i], y = my[i];
    if (!isfinite(x) || !isfinite(y)) st |= ST_NONFINITE;
    if (do_box && !p.periodic && !(st & ST_NONFINITE)) point_box(p, x, y, lox, hix, loy, hiy);
    if (do_box && p.periodic && !(x >= 0.f && x < p.L && y >= 0.f && y < p.L)) st |= ST_RANGE;
    if (!do_box || p.periodic) b = -1;
  }
  if (st) atomicOr(status, st);
  if (do_box) box_flush(b, lox, hix, loy, hiy, box);
}

// Keys of the 7 components of each moment.  With an obstacle, a component k != 0 that would land
// inside it gets the key `dead` (above every cell: it sorts last and belongs to no cell); its
// bounce parts are keyed by k_bounce_emit.
__global__ __launch_bounds__(TB) void k_keys(const float* __restrict__ mx, const float* __restrict__ my,
                                             const int* __restrict__ minst, int m, const int* __restrict__ mdev,
                                             const Inst* __restrict__ ins, int sub_bits, uint32_t* __restrict__ keys,
                                             unsigned* status, uint32_t dead, int* __restrict__ clear,
                                             long long nclear, float2* __restrict__ pxy) {
  // clear: the cell counts k_mark fills later in the step (in place of a memset launch)
  if (clear)
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nclear; q += (long long)gridDim.x * blockDim.x)
      clear[q] = 0;
  if (mdev) m = min(m, *mdev);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const Inst& p = ins[minst[i]];
  const float x0 = mx[i], y0 = my[i];
  // pxy: the positions again as pairs, for k_mark's gathers (one 8-byte load per component)
  if (pxy) pxy[i] = make_float2(x0, y0);
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    float x, y;
    comp_pos(x0, y0, k, p, x, y);
    if (p.obstacle && k && inside2(x, y, p)) {
      keys[7 * (long long)i + k] = dead;
      continue;
    }
    int qx, qy, sub;
    qcell_of(x, y, p, qx, qy, sub);
    clamp_cell(p, qx, qy, status);
    keys[7 * (long long)i + k] = (uint32_t)((qindex(p, qx, qy) << sub_bits) | sub);
  }
}

// First kernel of a step without obstacles, one pass over the moments: the moments collided
// (k_collide_box's work; in place, or into fo with pitch opitch), the components' keys and the
// moment positions as pairs written and the cell counts cleared (k_keys's).  SINGLE (a single
// simulation): the segment table travels as the argument p and is written to its device copy dst,
// with the box and displacement reset (k_set_inst's work); otherwise the table is ins and the
// first nreset boxes and displacements are reset (k_box_init's work).  pxy / pxyi (optional): the
// moment positions as pairs (a single simulation), or as {x, y, instance, -} records (several), for
// k_mark's gathers: one load per component instead of two or three.  A wave's 448 keys (AoS,
// component 7 i + k) go out through LDS as coalesced 16-byte stores: written as seven 4-byte
// stores per lane, 28 bytes apart, each store instruction touched 28 cache lines.
template <bool SINGLE>
__global__ __launch_bounds__(TB) void k_head(Inst p, Inst* __restrict__ dst, const Inst* __restrict__ ins,
                                             const int* __restrict__ minst, BBox* box, int* disp, int nreset,
                                             float* __restrict__ f, long long pitch, float* __restrict__ fo,
                                             long long opitch, const float* __restrict__ mx,
                                             const float* __restrict__ my, int m, const int* __restrict__ mdev,
                                             int sub_bits, uint32_t* __restrict__ keys, unsigned* status,
                                             int* __restrict__ clear, long long nclear, float2* __restrict__ pxy,
                                             float4* __restrict__ pxyi, int* __restrict__ toff, int ntab) {
  __shared__ __attribute__((aligned(16))) uint32_t sk[TB * 7];
  // toff (one-level grids): the first normaliser tile of each instance, and the total
  if (toff && blockIdx.x == 0 && threadIdx.x == 0) {
    if (SINGLE) {
      toff[0] = 0;
      toff[1] = (int)grid_tiles(p);
    } else {
      int acc = 0;
      for (int b = 0; b < ntab; ++b) {
        toff[b] = acc;
        acc += (int)grid_tiles(ins[b]);
      }
      toff[ntab] = acc;
    }
  }
  if (SINGLE) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *dst = p;
      *box = BBox{INT_MAX, INT_MIN, INT_MAX, INT_MIN};
      *disp = 0;
    }
  } else if (blockIdx.x == 0) {
    for (int b = threadIdx.x; b < nreset; b += blockDim.x) {
      box[b] = BBox{INT_MAX, INT_MIN, INT_MAX, INT_MIN};
      disp[b] = 0;
    }
  }
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nclear; q += (long long)gridDim.x * blockDim.x)
    clear[q] = 0;
  if (mdev) m = min(m, *mdev);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, base = i - lane;
  if (base >= m) return;  // the whole wave is past the end (a wave-uniform exit)
  const bool active = i < m;
  uint32_t* wk = sk + (threadIdx.x - lane) * 7;  // this wave's 448 words
  if (active) {
    const int b = SINGLE || !minst ? 0 : minst[i];
    const Inst& q = SINGLE ? p : ins[b];
    unsigned st = 0;
    float v[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] = f[k * pitch + i];
    collide7(v, q.tau, st);
    float* o = fo ? fo : f;
    const long long op = fo ? opitch : pitch;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      if (!isfinite(v[k])) st |= ST_NONFINITE;
      o[k * op + i] = v[k];
    }
    const float x0 = mx[i], y0 = my[i];
    if (!isfinite(x0) || !isfinite(y0)) st |= ST_NONFINITE;
    if (pxy) pxy[i] = make_float2(x0, y0);
    if (pxyi) pxyi[i] = make_float4(x0, y0, __int_as_float(b), 0.f);
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      float x, y;
      comp_pos(x0, y0, k, q, x, y);
      int qx, qy, sub;
      qcell_of(x, y, q, qx, qy, sub);
      clamp_cell(q, qx, qy, status);
      wk[7 * lane + k] = (uint32_t)((qindex(q, qx, qy) << sub_bits) | sub);
    }
    if (st) atomicOr(status, st);
  }
  // the wave's keys, [7 base, 7 base + cnt), from LDS: 16-byte stores (7 base is a multiple of 448,
  // so 16-byte aligned), then the last cnt % 4 words
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int cnt = 7 * min(64, m - base);
  uint32_t* out = keys + 7 * (long long)base;
  for (int j = lane; j < cnt / 4; j += 64)
    reinterpret_cast<uint4*>(out)[j] = reinterpret_cast<const uint4*>(wk)[j];
  const int t = (cnt & ~3) + lane;
  if (t < cnt) out[t] = wk[t];
}

// Keys of free points (row-major (n, 2)), instance 0.  Used by the stage entry points.
__global__ __launch_bounds__(TB) void k_keys_points(const float* __restrict__ pos, int n, const Inst* __restrict__ ins,
                                                    uint32_t* __restrict__ keys, unsigned* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Inst& p = ins[0];
  int qx, qy, sub;
  qcell_of(pos[2 * i], pos[2 * i + 1], p, qx, qy, sub);
  clamp_cell(p, qx, qy, status);
  keys[i] = (uint32_t)qindex(p, qx, qy);
}

// Number of bounce parts of each moment (obstacles only), as (count, 0) for the int2 scan.
__global__ __launch_bounds__(TB) void k_bounce_count(const float* __restrict__ mx, const float* __restrict__ my,
                                                     const int* __restrict__ minst, int m,
                                                     const Inst* __restrict__ ins, int2* __restrict__ bcnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const Inst& p = ins[minst[i]];
  int cnt = 0;
  if (p.obstacle) {
    const float x0 = mx[i], y0 = my[i];
    for (int k = 1; k < 7; ++k) {
      float x, y;
      comp_pos(x0, y0, k, p, x, y);
      if (!inside2(x, y, p)) continue;
      int pj[2];
      double pw[2];
      cnt += bounce2(x0, y0, x, y, k, p, pj, pw);
    }
  }
  bcnt[i] = make_int2(cnt, 0);
}

// Bounce parts as explicit components after the N implicit ones: record (x, y, value, direction |
// instance << 3) and key at component index N + offset, so that within a cell the parts follow the
// regular components in (moment, direction, part) order.  A part whose position is inside the
// obstacle stays at its moment's position (direbm/reference/simulator.py:113-115).  The value is
// the FP64 product f_k w rounded to fp32.
__global__ __launch_bounds__(TB) void k_bounce_emit(const float* __restrict__ mx, const float* __restrict__ my,
                                                    const int* __restrict__ minst, const float* __restrict__ f,
                                                    long long pitch, int m, const Inst* __restrict__ ins,
                                                    int sub_bits, const int2* __restrict__ boff, int N,
                                                    uint32_t* __restrict__ keys, float4* __restrict__ ex,
                                                    unsigned* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int o = boff[i].x;
  if (boff[i + 1].x == o) return;
  const int b = minst[i];
  const Inst& p = ins[b];
  const float x0 = mx[i], y0 = my[i];
  for (int k = 1; k < 7; ++k) {
    float x1, y1;
    comp_pos(x0, y0, k, p, x1, y1);
    if (!inside2(x1, y1, p)) continue;
    int pj[2];
    double pw[2];
    const int np = bounce2(x0, y0, x1, y1, k, p, pj, pw);
    const double v = (double)f[k * pitch + i];
    for (int a = 0; a < np; ++a) {
      float qx, qy;
      comp_pos(x0, y0, pj[a], p, qx, qy);
      if (inside2(qx, qy, p)) {
        qx = x0;
        qy = y0;
      }
      int cx, cy, sub;
      qcell_of(qx, qy, p, cx, cy, sub);
      clamp_cell(p, cx, cy, status);
      keys[(size_t)N + o] = (uint32_t)((qindex(p, cx, cy) << sub_bits) | sub);
      ex[o] = make_float4(qx, qy, (float)(v * pw[a]), __int_as_float(pj[a] | (b << 3)));
      ++o;
    }
  }
}

// Control points that stay after resampling: those outside their instance's obstacle
// (direbm/reference/simulator.py:79-80), as (flag, 0) over the whole capacity for the int2 scan.
__global__ __launch_bounds__(TB) void k_keep(const int* __restrict__ Ptot, const float* __restrict__ cpx,
                                             const float* __restrict__ cpy, const int* __restrict__ cpinst,
                                             long long pcap, const Inst* __restrict__ ins, int2* __restrict__ kc) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= pcap) return;
  bool keep = r < *Ptot;
  if (keep) {
    const Inst& p = ins[cpinst ? cpinst[r] : 0];
    keep = !(p.obstacle && inside2(cpx[r], cpy[r], p));
  }
  kc[r] = make_int2(keep ? 1 : 0, 0);
}

// Moments per instance after the obstacle drop: kept points within the instance's range of
// control points (the instances' cells, hence their points, are contiguous in sorted order).
__global__ void k_counts_keep(const int2* __restrict__ rp, const int2* __restrict__ koff,
                              const Inst* __restrict__ ins, int ninst, int* __restrict__ counts) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ninst) return;
  const Inst& p = ins[b];
  const long long c0 = p.cell_off, c1 = c0 + (long long)p.nqx * p.nqy;
  counts[b] = koff[rp[c1].y].x - koff[rp[c0].y].x;
}

__global__ __launch_bounds__(TB) void k_box_points(const float* __restrict__ pos, int n, const Inst* __restrict__ ins,
                                                   BBox* box, unsigned* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = -1, lox = INT_MAX, hix = INT_MIN, loy = INT_MAX, hiy = INT_MIN;
  if (i < n) {
    float x = pos[2 * i], y = pos[2 * i + 1];
    if (!isfinite(x) || !isfinite(y)) {
      atomicOr(status, ST_NONFINITE);
    } else {
      b = 0;
      int qx, qy, sub;
      qcell_of(x, y, ins[0], qx, qy, sub);
      lox = hix = qx;
      loy = hiy = qy;
    }
  }
  box_flush(b, lox, hix, loy, hiy, box);
}

__global__ __launch_bounds__(TB) void k_check_dirs(const int32_t* __restrict__ dirs, int n, unsigned* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (unsigned)dirs[i] > 6u) atomicOr(status, ST_RANGE);
}

// Gather sorted components into records (position, exp(value), direction | instance) and their
// values; the first component of each query cell writes the counts of its cell's components (and,
// with count_runs, of its thinning runs = distinct keys).  The counts come from wave ballots: a
// cell's components end at the next cell start within the wave, and the wave's last cell is
// followed into the next chunks of 64 keys by the whole wave.
// Chunks of 64 components per wave in the step's k_mark launches.
constexpr int kMarkChunks = 2;

template <int CH>
__global__ __launch_bounds__(TB) void k_mark(const uint32_t* __restrict__ K, const uint32_t* __restrict__ V, int n,
                                             const int* __restrict__ ndev,
                                             int sub_bits, const float* __restrict__ mx, const float* __restrict__ my,
                                             const float* __restrict__ f, long long pitch,
                                             const int* __restrict__ minst, const float* __restrict__ pc,
                                             const float* __restrict__ vc, const int* __restrict__ dc,
                                             const Inst* __restrict__ ins, Comp* __restrict__ comp,
                                             float* __restrict__ cv, int* __restrict__ cnt2, int count_runs,
                                             int nimp, const float4* __restrict__ ex, const float2* __restrict__ pxy,
                                             const float4* __restrict__ pxyi) {
  if (ndev) n = min(n, *ndev);
  const int lane = threadIdx.x & 63;
  // a wave takes CH chunks of 64 consecutive components; every chunk's loads are issued before
  // any chunk's record is written, so the gathers of the chunks overlap
  const int base0 = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x - lane) * CH);
  if (base0 >= n) return;  // the whole wave is past the end
  uint32_t cidx[CH], keyv[CH];
#pragma unroll
  for (int ch = 0; ch < CH; ++ch) {
    const int s = base0 + 64 * ch + lane;
    cidx[ch] = s < n ? V[s] : 0u;
    keyv[ch] = s < n ? K[s] : 0xffffffffu;
  }
  Comp rec[CH];
  float val[CH];
  if (!pc && !ex && !pxyi && !minst && pxy) {
    // one instance, moment positions in pxy (the plain step): every chunk's gathers issued before
    // any is used, the instance's parameters as uniform loads; a lane past the end reads
    // component 0 (it exists: base0 < n) and writes nothing
    const Inst& p0 = ins[0];
    float2 m2[CH];
    float fv[CH];
    uint32_t kk[CH];
#pragma unroll
    for (int ch = 0; ch < CH; ++ch) {
      const uint32_t c = cidx[ch], i = c / 7u;
      kk[ch] = c - 7u * i;
      m2[ch] = pxy[i];
      fv[ch] = f[kk[ch] * pitch + i];
    }
#pragma unroll
    for (int ch = 0; ch < CH; ++ch) {
      comp_pos(m2[ch].x, m2[ch].y, (int)kk[ch], p0, rec[ch].x, rec[ch].y);
      val[ch] = fv[ch];
      rec[ch].bits = (int)kk[ch];
    }
  } else if (!pc && !ex && pxyi) {
    // several instances, position and instance in one record (the batched step): as above, each
    // chunk's instance parameters loaded once its record has arrived
    float4 m4[CH];
    float fv[CH];
    uint32_t kk[CH];
#pragma unroll
    for (int ch = 0; ch < CH; ++ch) {
      const uint32_t c = cidx[ch], i = c / 7u;
      kk[ch] = c - 7u * i;
      m4[ch] = pxyi[i];
      fv[ch] = f[kk[ch] * pitch + i];
    }
#pragma unroll
    for (int ch = 0; ch < CH; ++ch) {
      const int b = __float_as_int(m4[ch].z);
      comp_pos(m4[ch].x, m4[ch].y, (int)kk[ch], ins[b], rec[ch].x, rec[ch].y);
      val[ch] = fv[ch];
      rec[ch].bits = (int)kk[ch] | (b << 3);
    }
  } else {
#pragma unroll
  for (int ch = 0; ch < CH; ++ch) {
    const int s = base0 + 64 * ch + lane;
    if (s >= n) continue;
    const uint32_t c = cidx[ch];
    Comp& r = rec[ch];
    float& v = val[ch];
    if (pc) {
      r.x = pc[2 * (size_t)c];
      r.y = pc[2 * (size_t)c + 1];
      v = vc ? vc[c] : 0.f;
      r.bits = dc ? (dc[c] & 7) : 0;
    } else if (ex && c >= (uint32_t)nimp) {  // an obstacle's bounce part (k_bounce_emit)
      const float4 e = ex[c - nimp];
      r.x = e.x;
      r.y = e.y;
      v = e.z;
      r.bits = __float_as_int(e.w);
    } else {
      const uint32_t i = c / 7u, k = c - 7u * i;
      int b;
      if (pxyi) {  // several instances: position and instance in one record
        const float4 m4 = pxyi[i];
        b = __float_as_int(m4.z);
        comp_pos(m4.x, m4.y, (int)k, ins[b], r.x, r.y);
      } else {
        b = minst ? minst[i] : 0;  // minst: several instances
        if (pxy) {
          const float2 m2 = pxy[i];
          comp_pos(m2.x, m2.y, (int)k, ins[b], r.x, r.y);
        } else {
          comp_pos(mx[i], my[i], (int)k, ins[b], r.x, r.y);
        }
      }
      v = f[k * pitch + i];
      r.bits = (int)k | (b << 3);
    }
  }
  }
#pragma unroll
  for (int ch = 0; ch < CH; ++ch) {
    const int s = base0 + 64 * ch + lane;
    if (s >= n) continue;
    rec[ch].w = expf(val[ch]);  // refine weight exp(f) (direbm/warp/propagation.py:100)
    comp[s] = rec[ch];
    cv[s] = val[ch];
  }
#pragma unroll
  for (int ch = 0; ch < CH; ++ch) {
    const int base = base0 + 64 * ch;
    if (base >= n) break;  // wave-uniform
    const int s = base + lane;
    const bool active = s < n;
    const uint32_t key = keyv[ch];
    uint32_t prev = __shfl_up(key, 1, 64);
    if (lane == 0) prev = base > 0 ? K[base - 1] : ~key;  // ~key: a different cell
    const uint32_t q = key >> sub_bits;
    const bool first = active && q != (prev >> sub_bits);
    const unsigned long long fm = __ballot(first), rm = __ballot(active && key != prev);
    if (!fm) continue;  // every key here continues a cell that started in an earlier chunk
    // the chunk's last cell, followed past the chunk: components and runs beyond it
    const int lf = 63 - __clzll((long long)fm);
    int tail_c = 0, tail_r = 0;
    if (base + 64 < n) {
      const uint32_t tq = __shfl(q, lf, 64);
      uint32_t pk = __shfl(key, 63, 64);
      for (int t0 = base + 64; t0 < n; t0 += 64) {
        const int t = t0 + lane;
        const uint32_t kt = t < n ? K[t] : 0xffffffffu;
        uint32_t kp = __shfl_up(kt, 1, 64);
        if (lane == 0) kp = pk;
        // keys ascend, so the lanes still in cell tq are a prefix of the chunk
        const unsigned long long sm = __ballot(t < n && (kt >> sub_bits) == tq);
        const int len = __popcll(sm);
        tail_c += len;
        tail_r += __popcll(__ballot(lane < len && kt != kp));
        if (len < 64) break;
        pk = __shfl(kt, 63, 64);
      }
    }
    if (first) {
      const unsigned long long above = lane == 63 ? 0ull : fm & ~((2ull << lane) - 1ull);
      const int e = above ? __ffsll((long long)above) - 1 : min(64, n - base);  // end lane of the cell
      const unsigned long long upto = e >= 64 ? ~0ull : (1ull << e) - 1ull;
      int nc = e - lane, nr = __popcll(rm & upto & ~((1ull << lane) - 1ull));
      if (lane == lf) {
        nc += tail_c;
        nr += tail_r;
      }
      cnt2[2 * (size_t)q] = nc;
      if (count_runs) cnt2[2 * (size_t)q + 1] = nr;
    }
  }
}

// Sorted free control points (stage path): slot positions and per-cell control point counts.
__global__ __launch_bounds__(TB) void k_mark_cp(const uint32_t* __restrict__ K, const uint32_t* __restrict__ V, int n,
                                                const float* __restrict__ cp, float* __restrict__ sx,
                                                float* __restrict__ sy, int* __restrict__ cnt2) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t c = V[s];
  sx[s] = cp[2 * (size_t)c];
  sy[s] = cp[2 * (size_t)c + 1];
  const uint32_t q = K[s];
  if (s == 0 || K[s - 1] != q) {
    int nc = 0;
    for (int t = s; t < n && K[t] == q; ++t) ++nc;
    cnt2[2 * (size_t)q + 1] = nc;
  }
}

// Per-instance counts after the scan: counts[b] = control points of instance b.
__global__ void k_counts(const int2* __restrict__ rp, const Inst* __restrict__ ins, int ninst, int* __restrict__ counts,
                         const int* __restrict__ ptot) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ninst) return;
  if (ptot) {  // a single simulation with spawned points: every point is emitted
    counts[b] = *ptot;
    return;
  }
  const Inst& p = ins[b];
  long long c0 = p.cell_off, c1 = p.cell_off + (long long)p.nqx * p.nqy;
  counts[b] = rp[c1].y - rp[c0].y;
}

// Thinning centroids, one lane per run (direbm/warp/propagation.py:62-68).  A wave holds 64
// consecutive sorted components; a lane whose key differs from the one before heads a run (the
// members of one thinning cell, contiguous in key order) and sums its members' positions in fp32 in
// member order -- within the wave by shuffles, past its end from global memory -- then divides by
// float(n).  A run's control-point index is its cell's first index (rp.y) plus the runs of the
// cell before it.  Keys past the grid (an obstacle's dead components, sorted last) form no run.
// With crun given (a recorded step) it also writes the run of every component, the run lengths
// and starts.  bclear: int2 counts cleared by a grid-stride loop (the refined points' bin counts,
// in place of a memset launch).  counts (optional): each instance's control points, as k_counts
// writes them, in place of its launch.
__global__ __launch_bounds__(TB) void k_runs(const int2* __restrict__ rp, long long ncells, int n,
                                             const int* __restrict__ ndev, int sub_bits,
                                             const uint32_t* __restrict__ K, const Comp* __restrict__ comp,
                                             float* __restrict__ cpx, float* __restrict__ cpy, int* __restrict__ cpq,
                                             int* __restrict__ cpinst, int* __restrict__ crun, int* __restrict__ cpn,
                                             int* __restrict__ cps, int2* __restrict__ bclear, long long nclear,
                                             const Inst* __restrict__ cins, int ninst, int* __restrict__ counts) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nclear; i += (long long)gridDim.x * blockDim.x)
    bclear[i] = make_int2(0, 0);
  // counts: the control points of each instance (the next step's moments), as k_counts
  if (counts && blockIdx.x == 0)
    for (int b = threadIdx.x; b < ninst; b += blockDim.x) {
      const Inst& p = cins[b];
      counts[b] = rp[p.cell_off + (long long)p.nqx * p.nqy].y - rp[p.cell_off].y;
    }
  if (ndev) n = min(n, *ndev);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, base = s - lane;
  if (base >= n) return;  // the whole wave is past the end
  const bool active = s < n;
  const uint32_t key = active ? K[s] : 0xffffffffu;
  float x = 0.f, y = 0.f;
  int bits = 0;
  if (active) {
    const Comp c = comp[s];
    x = c.x;
    y = c.y;
    bits = c.bits;
  }
  uint32_t prev = __shfl_up(key, 1, 64);
  if (lane == 0) prev = base > 0 ? K[base - 1] : ~key;
  const bool head = active && key != prev;
  const unsigned long long hm = __ballot(head);
  if (!hm) return;  // the wave lies inside a run headed by an earlier wave
  const int nv = min(64, n - base);
  const int h0 = __ffsll((long long)hm) - 1;
  const long long q0 = (long long)(__shfl(key, h0, 64) >> sub_bits);
  if (q0 >= ncells) return;  // dead keys only
  // runs of the first head's cell before it: that cell may have started in an earlier wave (the
  // lanes below h0 hold no head)
  const int cs = rp[q0].x;
  int before = 0;
  for (int t0 = cs; t0 < base; t0 += 64) {
    const int t = t0 + lane;
    const uint32_t kt = t < base ? K[t] : 0u;
    uint32_t kp = __shfl_up(kt, 1, 64);
    if (lane == 0) kp = t > cs ? K[t - 1] : ~kt;  // cs heads the cell's first run
    before += __popcll(__ballot(t < base && kt != kp));
  }
  const unsigned long long below = hm & ((1ull << lane) - 1ull);
  const int r = rp[q0].y + before + __popcll(below);  // meaningful on head lanes
  // each head's members in the wave: lanes [lane, e), e the next head or the end of the valid lanes
  const unsigned long long above = lane == 63 ? 0ull : hm & ~((2ull << lane) - 1ull);
  const int e = head ? (above ? __ffsll((long long)above) - 1 : nv) : lane;
  const int len = e - lane;
  int maxlen = len;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) maxlen = max(maxlen, __shfl_xor(maxlen, o, 64));
  float ax = 0.f, ay = 0.f;
  for (int i = 0; i < maxlen; ++i) {
    const int src = min(lane + i, 63);
    const float xi = __shfl(x, src, 64), yi = __shfl(y, src, 64);
    if (i < len) {
      ax = fadd(ax, xi);
      ay = fadd(ay, yi);
    }
  }
  // crun of the wave's members of runs headed here: each lane takes its head's index
  if (crun) {
    const unsigned long long hb = lane == 63 ? hm : hm & ((2ull << lane) - 1ull);
    const int hl = hb ? 63 - __clzll((long long)hb) : 0;
    const int rr = __shfl(r, hl, 64);
    if (active && hb && (long long)(key >> sub_bits) < ncells) crun[s] = rr;
  }
  // The last run may reach past the wave: its further members, 64 at a time (coalesced loads, then
  // the same shuffle walk for the run's head).
  int cnt = len;
  const int hlast = 63 - __clzll((long long)hm);
  const uint32_t lkey = __shfl(key, hlast, 64);
  if (nv == 64 && __shfl(e, hlast, 64) == 64 && (long long)(lkey >> sub_bits) < ncells) {
    const int rl = __shfl(r, hlast, 64);
    for (int t0 = base + 64; t0 < n; t0 += 64) {
      const int t = t0 + lane;
      const uint32_t kt = t < n ? K[t] : 0xffffffffu;
      float xt = 0.f, yt = 0.f;
      if (t < n && kt == lkey) {
        const Comp c = comp[t];
        xt = c.x;
        yt = c.y;
      }
      // keys ascend: the lanes still in the run are a prefix of the chunk
      const int m = __popcll(__ballot(t < n && kt == lkey));
      for (int i = 0; i < m; ++i) {
        const float xi = __shfl(xt, i, 64), yi = __shfl(yt, i, 64);
        if (lane == hlast) {
          ax = fadd(ax, xi);
          ay = fadd(ay, yi);
        }
      }
      if (lane == hlast) cnt += m;
      if (crun && lane < m) crun[t] = rl;
      if (m < 64) break;
    }
  }
  if (!head) return;
  const long long q = key >> sub_bits;
  if (q >= ncells) return;
  cpx[r] = fdiv(ax, (float)cnt);
  cpy[r] = fdiv(ay, (float)cnt);
  if (cpq) cpq[r] = (int)q;
  if (cpinst) cpinst[r] = bits >> 3;
  if (crun) {
    cpn[r] = cnt;
    cps[r] = s;
  }
}

// Neighbourhood kernels.  A group of kGroup lanes cooperates on one query point: the lanes
// stride through the contiguous sorted ranges of the point's disk window (coalesced loads, one
// window walk per group instead of per lane), each keeps partial sums, and the group combines
// them with a butterfly over lanes (__shfl_xor: both partners add the same two values, so every
// lane ends with identical bits).  The visiting order and the reduction tree are fixed, so the
// results are bit-for-bit reproducible.


// Refine (direbm/warp/propagation.py:88-112): kappa from the direction mask of components within
// dx; kappa <= kappa_hard keeps the point, else the exp(val)-weighted centroid.  wsum (recorded
// steps) receives the weight sum.
template <int PM, int G, int K>
__global__ __launch_bounds__(TB) void k_refine(const int* __restrict__ Ptot, int pcap, const float* __restrict__ x0s,
                                               const float* __restrict__ y0s, const int* __restrict__ cpinst,
                                               const int* __restrict__ cpq, const Inst* __restrict__ ins,
                                               const int* __restrict__ rp2, const Comp* __restrict__ comp,
                                               float* __restrict__ ox, float* __restrict__ oy,
                                               unsigned char* __restrict__ kap, int* __restrict__ disp,
                                               float* __restrict__ wsum, unsigned* __restrict__ status,
                                               uint32_t* __restrict__ bkey, int2* __restrict__ bcnt,
                                               int* __restrict__ bsl) {
  const int lane = threadIdx.x & (G - 1);
  const int P = Ptot ? *Ptot : pcap;
  const int ngroups = gridDim.x * (blockDim.x / G);
  const int rlast = ((P + ngroups - 1) / ngroups) * ngroups;  // every group runs the same trip count
  for (int r = (xcd_block(gridDim.x) * blockDim.x + threadIdx.x) / G; r < rlast; r += ngroups) {
  const bool active = r < P;
  const int b = active ? (cpinst ? cpinst[r] : 0) : 0;
  const Inst& p = ins[b];
  const float x0 = active ? x0s[r] : 0.f, y0 = active ? y0s[r] : 0.f;
  const float t2 = p.t_dx, L = p.L;
  const int per = PM == 2 ? p.periodic : PM;
  unsigned mask = 0;
  float sw = 0.f, sx = 0.f, sy = 0.f;
  if (active)
    for_window(p, rp2, x0, y0, p.dx, 0, [&](int a, int e, int wc) {
      auto run = [&](auto plain) {
        auto take = [&](const Comp& c, bool in) {
          // reference: wp.length(q - p) <= dx with q the component (minimum image: wrap code)
          float ddx = fsub(c.x, x0), ddy = fsub(c.y, y0);
          if (!decltype(plain)::value) {
            ddx = wrap_shift(ddx, wc & 3, L);
            ddy = wrap_shift(ddy, wc >> 2, L);
          }
          // Branch-free: a rejected candidate adds weight +0, which leaves every sum's bits
          // unchanged (no sum is ever -0).  A branch here would always be taken by some lane of
          // the wave, and it let the compiler sink the record's second half into it, behind a
          // wait for every load in flight.
          const bool ok = in && fadd(fmul(ddx, ddx), fmul(ddy, ddy)) <= t2;
          const float w = ok ? c.w : 0.f;
          // The direction bit, shifted by the record's low five bits (dir | inst << 3): it lands at
          // bit dir + 8 j, j < 4, and the four bytes are folded onto dir after the walk (one select
          // and one shift-or per candidate).
          mask |= (ok ? 1u : 0u) << (c.bits & 31);
          // the weighted sums are continuous quantities: fused multiply-adds (one rounding each)
          sw = fadd(sw, w);
          if (per) {
            sx = __builtin_fmaf(w, ddx, sx);
            sy = __builtin_fmaf(w, ddy, sy);
          } else {
            sx = __builtin_fmaf(w, c.x, sx);
            sy = __builtin_fmaf(w, c.y, sy);
          }
        };
        // K candidates per lane in flight: their records are loaded before any is used, and
        // they are added in ascending index order as one at a time would be.  Full passes need no
        // range test or clamp (one address, immediate offsets); the last, partial pass clamps its
        // loads to the range and masks the candidates past its end.
        int s0 = a + lane;
        for (; s0 + (K - 1) * G < e; s0 += K * G) {
          Comp cc[K];
#pragma unroll
          for (int u = 0; u < K; ++u) cc[u] = comp[s0 + u * G];
#pragma unroll
          for (int u = 0; u < K; ++u) take(cc[u], true);
        }
        if (s0 < e) {
          Comp cc[K];
#pragma unroll
          for (int u = 0; u < K; ++u) cc[u] = comp[min(s0 + u * G, e - 1)];
#pragma unroll
          for (int u = 0; u < K; ++u) take(cc[u], s0 + u * G < e);
        }
      };
      if (PM == 0 || wc == 0) run(std::true_type{});
      else run(std::false_type{});
    });
  mask = group_or<G>(mask);
  mask = (mask | mask >> 8 | mask >> 16 | mask >> 24) & 0x7fu;
  sw = group_sum<G>(sw);
  sx = group_sum<G>(sx);
  sy = group_sum<G>(sy);
  if (!active || lane != 0) continue;
  const int kappa = __popc(mask);
  float nx = x0, ny = y0;
  if (kappa > p.kappa_hard) {
    if (per) {
      nx = wrap_coord(fadd(x0, fdiv(sx, sw)), L, p.cs, p.nfine);
      ny = wrap_coord(fadd(y0, fdiv(sy, sw)), L, p.cs, p.nfine);
    } else {
      nx = fdiv(sx, sw);
      ny = fdiv(sy, sw);
    }
  }
  ox[r] = nx;
  oy[r] = ny;
  if (kap) kap[r] = (unsigned char)kappa;
  if (wsum) wsum[r] = sw;
  // Weights exp(val) overflow for val > 88.7 (an unstable run): the mean is then not finite.
  const bool finite = isfinite(nx) && isfinite(ny);
  if (!finite && status) atomicOr(status, ST_NONFINITE);
  if (disp && cpq && finite) {
    // Chebyshev distance in query cells between the refined position and the registered cell.
    const long long rel = (long long)cpq[r] - p.cell_off;
    const int rx = (int)(rel / p.nqy) + p.qx0, ry = (int)(rel % p.nqy) + p.qy0;
    int qx, qy, sub;
    qcell_of(nx, ny, p, qx, qy, sub);
    int ddx = abs(qx - rx), ddy = abs(qy - ry);
    if (per) {
      ddx = min(ddx, p.nq - ddx);
      ddy = min(ddy, p.nq - ddy);
    }
    // A refined point is a weighted mean of components within dx of its old position, so it
    // moves at most dx; the cap keeps every window that disp widens bounded whatever the data.
    const int dd = min(max(ddx, ddy), (int)(p.dx * p.iq) + 2);
    if (dd > 0 && dd > disp[b]) atomicMax(&disp[b], dd);
  }
  if (bkey) {
    // the bin of the refined point, as k_cpkeys gives it, and its arrival slot in that bin
    int qx, qy, sub;
    qcell_of(nx, ny, p, qx, qy, sub);
    clamp_cell(p, qx, qy, status);
    const uint32_t q = (uint32_t)qindex(p, qx, qy);
    bkey[r] = q;
    bsl[r] = atomicAdd(&bcnt[q].y, 1);
  }
  }
}

// Scatter normaliser per component (direbm/warp/propagation.py:168-178):
// sw = sum over control points within r of (r - d).  Stores q = val / sw (0 when sw <= 0: the
// component is dropped), so the gather needs no division per pair.
//
// One lane per component.  A wave holds 64 consecutive sorted components, which lie in a few
// cells of one column: the wave walks the window of the box spanned by their positions with
// wave-uniform bounds (scalar loads of the candidate control points, scalar loop control) and
// every lane tests every candidate.  A lane's hits, and their order, are those of its own
// window; a wave spanning two instances walks per-lane windows instead.
// q also replaces the record's refine weight (comp[s].w): refine, its only reader, has run, and the
// gather then reads position, q and direction in one 16-byte load.
// The normaliser sum of each lane's component c (active lanes): over the refined control points
// within r of c, the sum of (r - d).  When the wave's active lanes share one instance and span a
// compact box, the wave walks the window of that box with wave-uniform bounds (scalar loads) and
// every lane sees every candidate -- in its own window's order, so its sum does not depend on the
// box; otherwise each active lane walks its own window.  Inactive lanes return 0.
template <int PM>
__device__ __forceinline__ float norm_sum(bool active, const Comp& c, const Inst* __restrict__ ins,
                                          const int* __restrict__ rp2, const float* __restrict__ cpx,
                                          const float* __restrict__ cpy, const int* __restrict__ disp) {
  const int b = c.bits >> 3;
  const Inst& p = ins[b];
  // wave-uniform summary of the lanes' instance and position box
  int bmin = active ? b : INT_MAX, bmax = active ? b : INT_MIN;
  float x0 = active ? c.x : INFINITY, x1 = active ? c.x : -INFINITY;
  float y0 = active ? c.y : INFINITY, y1 = active ? c.y : -INFINITY;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    bmin = min(bmin, __shfl_xor(bmin, o, 64));
    bmax = max(bmax, __shfl_xor(bmax, o, 64));
    x0 = fminf(x0, __shfl_xor(x0, o, 64));
    x1 = fmaxf(x1, __shfl_xor(x1, o, 64));
    y0 = fminf(y0, __shfl_xor(y0, o, 64));
    y1 = fmaxf(y1, __shfl_xor(y1, o, 64));
  }
  if (bmax == INT_MIN) return 0.f;  // no active lane
  const float rr = p.rr, t2 = p.t_rr, L = p.L;
  const int per = PM == 2 ? p.periodic : PM;
  // A compact box only: a wave that wraps to the next column (or across a periodic seam) spans a
  // tall box and walks per-lane windows instead.
  const float lim = 3.f * (rr + p.qcs);
  const bool uniform = bmin == bmax && x1 - x0 <= lim && y1 - y0 <= lim;
  float sw = 0.f;
  if (uniform) {
    const int ub = __builtin_amdgcn_readfirstlane(bmin);
    const Inst& pu = ins[ub];
    const float ux0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x0)));
    const float ux1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x1)));
    const float uy0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(y0)));
    const float uy1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(y1)));
    const int uext = disp ? __builtin_amdgcn_readfirstlane(disp[ub]) : 0;
    // Every caller lists each candidate under the cell of its own position (the step bins the
    // refined control points first), so the ranges' wrap codes give the minimum image.
    for_box_window(pu, rp2 + 1, ux0, ux1, uy0, uy1, rr, uext, [&](int a, int e, int wc) {
      const int ua = __builtin_amdgcn_readfirstlane(a), ue = __builtin_amdgcn_readfirstlane(e);
      const int uw = __builtin_amdgcn_readfirstlane(wc);
      // unrolled so the scalar loads of several candidates are in flight together; the sum
      // still runs in ascending j
      auto run = [&](auto plain) {
#pragma unroll DRB_NORM_UNROLL
        for (int j = ua; j < ue; ++j) {
          // reference: wp.length(cp - p) with p the component
          const float s2 = dist2_w<decltype(plain)::value>(cpx[j], cpy[j], c.x, c.y, uw, L);
          if (s2 <= t2) sw = fadd(sw, fsub(rr, fast_sqrt(s2)));
        }
      };
      if (PM == 0 || uw == 0) run(std::true_type{});
      else run(std::false_type{});
    });
  } else if (active) {
    for_window(p, rp2 + 1, c.x, c.y, rr, disp ? disp[b] : 0, [&](int a, int e, int wc) {
      auto run = [&](auto plain) {
        for (int j = a; j < e; ++j) {
          const float s2 = dist2_w<decltype(plain)::value>(cpx[j], cpy[j], c.x, c.y, wc, L);
          if (s2 <= t2) sw = fadd(sw, fsub(rr, fast_sqrt(s2)));
        }
      };
      if (PM == 0 || wc == 0) run(std::true_type{});
      else run(std::false_type{});
    });
  }
  return sw;
}

template <int PM>
__global__ __launch_bounds__(TB) void k_norm(int n, const int* __restrict__ ndev, Comp* __restrict__ comp,
                                             const float* __restrict__ cv,
                                             const Inst* __restrict__ ins, const int* __restrict__ rp2,
                                             const float* __restrict__ cpx, const float* __restrict__ cpy,
                                             const int* __restrict__ disp, float* __restrict__ swc) {
  if (ndev) n = min(n, *ndev);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = s < n;
  if (s - (int)(threadIdx.x & 63) >= n) return;  // the whole wave is past the end
  Comp c;
  c.x = c.y = 0.f;
  c.bits = 0;
  if (active) c = comp[s];
  const float sw = norm_sum<PM>(active, c, ins, rp2, cpx, cpy, disp);
  if (!active) return;
  const float q = sw > 0.f ? fdiv(cv[s], sw) : 0.f;
  swc[s] = sw;
  comp[s].w = q;
}

// ------------------------------------------------------------- normaliser over cell tiles
// One-level grids hold about 7 components per cell, so k_norm's 64 consecutive sorted components are
// a strip of ~9 cells of one column, whose box window is ~9 x 17 cells at a radius of 4 cells.
// Renumbered tile-major (tiles of kTileX x kTileY cells, ~42 components), 64 consecutive
// components span one or two tiles stacked in y: a box of ~3 x 3 cells, a window of ~11 x 11, a
// fifth fewer candidates for the same components.  The sums are k_norm's bit for bit (norm_sum:
// every lane meets its own candidates in its own window's order whatever the box).
//   k_tile_counts: the components of each tile (.x), for a scan into each tile's first position;
//   k_wave_start:  the tile holding each wave's first position;
//   k_norm_tiles:  wave w takes positions [64 w, 64 w + 64), each lane's component found from
//                  its tile's column segments.
// toff: the first tile of each instance and the total (k_head); crp: the component cell prefix.
__device__ __forceinline__ int tile_instance(const int* __restrict__ toff, int ninst, int t) {
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (toff[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
// Cell range of column k of tile t of instance p: components [crp[c0], crp[c1]) (.x).  The bands of
// kTileX columns are walked up and down in turn (odd bands from the top), so that consecutive
// tiles are neighbours also across a band's end, and a wave spanning two tiles keeps a compact box.
__device__ __forceinline__ void tile_column(const Inst& p, int t, int k, long long& c0, long long& c1) {
  const int nty = (p.nqy + kTileY - 1) / kTileY;
  const int band = t / nty, r = t % nty, ty = (band & 1) ? nty - 1 - r : r;
  const int cx = band * kTileX + k, cy0 = ty * kTileY, cy1 = min(cy0 + kTileY, p.nqy);
  if (cx >= p.nqx) {
    c0 = c1 = 0;
    return;
  }
  const long long base = p.cell_off + (long long)cx * p.nqy;
  c0 = base + cy0;
  c1 = base + cy1;
}
// Per tile: tcnt = {components, those of its first two columns}, tseg = {first column's start,
// second's start less the first's count, third's start less the first two's counts, first's
// count}, so that tile position i is component (i < n0 ? a0 : i < n01 ? b1 : b2) + i.
__global__ __launch_bounds__(TB) void k_tile_counts(const int* __restrict__ toff, int ninst, const Inst* __restrict__ ins,
                                                    const int* __restrict__ crp, int ntiles_cap, int2* __restrict__ tcnt,
                                                    int4* __restrict__ tseg) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles_cap) return;
  int a[kTileX], n[kTileX];
#pragma unroll
  for (int k = 0; k < kTileX; ++k) a[k] = n[k] = 0;
  if (t < toff[ninst]) {
    const int b = tile_instance(toff, ninst, t);
    const Inst& p = ins[b];
#pragma unroll
    for (int k = 0; k < kTileX; ++k) {
      long long c0, c1;
      tile_column(p, t - toff[b], k, c0, c1);
      a[k] = crp[2 * c0];
      n[k] = crp[2 * c1] - a[k];
    }
  }
  static_assert(kTileX == 3, "three column segments per tile");
  tcnt[t] = make_int2(n[0] + n[1] + n[2], n[0] + n[1]);
  tseg[t] = make_int4(a[0], a[1] - n[0], a[2] - n[0] - n[1], n[0]);
}
__global__ __launch_bounds__(TB) void k_wave_start(const int* __restrict__ toff, int ninst, const int2* __restrict__ tpos,
                                                   int* __restrict__ wstart) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= toff[ninst]) return;
  const int a = tpos[t].x, e = tpos[t + 1].x;
  for (int w = (a + 63) >> 6; (long long)w * 64 < e; ++w) wstart[w] = t;
}
template <int PM>
__global__ __launch_bounds__(TB) void k_norm_tiles(const int* __restrict__ toff, int ninst, const int2* __restrict__ tpos,
                                                   const int* __restrict__ wstart, const int2* __restrict__ tcnt,
                                                   const int4* __restrict__ tseg,
                                                   Comp* __restrict__ comp, const float* __restrict__ cv,
                                                   const Inst* __restrict__ ins, const int* __restrict__ rp2,
                                                   const float* __restrict__ cpx, const float* __restrict__ cpy,
                                                   float* __restrict__ swc) {
  const int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  const int nt = toff[ninst];
  const int total = tpos[nt].x;
  if ((long long)w * 64 >= total) return;  // the whole wave is past the end
  const int pos = w * 64 + lane;
  const bool active = pos < total;
  Comp c;
  c.x = c.y = 0.f;
  c.bits = 0;
  int s = 0;
  if (active) {
    int t = wstart[w];
    while (pos >= tpos[t + 1].x) ++t;  // a wave spans one or two tiles
    const int i = pos - tpos[t].x;
    const int4 g = tseg[t];
    const int n01 = tcnt[t].y;
    s = (i < g.w ? g.x : i < n01 ? g.y : g.z) + i;
    c = comp[s];
  }
  const float sw = norm_sum<PM>(active, c, ins, rp2, cpx, cpy, (const int*)nullptr);
  if (!active) return;
  swc[s] = sw;
  comp[s].w = sw > 0.f ? fdiv(cv[s], sw) : 0.f;
}

// Scatter as a gather (replaces the atomics of direbm/warp/propagation.py:181-186): each control
// point sums (r - d) * (val_c / sw_c) over components within r.  A component's record holds
// val_c / sw_c in its weight slot (written by k_norm).  The direction of a component is
// data-dependent: each lane keeps its seven sums in LDS (slot k * TB + thread, bank-conflict free
// whatever the directions) and adds t to its own direction's slot.  With bem, the empty-direction
// mask is also written at the point's bin slot (slot[r]).
template <int PM, int G, int K>
__global__ __launch_bounds__(TB) void k_gather(const int* __restrict__ Ptot, int pcap, const float* __restrict__ cpx,
                                               const float* __restrict__ cpy, const int* __restrict__ cpinst,
                                               const Inst* __restrict__ ins, const int* __restrict__ rp2,
                                               const Comp* __restrict__ comp, float* __restrict__ cpf, long long pitch,
                                               unsigned char* __restrict__ emask, const int* __restrict__ slot,
                                               unsigned char* __restrict__ bem) {
  const int lane = threadIdx.x & (G - 1);
  const int P = Ptot ? *Ptot : pcap;
  const int ngroups = gridDim.x * (blockDim.x / G);
  const int rlast = ((P + ngroups - 1) / ngroups) * ngroups;  // every group runs the same trip count
  // seven direction sums per lane, and an eighth slot that no result reads: direction bits 7 (no
  // record carries them) land there, so the slot index needs no test
  __shared__ float sacc[8 * TB];
  float* my = sacc + threadIdx.x;
  for (int r = (xcd_block(gridDim.x) * blockDim.x + threadIdx.x) / G; r < rlast; r += ngroups) {
  const bool active = r < P;
  const Inst& p = ins[active && cpinst ? cpinst[r] : 0];
  const float x = active ? cpx[r] : 0.f, y = active ? cpy[r] : 0.f, rr = p.rr, t2 = p.t_rr, L = p.L;
  const int per = PM == 2 ? p.periodic : PM;
  float acc[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) my[k * TB] = 0.f;
  if (active)
    for_window(p, rp2, x, y, rr, 0, [&](int a, int e, int wc) {
      auto run = [&](auto plain) {
        auto take = [&](const Comp& c, bool in) {
          // the squared distance has the same bits either way round
          const float s2 = dist2_w<decltype(plain)::value>(c.x, c.y, x, y, wc, L);
          // c.w = val / sw.  Branch-free: a rejected candidate adds a zero product (r - d) * 0 to
          // its direction's sum in one fused multiply-add, which leaves the sum's bits unchanged
          // (no sum is ever -0); so does an accepted one with c.w = 0 (a dropped component).  Some
          // lane of the wave takes any branch here.
          const bool ok = in && s2 <= t2;
          float* slot = my + (c.bits & 7) * TB;
          *slot = __builtin_fmaf(fsub(rr, fast_sqrt(s2)), ok ? c.w : 0.f, *slot);
        };
        // as in k_refine: full passes unclamped, the last partial pass clamped and masked
        int s0 = a + lane;
        for (; s0 + (K - 1) * G < e; s0 += K * G) {
          Comp cc[K];
#pragma unroll
          for (int u = 0; u < K; ++u) cc[u] = comp[s0 + u * G];
#pragma unroll
          for (int u = 0; u < K; ++u) take(cc[u], true);
        }
        if (s0 < e) {
          Comp cc[K];
#pragma unroll
          for (int u = 0; u < K; ++u) cc[u] = comp[min(s0 + u * G, e - 1)];
#pragma unroll
          for (int u = 0; u < K; ++u) take(cc[u], s0 + u * G < e);
        }
      };
      if (PM == 0 || wc == 0) run(std::true_type{});
      else run(std::false_type{});
    });
#pragma unroll
  for (int k = 0; k < 7; ++k) acc[k] = group_sum<G>(my[k * TB]);
  if (!active || lane != 0) continue;
  unsigned em = 0;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    cpf[k * pitch + r] = acc[k];
    if (acc[k] < p.eps) em |= 1u << k;
  }
  emask[r] = (unsigned char)em;
  if (bem) bem[slot[r]] = (unsigned char)em;  // the mask again, in bin order, for the gap-fill walk
  }
}

// Gap-fill and emit (direbm/warp/propagation.py:203-218): an empty direction i (< eps) becomes
// (r / sw) f_rest[i], sw = sum of (r - d) over control points within r also empty in i (self
// included).  Writes the new moment to slot dest = perm ? perm[r] : r.  A grid-stride loop, one
// lane per control point at a time (only points with a gap walk a window).
// gsum (recorded steps, same pitch as cpf) receives the gap-fill sums of the empty directions.
template <int PM>
__global__ __launch_bounds__(TB) void k_gapfill_emit(
    const int* __restrict__ Ptot, int pcap, const float* __restrict__ cpx, const float* __restrict__ cpy,
    const int* __restrict__ cpinst, const Inst* __restrict__ ins, const int* __restrict__ rp2,
    const float* __restrict__ cpf, long long pitch, const unsigned char* __restrict__ emask,
    const uint32_t* __restrict__ perm, float* __restrict__ ox, float* __restrict__ oy, long long pos_stride,
    float* __restrict__ of, long long fsi, long long fsk, int* __restrict__ oinst, unsigned* status,
    float* __restrict__ gsum, const int2* __restrict__ koff, const unsigned char* __restrict__ wm,
    const float* __restrict__ wx, const float* __restrict__ wy) {
  const int P = Ptot ? *Ptot : pcap;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < P; r += gridDim.x * blockDim.x) {
    // with obstacles (koff), the points outside them are emitted in order at their prefix index
    if (koff && koff[r + 1].x == koff[r].x) continue;
    const int b = cpinst ? cpinst[r] : 0;
    const Inst& p = ins[b];
    const float x = cpx[r], y = cpy[r], rr = p.rr, t2 = p.t_rr, L = p.L;
    const int per = PM == 2 ? p.periodic : PM;
    float out[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) out[k] = cpf[k * pitch + r];
    const unsigned e = emask[r];
    if (e) {
      float sw[7];
#pragma unroll
      for (int k = 0; k < 7; ++k) sw[k] = 0.f;
      // The window runs over lists of the points by the cell of their own position (wm, wx, wy:
      // empty mask and position per list slot), so it needs no displacement margin.  kInflight
      // slots are loaded before any is used: the walk is a chain of dependent loads otherwise,
      // and the few points with a gap set the time of the whole launch.
      for_window(p, rp2 + 1, x, y, rr, 0, [&](int a, int en) {
        for (int j0 = a; j0 < en; j0 += kInflight) {
          unsigned mm[kInflight];
          float xx[kInflight], yy[kInflight];
#pragma unroll
          for (int u = 0; u < kInflight; ++u) {
            const int j = min(j0 + u, en - 1);
            mm[u] = wm[j];
            xx[u] = wx[j];
            yy[u] = wy[j];
          }
#pragma unroll
          for (int u = 0; u < kInflight; ++u) {
            const unsigned m = j0 + u < en ? e & mm[u] : 0u;
            if (m) {
              const float s2 = dist2_32(xx[u], yy[u], x, y, L, per);
              if (s2 <= t2) {
                const float w = fsub(rr, fast_sqrt(s2));
#pragma unroll
                for (int k = 0; k < 7; ++k)
                  if ((m >> k) & 1u) sw[k] = fadd(sw[k], w);
              }
            }
          }
        }
      });
#pragma unroll
      for (int k = 0; k < 7; ++k)
        if (((e >> k) & 1u) && sw[k] > 0.f) out[k] = fmul(fdiv(rr, sw[k]), p.frest[k]);
      if (gsum)
#pragma unroll
        for (int k = 0; k < 7; ++k) gsum[k * pitch + r] = ((e >> k) & 1u) ? sw[k] : 0.f;
    } else if (gsum) {
#pragma unroll
      for (int k = 0; k < 7; ++k) gsum[k * pitch + r] = 0.f;
    }
    const long long dst = koff ? (long long)koff[r].x : perm ? (long long)perm[r] : (long long)r;
    ox[dst * pos_stride] = x;
    oy[dst * pos_stride] = y;
#pragma unroll
    for (int k = 0; k < 7; ++k) of[dst * fsi + k * fsk] = out[k];
    if (oinst) oinst[dst] = b;
    const bool finite = isfinite(x) && isfinite(y);
    if (!finite && status) atomicOr(status, ST_NONFINITE);
  }
}

// Copy SoA moments to row-major host-layout device buffers.
__global__ __launch_bounds__(TB) void k_pack(const float* __restrict__ mx, const float* __restrict__ my,
                                             const float* __restrict__ f, long long pitch, int m,
                                             float* __restrict__ pos, float* __restrict__ fo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  pos[2 * (size_t)i] = mx[i];
  pos[2 * (size_t)i + 1] = my[i];
#pragma unroll
  for (int k = 0; k < 7; ++k) fo[7 * (size_t)i + k] = f[k * pitch + i];
}

__global__ __launch_bounds__(TB) void k_unpack(const float* __restrict__ pos, const float* __restrict__ fi, int m,
                                               float* __restrict__ mx, float* __restrict__ my, float* __restrict__ f,
                                               long long pitch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  mx[i] = pos[2 * (size_t)i];
  my[i] = pos[2 * (size_t)i + 1];
#pragma unroll
  for (int k = 0; k < 7; ++k) f[k * pitch + i] = fi[7 * (size_t)i + k];
}

__global__ void k_fill_inst(int* __restrict__ minst, const long long* __restrict__ offs, int ninst) {
  // one block per instance
  const int b = blockIdx.x;
  if (b >= ninst) return;
  for (long long i = offs[b] + threadIdx.x; i < offs[b + 1]; i += blockDim.x) minst[i] = b;
}

// FP64 diagnostics: kDiagChunks blocks per instance each reduce a slice (fixed-order tree), then
// one thread per (instance, quantity) sums the slices in order.  Deterministic.
constexpr int kDiagChunks = 128;

__global__ __launch_bounds__(TB) void k_diag(const float* __restrict__ f, long long pitch,
                                             const long long* __restrict__ moff, const float* __restrict__ cv,
                                             const float* __restrict__ swc, const int2* __restrict__ rp,
                                             const Inst* __restrict__ ins, const float* __restrict__ cpf,
                                             long long cpitch, const unsigned char* __restrict__ emask, int have_step,
                                             const int2* __restrict__ koff, const int* __restrict__ ptot,
                                             double* __restrict__ part) {
  __shared__ double sh[5][TB];
  const int b = blockIdx.y, chunk = blockIdx.x, t = threadIdx.x;
  double a[5] = {0, 0, 0, 0, 0};
  auto slice = [&](long long lo, long long hi, long long& s0, long long& s1) {
    const long long per = (hi - lo + kDiagChunks - 1) / kDiagChunks;
    s0 = min(hi, lo + per * chunk);
    s1 = min(hi, s0 + per);
  };
  long long i0, i1;
  slice(moff[b], moff[b + 1], i0, i1);
  for (long long i = i0 + t; i < i1; i += TB) {
    for (int k = 0; k < 7; ++k) {
      double v = f[k * pitch + i];
      a[0] += v;
      a[1] += v * (double)kCx[k];
      a[2] += v * (double)kCy[k];
    }
  }
  if (have_step) {
    const Inst& p = ins[b];
    const long long c0 = p.cell_off, c1 = p.cell_off + (long long)p.nqx * p.nqy;
    slice(rp[c0].x, rp[c1].x, i0, i1);
    for (long long s = i0 + t; s < i1; s += TB)
      if (!(swc[s] > 0.f)) a[3] += cv[s];
    // with spawned points (a single simulation), the points are [0, ptot)
    slice(ptot ? 0 : rp[c0].y, ptot ? *ptot : rp[c1].y, i0, i1);
    for (long long r = i0 + t; r < i1; r += TB) {
      // with obstacles the emitted moments are the kept points, at their prefix index
      long long o = r;
      if (koff) {
        if (koff[r + 1].x == koff[r].x) continue;
        o = koff[r].x;
      }
      const unsigned e = emask[r];
      for (int k = 0; k < 7; ++k)
        if ((e >> k) & 1u) a[4] += (double)f[k * pitch + o] - (double)cpf[k * cpitch + r];
    }
  }
  for (int q = 0; q < 5; ++q) sh[q][t] = a[q];
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int q = 0; q < 5; ++q) sh[q][t] += sh[q][t + w];
    __syncthreads();
  }
  if (t == 0)
    for (int q = 0; q < 5; ++q) part[((size_t)b * kDiagChunks + chunk) * 5 + q] = sh[q][0];
}

__global__ void k_diag_final(const double* __restrict__ part, int ninst, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 5 * ninst) return;
  const int b = i / 5, q = i - 5 * b;
  double s = 0.0;
  for (int c = 0; c < kDiagChunks; ++c) s += part[((size_t)b * kDiagChunks + c) * 5 + q];
  out[i] = s;
}

// Graph-mode helpers (kernels only: a captured graph must contain no memset nodes).
// Query cell (global dense index) of each refined control point, and the number of points per
// such cell (integer atomics: the counts, and so the bins, do not depend on their order).
__global__ __launch_bounds__(TB) void k_cpkeys(const int* __restrict__ Ptot, int pcap, const float* __restrict__ cpx,
                                               const float* __restrict__ cpy, const int* __restrict__ cpinst,
                                               const Inst* __restrict__ ins, uint32_t* __restrict__ keys,
                                               int2* __restrict__ cnt, unsigned* status, int* __restrict__ bsl) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (Ptot ? *Ptot : pcap)) return;
  const Inst& p = ins[cpinst ? cpinst[r] : 0];
  int qx, qy, sub;
  qcell_of(cpx[r], cpy[r], p, qx, qy, sub);
  clamp_cell(p, qx, qy, status);
  const uint32_t q = (uint32_t)qindex(p, qx, qy);
  keys[r] = q;
  const int slot = atomicAdd(&cnt[q].y, 1);
  if (bsl) bsl[r] = slot;
}

// Each point at its bin's start plus the arrival slot the counting atomic gave it: an order
// within a bin that differs from run to run, which k_binsort turns into ascending index.
__global__ __launch_bounds__(TB) void k_binplace(const int* __restrict__ Ptot, int pcap, const uint32_t* __restrict__ K,
                                                 const int* __restrict__ bsl, const int2* __restrict__ brp,
                                                 int* __restrict__ bidx) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (Ptot ? *Ptot : pcap)) return;
  bidx[brp[K[r]].y + bsl[r]] = r;
}

// Per-block partial of box_flush_block (every thread of the block calls it): the box of the
// block's lanes of one instance goes to (*pb, *pq) with no atomics; lanes of other instances (a
// block over an instance boundary) widen their boxes with atomics.  *pb = -1: no partial.
__device__ __forceinline__ void box_partial_block(int b, int lox, int hix, int loy, int hiy, BBox* box, int* pb,
                                                  int4* pq) {
  __shared__ int sb[TB / 64];
  __shared__ int4 sq[TB / 64];
  int bmin = b < 0 ? INT_MAX : b, bmax = b;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    bmin = min(bmin, __shfl_xor(bmin, o, 64));
    bmax = max(bmax, __shfl_xor(bmax, o, 64));
  }
  const bool uni = bmax >= 0 && bmin == bmax;
  if (bmax >= 0 && !uni && b >= 0) box_atomic(b, lox, hix, loy, hiy, box);
  if (uni) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lox = min(lox, __shfl_xor(lox, o, 64));
      hix = max(hix, __shfl_xor(hix, o, 64));
      loy = min(loy, __shfl_xor(loy, o, 64));
      hiy = max(hiy, __shfl_xor(hiy, o, 64));
    }
  }
  const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sb[wv] = uni ? bmin : -1;
    sq[wv] = make_int4(lox, hix, loy, hiy);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int pbv = -1;
    int4 a = make_int4(INT_MAX, INT_MIN, INT_MAX, INT_MIN);
    for (int i = 0; i < nw; ++i) {
      if (sb[i] < 0) continue;
      const int4 c = sq[i];
      if (pbv < 0 || sb[i] == pbv) {
        pbv = sb[i];
        a = make_int4(min(a.x, c.x), max(a.y, c.y), min(a.z, c.z), max(a.w, c.w));
      } else {
        box_atomic(sb[i], c.x, c.y, c.z, c.w, box);
      }
    }
    *pb = pbv;
    *pq = a;
  }
}

// One block (every thread calls it): the box partials of k_binsort into the boxes, walked in
// block order (an instance's partials are consecutive), so each instance is widened about once.
__device__ __forceinline__ void box_merge_parts(int nparts, const int* __restrict__ pb, const int4* __restrict__ pq,
                                                BBox* box) {
  const int per = (nparts + TB - 1) / TB;
  const int i0 = threadIdx.x * per, i1 = min(i0 + per, nparts);
  int cb = -1, lox = INT_MAX, hix = INT_MIN, loy = INT_MAX, hiy = INT_MIN;
  for (int i = i0; i < i1; ++i) {
    const int b = pb[i];
    if (b < 0) continue;
    if (b != cb) {
      if (cb >= 0) box_atomic(cb, lox, hix, loy, hiy, box);
      cb = b;
      lox = loy = INT_MAX;
      hix = hiy = INT_MIN;
    }
    const int4 c = pq[i];
    lox = min(lox, c.x);
    hix = max(hix, c.y);
    loy = min(loy, c.z);
    hiy = max(hiy, c.w);
  }
  box_flush_block(cb, lox, hix, loy, hiy, box);
}
// One thread per bin: the bin's points in ascending index (an insertion sort of the few arrivals:
// the order a stable sort by bin gives), their positions in bin order and the bin slot of each.
// With box: the next step's bounding box of every open instance, as one partial per block (pb,
// pq; k_box_merge combines them): the binned points are the next step's moments, and a point's
// bin is its query cell (point_box's, for every finite point).  (Merging in the last block to
// finish instead, after a device-scope fence in every block, made the bin stage 52 -> 134 us: the
// fence writes the L2 back.)
__global__ __launch_bounds__(TB) void k_binsort(long long ncells, const int2* __restrict__ brp, int* __restrict__ bidx,
                                                const float* __restrict__ cpx, const float* __restrict__ cpy,
                                                float* __restrict__ bx, float* __restrict__ by,
                                                int* __restrict__ slot, const int* __restrict__ cpinst,
                                                const Inst* __restrict__ ins, BBox* box, int* __restrict__ pb,
                                                int4* __restrict__ pq) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int b = -1, lox = INT_MAX, hix = INT_MIN, loy = INT_MAX, hiy = INT_MIN;
  if (q < ncells) {
    const int a = brp[q].y, e = brp[q + 1].y;
    for (int i = a + 1; i < e; ++i) {
      const int v = bidx[i];
      int j = i - 1;
      while (j >= a && bidx[j] > v) {
        bidx[j + 1] = bidx[j];
        --j;
      }
      bidx[j + 1] = v;
    }
    for (int i = a; i < e; ++i) {
      const int r = bidx[i];
      bx[i] = cpx[r];
      by[i] = cpy[r];
      slot[r] = i;
    }
    if (box && e > a) {
      b = cpinst ? cpinst[bidx[a]] : 0;  // cpinst: several instances
      const Inst& p = ins[b];
      if (p.periodic) {
        b = -1;
      } else {
        // components lie within dx of their moment: widen by ceil(dx / qcs) + 2 query cells
        const long long rel = q - p.cell_off;
        const int qx = (int)(rel / p.nqy) + p.qx0, qy = (int)(rel % p.nqy) + p.qy0;
        const int w = (int)ceilf(p.dx / p.qcs) + 2;
        lox = qx - w;
        hix = qx + w;
        loy = qy - w;
        hiy = qy + w;
      }
    }
  }
  if (box) box_partial_block(b, lox, hix, loy, hiy, box, pb + blockIdx.x, pq + blockIdx.x);
}

__global__ __launch_bounds__(TB) void k_box_merge(int nparts, const int* __restrict__ pb, const int4* __restrict__ pq,
                                                  BBox* box) {
  box_merge_parts(nparts, pb, pq, box);
}

// Refine weights of the component records, exp(val) as k_mark writes them (k_norm replaces them).
__global__ __launch_bounds__(TB) void k_weights(const float* __restrict__ cv, int n, Comp* __restrict__ comp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) comp[s].w = expf(cv[s]);
}

__global__ void k_fill_i32(int* __restrict__ p, long long n, int v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
// Graph mode with open instances: a step's segment table on the device, from the boxes the step
// before left -- the grid layout() builds on the host from the same boxes, field for field (origin,
// extent, cell offset, window margin) -- then the boxes and displacements reset for the step.  A
// grid beyond the captured capacity of ncap cells (or a box that layout() refuses) makes the step
// run empty: its moment count becomes 0 and every entry an empty grid, so no kernel of it or of a
// later step of the replay reads the state entering it or writes the other buffer, and sv records
// the step for the host to resume from.  counts: the instances' moments entering the step.
// sv = {a step ran empty, steps completed, moments entering the empty step, its control points
// (k_graph_counts; 0: its grid did not fit), its counts[ninst]}.  One thread.
__device__ void graph_layout(Inst* __restrict__ ins, int ninst, BBox* __restrict__ box, int* __restrict__ disp,
                             const int* __restrict__ counts, int* __restrict__ cnt, long long ncap,
                             int* __restrict__ sv, unsigned* __restrict__ status) {
  if (sv[0]) {  // an earlier step of the replay ran empty: so do the rest
    cnt[0] = cnt[1] = 0;
    return;
  }
  long long off = 0;
  bool fits = true;
  for (int b = 0; b < ninst && fits; ++b) {
    Inst& p = ins[b];
    int qx0 = 0, qy0 = 0, nqx = 0, nqy = 0;
    if (p.periodic) {
      nqx = nqy = p.nq;
    } else if (counts[b] > 0) {
      const BBox bb = box[b];
      if (bb.qxmin > bb.qxmax) {
        atomicOr(status, ST_NONFINITE);  // moments but no finite position
        fits = false;
        break;
      }
      const long long nx = (long long)bb.qxmax - bb.qxmin + 1, ny = (long long)bb.qymax - bb.qymin + 1;
      if (nx > (1LL << 28) || ny > (1LL << 28) || nx * ny > (1LL << 31)) {
        fits = false;
        break;
      }
      qx0 = bb.qxmin;
      qy0 = bb.qymin;
      nqx = (int)nx;
      nqy = (int)ny;
    }
    p.qx0 = qx0;
    p.qy0 = qy0;
    p.nqx = nqx;
    p.nqy = nqy;
    p.cell_off = off;
    off += (long long)nqx * nqy;
    const double far = fmax(fmax(fmax(fabs((double)qx0), fabs((double)qx0 + nqx)),
                                 fmax(fabs((double)qy0), fabs((double)qy0 + nqy))),
                            (double)p.nq) * (p.two_level ? 2.0 : 1.0);
    p.margin = (float)((double)p.qcs * fmax(1e-3, 16.0 * far * 0x1p-24));
  }
  if (!fits || off > ncap) {
    sv[0] = 1;
    sv[2] = cnt[0];
    sv[3] = 0;
    for (int b = 0; b < ninst; ++b) {
      sv[4 + b] = counts[b];
      ins[b].qx0 = ins[b].qy0 = ins[b].nqx = ins[b].nqy = 0;
      ins[b].cell_off = 0;
    }
    cnt[0] = cnt[1] = 0;
    return;
  }
  for (int b = 0; b < ninst; ++b) {
    box[b] = BBox{INT_MAX, INT_MIN, INT_MAX, INT_MIN};
    disp[b] = 0;
  }
}

// End of a replayed step: cnt = {moments, components} of the next step from this step's
// control-point total; upd accumulates the moments entering each step (the point-update count of
// a drb_sim_step call).  Open instances: the step is counted as completed unless it ran empty,
// and the next step is laid out (its boxes are final here).
__global__ void k_graph_next(const int* __restrict__ Ptot, int* __restrict__ cnt, unsigned long long* __restrict__ upd,
                             GraphLayout g) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  *upd += (unsigned long long)cnt[0];
  cnt[0] = *Ptot;
  cnt[1] = 7 * *Ptot;
  if (g.ins) {
    if (!g.sv[0]) ++g.sv[1];
    graph_layout(g.ins, g.ninst, g.box, g.disp, g.counts, cnt, g.ncap, g.sv, g.status);
  }
}
// Start of a replay of m moments: counts, update sum and record reset; open instances: the first
// step laid out.
__global__ void k_graph_set(int* __restrict__ cnt, int m, unsigned long long* __restrict__ upd, GraphLayout g) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  cnt[0] = m;
  cnt[1] = 7 * m;
  *upd = 0ull;
  if (g.ins) {
    g.sv[0] = g.sv[1] = 0;
    graph_layout(g.ins, g.ninst, g.box, g.disp, g.counts, cnt, g.ncap, g.sv, g.status);
  }
}

// Graph mode with open instances, after the cell scan (one block): a step whose control points
// exceed the captured capacity pcap runs empty from here on -- its point total, moment and component
// counts become 0, so the thinning, refine, resample and emit kernels after it write nothing -- and
// sv records it as graph_layout does, with the moments of each instance entering the step (counts,
// before this kernel overwrites them).  Then the moments of each instance leaving the step, as
// k_counts.
__global__ __launch_bounds__(TB) void k_graph_counts(const int2* __restrict__ rp, const Inst* __restrict__ ins,
                                                     int ninst, int* __restrict__ counts, int* __restrict__ Ptot,
                                                     int pcap, int* __restrict__ cnt, int* __restrict__ sv) {
  if (threadIdx.x == 0 && !sv[0] && *Ptot > pcap) {
    sv[0] = 1;
    sv[2] = cnt[0];
    sv[3] = *Ptot;
    for (int b = 0; b < ninst; ++b) sv[4 + b] = counts[b];
    *Ptot = 0;
    cnt[0] = cnt[1] = 0;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < ninst; b += blockDim.x) {
    const Inst& p = ins[b];
    const long long c0 = p.cell_off, c1 = p.cell_off + (long long)p.nqx * p.nqy;
    counts[b] = rp[c1].y - rp[c0].y;
  }
}

// ------------------------------------------------------------------------ soft_outer spawn
// The per-object solver's soft_outer spawn (direbm/reference/simulator.py:141-166) in this step:
// a control point with kappa_hard < kappa <= kappa_soft adds a point at its pre-refine position +
// n SOFT_OFFSET dx, n the unit vector of the summed directions of its components within dx.  The
// sum is kept as integers (2 c_x and the signed count of the c_y = +-sqrt(3)/2 directions), so it
// is exact, and |sum| is 0 (balanced: no spawn) or >= 1; the position is formed in FP64 and
// rounded once.  Only the soft points walk their window (one lane each); refine is unchanged.
__constant__ static const int kIx2[8] = {0, 2, 1, -1, -2, -1, 1, 0};
__constant__ static const int kIy[8] = {0, 0, 1, 1, 0, -1, -1, 0};

__global__ __launch_bounds__(TB) void k_spawn_flag(const int* __restrict__ Ptot, const float* __restrict__ x0s,
                                                   const float* __restrict__ y0s, const int* __restrict__ cpinst,
                                                   const Inst* __restrict__ ins, const int* __restrict__ rp2,
                                                   const Comp* __restrict__ comp, const unsigned char* __restrict__ kap,
                                                   double off_unit, int2* __restrict__ flag, float* __restrict__ sx,
                                                   float* __restrict__ sy) {
  const int P = *Ptot;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < P; r += gridDim.x * blockDim.x) {
    const Inst& p = ins[cpinst ? cpinst[r] : 0];
    const int kappa = kap[r];
    int f = 0;
    if (kappa > p.kappa_hard && kappa <= p.kappa_soft) {
      const float x0 = x0s[r], y0 = y0s[r], t2 = p.t_dx, L = p.L;
      int ix2 = 0, iy = 0;
      // the components of kappa: refine's predicate (fp32 distance to the pre-refine position)
      for_window(p, rp2, x0, y0, p.dx, 0, [&](int a, int e) {
        for (int s = a; s < e; ++s) {
          const Comp c = comp[s];
          if (dist2_32(c.x, c.y, x0, y0, L, p.periodic) <= t2) {
            ix2 += kIx2[c.bits & 7];
            iy += kIy[c.bits & 7];
          }
        }
      });
      if (ix2 != 0 || iy != 0) {
        const double cx = 0.5 * (double)ix2, cy = 0.8660254037844386 * (double)iy;
        const double n = sqrt(cx * cx + cy * cy), off = off_unit * (double)p.dx;
        float nx = (float)((double)x0 + cx / n * off), ny = (float)((double)y0 + cy / n * off);
        if (p.periodic) {
          nx = wrap_coord(nx, L, p.cs, p.nfine);
          ny = wrap_coord(ny, L, p.cs, p.nfine);
        }
        sx[r] = nx;
        sy[r] = ny;
        f = 1;
      }
    }
    flag[r] = make_int2(f, 0);
  }
}

// Spawned points follow the P refined ones, in the order of their parents: point P + soff[r].
__global__ __launch_bounds__(TB) void k_spawn_place(const int* __restrict__ Ptot, const int2* __restrict__ flag,
                                                    const int2* __restrict__ soff, const float* __restrict__ sx,
                                                    const float* __restrict__ sy, float* __restrict__ cpx,
                                                    float* __restrict__ cpy, int* __restrict__ cpi,
                                                    int* __restrict__ ptot) {
  const int P = *Ptot;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < P; r += gridDim.x * blockDim.x) {
    if (!flag[r].x) continue;
    const int j = P + soff[r].x;
    cpx[j] = sx[r];
    cpy[j] = sy[r];
    if (cpi) cpi[j] = cpi[r];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *ptot = P + soff[P].x;
}

// ------------------------------------------------------------------------------ host side
static float f32(double v) { return (float)v; }
// 2 (1 - sqrt(3)/2): the soft_outer spawn offset in units of dx (direbm/reference/simulator.py:25-27)
static const double kSoftOffset = 2.0 * (1.0 - std::sqrt(3.0) / 2.0);

static int make_inst(const drb_params& prm, Inst& p) {
  std::memset(&p, 0, sizeof(p));
  if (!(prm.tau > 0.f)) return fail(DRB_ERR_INVALID_ARG, "tau must be > 0");
  if (!(prm.dx > 0.f) || !(prm.alpha > 0.f)) return fail(DRB_ERR_INVALID_ARG, "dx and alpha must be > 0");
  p.tau = prm.tau;
  p.dx = prm.dx;
  p.cs = f32((double)prm.dx / (double)prm.alpha);
  p.rr = prm.r_res > 0.f ? prm.r_res : prm.dx;
  p.t_dx = sqrt_threshold(p.dx);
  p.t_rr = sqrt_threshold(p.rr);
  p.eps = prm.eps > 0.f ? prm.eps : 1e-6f;
  p.kappa_hard = prm.kappa_hard;
  p.kappa_soft = prm.kappa_soft > prm.kappa_hard ? prm.kappa_soft : 0;
  static const double W[7] = {0.5, 1.0 / 12, 1.0 / 12, 1.0 / 12, 1.0 / 12, 1.0 / 12, 1.0 / 12};
  for (int k = 0; k < 7; ++k) p.frest[k] = f32((double)prm.rho_rest * W[k]);
  p.two_level = prm.two_level ? 1 : 0;
  p.qcs = p.two_level ? 2.f * p.cs : p.cs;
  p.iq = 1.f / p.qcs;
  if (p.two_level) {
    p.cx = prm.cx;
    p.cy = prm.cy;
    p.rin2 = prm.r_in * prm.r_in;
  }
  p.periodic = prm.periodic ? 1 : 0;
  if (p.periodic) {
    p.L = prm.box;
    double n = std::round((double)prm.box / (double)p.cs);
    if (n < 1 || std::fabs(n * p.cs - prm.box) > 1e-5 * prm.box)
      return fail(DRB_ERR_INVALID_ARG, "periodic box must be a whole number of thinning cells");
    p.nfine = (int)n;
    if (p.two_level && (p.nfine & 1)) return fail(DRB_ERR_INVALID_ARG, "two-level periodic box needs an even cell count");
    p.nq = p.two_level ? p.nfine / 2 : p.nfine;
    if (!(std::max(p.rr, p.dx) < 0.5f * p.L)) return fail(DRB_ERR_INVALID_ARG, "radius must be < box / 2");
  }
  // Window margin: covers fp32 rounding of cell indices (positions up to ~1e5 query cells).
  p.margin = 1e-3f * p.qcs;
  return DRB_OK;
}

}  // namespace drb

using namespace drb;

// ================================================================================ simulator
drb_sim::~drb_sim() {
  graph_drop();
  if (gs) (void)hipStreamDestroy(gs);
  gcnt.release();
  gsv.release();
  ginst1.release();
  for (Buf* b : {&d_inst, &mx[0], &mx[1], &my[0], &my[1], &mf[0], &mf[1], &mi[0], &mi[1], &box, &disp, &status, &dcounts,
                 &moff, &diag, &diagpart, &bk, &bcells, &brp, &bx, &by, &bslot, &bsl, &bem, &bpb, &bpq, &pxy, &pxyi, &ntoff, &tcnt, &tpos, &tseg, &wstart, &bcnt, &boff, &ex, &kc, &koff, &spflag, &spoff,
                 &spx, &spy, &sptot, &bidx, &hio})
    b->release();
  w0.release();
  for (auto& w : tape) w.release();
  adj.release();
  sorter.release();
  scan.release();
  if (hpin) (void)hipHostFree(hpin);
  for (auto& set : ev)
    for (auto& e : set)
      if (e) (void)hipEventDestroy(e);
  if (ev_mid) (void)hipEventDestroy(ev_mid);
}

int drb_sim::ensure_state(int which, long long cap) {
  if (cap <= mcap[which]) return DRB_OK;
  long long c = std::max<long long>({cap + cap / 2, 2 * mcap[which], 1024});  // geometric, as Buf
  // Growing a SoA buffer: copy live planes (only the current buffer holds live data).
  Buf nx, ny, nf, ni;
  DRB_CHECK(nx.ensure((size_t)c * 4));
  DRB_CHECK(ny.ensure((size_t)c * 4));
  DRB_CHECK(nf.ensure((size_t)c * 7 * 4));
  DRB_CHECK(ni.ensure((size_t)c * 4));
  if (which == cur && M > 0 && mcap[which] > 0) {
    DRB_HIP(hipMemcpyAsync(nx.p, mx[which].p, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
    DRB_HIP(hipMemcpyAsync(ny.p, my[which].p, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
    DRB_HIP(hipMemcpyAsync(ni.p, mi[which].p, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
    DRB_HIP(hipMemcpy2DAsync(nf.p, (size_t)c * 4, mf[which].p, (size_t)mcap[which] * 4, (size_t)M * 4, 7,
                             hipMemcpyDeviceToDevice, s));
    DRB_HIP(hipStreamSynchronize(s));
  }
  mx[which].release(); my[which].release(); mf[which].release(); mi[which].release();
  mx[which] = nx; my[which] = ny; mf[which] = nf; mi[which] = ni;
  mcap[which] = c;
  return DRB_OK;
}


int drb_sim::layout(Work& w, bool reset_box, bool defer) {
  // Dense query grid of every instance from its bounding box (or its periodic box).
  long long off = 0;
  w.sub_bits = 0;
  for (int b = 0; b < ninst; ++b) {
    Inst& p = hin[b];
    if (p.two_level) w.sub_bits = 3;
    p.cell_off = off;
    if (p.periodic) {
      p.qx0 = p.qy0 = 0;
      p.nqx = p.nqy = p.nq;
    } else if (counts[b] == 0) {
      p.qx0 = p.qy0 = 0;
      p.nqx = p.nqy = 0;
    } else if (hbox()[b].qxmin > hbox()[b].qxmax) {
      return fail(DRB_ERR_NONFINITE, "instance has no finite position");
    } else {
      const BBox& bb = hbox()[b];
      p.qx0 = bb.qxmin;
      p.qy0 = bb.qymin;
      long long nx = (long long)bb.qxmax - bb.qxmin + 1, ny = (long long)bb.qymax - bb.qymin + 1;
      if (nx > (1LL << 28) || ny > (1LL << 28) || nx * ny > (1LL << 31))
        return fail(DRB_ERR_KEY_RANGE, "point cloud spans too many cells");
      p.nqx = (int)nx;
      p.nqy = (int)ny;
    }
    off += (long long)p.nqx * p.nqy;
    // Window margin: several fp32 ulps of the largest cell coordinate of the grid (a point's cell
    // index floor(x / cs) and the window bounds are both rounded), at least 1e-3 of a cell.
    const double far = std::max({std::fabs((double)p.qx0), std::fabs((double)p.qx0 + p.nqx), std::fabs((double)p.qy0),
                                 std::fabs((double)p.qy0 + p.nqy), (double)p.nq}) * (p.two_level ? 2.0 : 1.0);
    p.margin = (float)((double)p.qcs * std::max(1e-3, 16.0 * far * 0x1p-24));
    // Reach limit of the per-range image shifts (wrap_shift): L/2 less one query cell, and one
    // more against the fp32 rounding of the window bounds.
    p.wlim = p.periodic ? (float)(0.5 * (double)p.L - 2.0 * (double)p.qcs) : -1.f;
  }
  w.ncells = off;
  if (((unsigned long long)std::max<long long>(w.ncells, 1) << w.sub_bits) > (1ULL << 32))
    return fail(DRB_ERR_KEY_RANGE, "cell grid too large for 32-bit keys");
  w.hin = hin;
  DRB_CHECK(w.d_inst.ensure(sizeof(Inst) * ninst));
  if (defer && ninst == 1) return DRB_OK;  // the step's first kernel (k_head) writes the table and resets
  if (ninst <= kInstArg) {
    // the table travels as the launch's argument: no copy from pageable memory, and the boxes
    // and displacements of the step are reset by the same launch
    InstArg a;
    std::memcpy(a.v, w.hin.data(), sizeof(Inst) * ninst);
    hipLaunchKernelGGL(k_set_inst, dim3(1), dim3(64), 0, s, a, ninst, w.d_inst.as<Inst>(),
                       reset_box ? box.as<BBox>() : (BBox*)nullptr, disp.as<int>());
    DRB_HIP(hipGetLastError());
  } else {
    DRB_HIP(hipMemcpyAsync(w.d_inst.p, w.hin.data(), sizeof(Inst) * ninst, hipMemcpyHostToDevice, s));
    if (reset_box)
      hipLaunchKernelGGL(k_box_init, dim3((ninst + 63) / 64), dim3(64), 0, s, box.as<BBox>(), disp.as<int>(), ninst);
  }
  return DRB_OK;
}

int drb_sim::ensure_bins(long long ncells, long long pcap) {
  DRB_CHECK(bk.ensure((size_t)pcap * 4));
  DRB_CHECK(bx.ensure((size_t)pcap * 4));
  DRB_CHECK(by.ensure((size_t)pcap * 4));
  DRB_CHECK(bslot.ensure((size_t)pcap * 4));
  DRB_CHECK(bsl.ensure((size_t)pcap * 4));
  DRB_CHECK(bidx.ensure((size_t)pcap * 4));
  DRB_CHECK(bem.ensure((size_t)pcap));
  DRB_CHECK(bcells.ensure((size_t)(ncells + 1) * 8));
  DRB_CHECK(brp.ensure((size_t)(ncells + 1) * 8));
  DRB_CHECK(bpb.ensure((size_t)nblk(ncells) * 4));
  DRB_CHECK(bpq.ensure((size_t)nblk(ncells) * 16));
  return DRB_OK;
}

// Bin the refined control points of w by the query cell of their refined position: refine moves a
// point up to disp cells from the cell it is listed under (across a periodic seam too), so windows
// over the bins need no displacement margin, and every candidate lies in the cell it is listed
// under.  Within a bin the points are in ascending index, the order of a stable sort by bin.  The
// live count stays on the device.  keyed: the refine has written the bin keys, counts and arrival
// slots of the points (k_cpkeys does it otherwise, for every point up to Ptot, spawned ones too).
int drb_sim::bin_control_points(const Work& w, long long ncells, long long pcap, const int* Ptot, hipStream_t q,
                                bool graph, bool keyed, const Inst* I) {
  if (!I) I = w.d_inst.as<Inst>();
  // k_runs cleared the bin counts before the refine counted into them (keyed); otherwise clear here
  if (graph && !keyed)
    hipLaunchKernelGGL(k_fill_i32, dim3(std::min<unsigned>(nblk(2 * ncells), 4096)), dim3(TB), 0, q, bcells.as<int>(),
                       2 * ncells, 0);
  if (!keyed)
    hipLaunchKernelGGL(k_cpkeys, dim3(nblk(pcap)), dim3(TB), 0, q, Ptot, (int)pcap, w.cpx.as<float>(), w.cpy.as<float>(),
                       w.cpi.as<int>(), I, bk.as<uint32_t>(), bcells.as<int2>(), status.as<unsigned>(), bsl.as<int>());
  DRB_CHECK(scan.run(bcells.as<int2>(), brp.as<int2>(), ncells, q));
  hipLaunchKernelGGL(k_binplace, dim3(nblk(pcap)), dim3(TB), 0, q, Ptot, (int)pcap, bk.as<uint32_t>(), bsl.as<int>(),
                     brp.as<int2>(), bidx.as<int>());
  // open instances: the next step's boxes too (reset with this step's layout)
  const bool boxes = pmode != 1;
  hipLaunchKernelGGL(k_binsort, dim3(nblk(ncells)), dim3(TB), 0, q, ncells, brp.as<int2>(), bidx.as<int>(),
                     w.cpx.as<float>(), w.cpy.as<float>(), bx.as<float>(), by.as<float>(), bslot.as<int>(),
                     ninst > 1 ? w.cpi.as<int>() : (const int*)nullptr, I, boxes ? box.as<BBox>() : (BBox*)nullptr,
                     bpb.as<int>(), bpq.as<int4>());
  if (boxes)
    hipLaunchKernelGGL(k_box_merge, dim3(1), dim3(TB), 0, q, (int)nblk(ncells), bpb.as<int>(), bpq.as<int4>(),
                       box.as<BBox>());
  return DRB_OK;
}

// Buffers of the tile-major normaliser: the instances' first tiles, each tile's components and
// first position (with the scan's partials), the first tile of each 64-component wave.
int drb_sim::tile_buffers(long long ntiles, long long ncomp) {
  DRB_CHECK(ntoff.ensure(sizeof(int) * (ninst + 1)));
  DRB_CHECK(tcnt.ensure((size_t)(ntiles + 1) * 8));
  DRB_CHECK(tpos.ensure((size_t)(ntiles + 1) * 8));
  DRB_CHECK(tseg.ensure((size_t)(ntiles + 1) * 16));
  DRB_CHECK(wstart.ensure((size_t)(ncomp / 64 + 2) * 4));
  return scan.reserve(ntiles);
}

// The normaliser over the components renumbered tile-major (one-level grids): tile counts, their
// scan, each wave's first tile, then one wave per 64 positions.  ntiles: the tiles of the layout
// (launch size; the device's total is toff[ninst]); ncomp: the components (launch size).
int drb_sim::norm_tiles(const Work& w, const Inst* I, long long ntiles, long long ncomp, hipStream_t q) {
  hipLaunchKernelGGL(k_tile_counts, dim3(nblk(ntiles)), dim3(TB), 0, q, (const int*)ntoff.as<int>(), ninst, I,
                     (const int*)w.rp.as<int>(), (int)ntiles, tcnt.as<int2>(), tseg.as<int4>());
  DRB_CHECK(scan.run(tcnt.as<int2>(), tpos.as<int2>(), ntiles, q));
  hipLaunchKernelGGL(k_wave_start, dim3(nblk(ntiles)), dim3(TB), 0, q, (const int*)ntoff.as<int>(), ninst,
                     (const int2*)tpos.as<int2>(), wstart.as<int>());
  DRB_LAUNCH_PM(k_norm_tiles, pmode, dim3(nblk(ncomp)), dim3(TB), 0, q, (const int*)ntoff.as<int>(), ninst,
                (const int2*)tpos.as<int2>(), (const int*)wstart.as<int>(), (const int2*)tcnt.as<int2>(),
                (const int4*)tseg.as<int4>(), w.comp.as<Comp>(), w.cv.as<float>(), I, brp.as<int>(), bx.as<float>(), by.as<float>(),
                w.swc.as<float>());
  return DRB_OK;
}

int drb_sim::step_once() {
  if (M == 0) return DRB_OK;
  const bool rec = tape_pos < tape_n;
  Work& w = rec ? tape[tape_pos] : w0;
  const long long N = 7 * M;
  if (N > INT_MAX / 2) return fail(DRB_ERR_CAPACITY, "too many components");
  if (obst && rec) return fail(DRB_ERR_INVALID_ARG, "obstacles are not supported in recorded steps");
  if (spawn && rec) return fail(DRB_ERR_INVALID_ARG, "the soft_outer spawn is not supported in recorded steps");
  // A single simulation without obstacles uploads its table and resets the box in the first kernel.
  const bool head = ninst == 1 && !obst;
  DRB_CHECK(layout(w, true, head));  // also resets the boxes the emit accumulates, and disp
  const long long ncells = w.ncells;
  const int sub_bits = w.sub_bits;
  // Obstacles: the number of bounce parts (explicit components after the N implicit ones), read by
  // the host to size the component arrays; a bounced component gets the key dead = ncells << sub_bits.
  long long E = 0;
  if (obst) {
    DRB_CHECK(bcnt.ensure((size_t)(M + 1) * 8));
    DRB_CHECK(boff.ensure((size_t)(M + 1) * 8));
    hipLaunchKernelGGL(k_bounce_count, dim3(nblk(M)), dim3(TB), 0, s, mx[cur].as<float>(), my[cur].as<float>(),
                       mi[cur].as<int>(), (int)M, w.d_inst.as<Inst>(), bcnt.as<int2>());
    DRB_CHECK(scan.run(bcnt.as<int2>(), boff.as<int2>(), M, s));
    DRB_HIP(hipMemcpyAsync(hE(), boff.as<int2>() + M, sizeof(int), hipMemcpyDeviceToHost, s));
    DRB_HIP(hipStreamSynchronize(s));
    E = *hE();
    if (N + E > INT_MAX / 2) return fail(DRB_ERR_CAPACITY, "too many components");
    DRB_CHECK(ex.ensure((size_t)std::max<long long>(E, 1) * sizeof(float4)));
  }
  const long long NT = N + E;
  const uint32_t dead = (uint32_t)((unsigned long long)ncells << sub_bits);
  int bits = 1;
  const unsigned long long kmax = ((unsigned long long)ncells + (obst ? 1 : 0)) << sub_bits;
  while (bits < 32 && (1ULL << bits) < kmax) ++bits;
  if (kmax > (1ULL << 32)) return fail(DRB_ERR_KEY_RANGE, "cell grid too large for 32-bit keys");
  const long long pcap = std::max<long long>(1, std::min<long long>(NT, ncells * (sub_bits ? 5 : 1)));
  // with the spawn, at most one spawned point per refined one
  const long long pcap_s = spawn ? 2 * pcap : pcap;
  const int nxt = cur ^ 1;
  DRB_CHECK(ensure_state(nxt, pcap_s));
  DRB_CHECK(w.keys.ensure((size_t)NT * 4));
  DRB_CHECK(w.vals.ensure((size_t)NT * 4));
  if (ninst == 1) DRB_CHECK(pxy.ensure((size_t)M * 8));
  // several instances without obstacles: the head kernel writes {x, y, instance} per moment
  const bool head_rec = !obst && ninst > 1;
  if (head_rec) DRB_CHECK(pxyi.ensure((size_t)M * 16));
  // one-level grids of 2^20 components or more: the normaliser runs over the components renumbered
  // tile-major (k_head writes the instances' first tiles); its four extra launches (~16 us) would
  // cost a small step more than the shorter walks save
  const bool tiled = sub_bits == 0 && !obst && NT >= kTileMin;
  long long ntiles = 0;
  if (tiled) {
    for (const Inst& q : w.hin) ntiles += grid_tiles(q);
    if (ntiles > INT_MAX / 2) return fail(DRB_ERR_CAPACITY, "too many cells");
    DRB_CHECK(tile_buffers(ntiles, NT));
  }
  DRB_CHECK(w.comp.ensure((size_t)NT * sizeof(Comp)));
  DRB_CHECK(w.swc.ensure((size_t)NT * 4));
  DRB_CHECK(w.cv.ensure((size_t)NT * 4));
  DRB_CHECK(w.cells.ensure((size_t)(ncells + 1) * 8));
  DRB_CHECK(w.rp.ensure((size_t)(ncells + 1) * 8));
  DRB_CHECK(ensure_bins(ncells, pcap_s));
  if (pcap_s > w.cpcap) {
    long long c = std::max(pcap_s + pcap_s / 2, 2 * w.cpcap);  // geometric, as Buf
    for (Buf* b : {&w.cp0x, &w.cp0y, &w.cpq, &w.cpi, &w.cpx, &w.cpy}) DRB_CHECK(b->ensure((size_t)c * 4));
    DRB_CHECK(w.cpf.ensure((size_t)c * 7 * 4));
    DRB_CHECK(w.emask.ensure((size_t)c));
    DRB_CHECK(w.kap.ensure((size_t)c));
    if (rec) {
      DRB_CHECK(w.cpn.ensure((size_t)c * 4));
      DRB_CHECK(w.cps.ensure((size_t)c * 4));
      DRB_CHECK(w.cpw.ensure((size_t)c * 4));
      DRB_CHECK(w.cpg.ensure((size_t)c * 7 * 4));
    }
    w.cpcap = c;
  }
  const long long pitch = mcap[cur];
  if (rec) {
    // input state of the step: positions, pre-collision f (SoA, pitch M), instance
    DRB_CHECK(w.sx.ensure((size_t)M * 4));
    DRB_CHECK(w.sy.ensure((size_t)M * 4));
    DRB_CHECK(w.si.ensure((size_t)M * 4));
    DRB_CHECK(w.sf.ensure((size_t)M * 7 * 4));
    DRB_CHECK(w.crun.ensure((size_t)N * 4));
    DRB_HIP(hipMemcpyAsync(w.sx.p, mx[cur].p, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
    DRB_HIP(hipMemcpyAsync(w.sy.p, my[cur].p, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
    DRB_HIP(hipMemcpyAsync(w.si.p, mi[cur].p, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
    DRB_HIP(hipMemcpy2DAsync(w.sf.p, (size_t)M * 4, mf[cur].p, (size_t)pitch * 4, (size_t)M * 4, 7,
                             hipMemcpyDeviceToDevice, s));
  }
  w.N = NT;  // components sorted (with bounce parts); the live ones end at rp[ncells].x
  w.M = M;
  w.taped = rec;
  w.counts_in = counts;
  const Inst* I = w.d_inst.as<Inst>();
  const int m = (int)M;
  // launch size of the control-point kernels: the count changes slowly from step to step, so
  // the current moment count (+1/8) covers it; the kernels loop over any excess on the device
  const long long pguess = std::min<long long>(pcap, M + M / 8 + 1024);
  const long long cpcap = w.cpcap;
  mark(0);
  if (!obst) {
    // collide and keys in one kernel: the "keys" stage is empty
    if (head)
      hipLaunchKernelGGL(k_head<true>, dim3(nblk(m)), dim3(TB), 0, s, w.hin[0], w.d_inst.as<Inst>(), (const Inst*)nullptr,
                         (const int*)nullptr, box.as<BBox>(), disp.as<int>(), 0, mf[cur].as<float>(), pitch,
                         (float*)nullptr, 0LL, mx[cur].as<float>(), my[cur].as<float>(), m, (const int*)nullptr,
                         sub_bits, w.keys.as<uint32_t>(), status.as<unsigned>(), w.cells.as<int>(), 2 * ncells,
                         pxy.as<float2>(), (float4*)nullptr, tiled ? ntoff.as<int>() : (int*)nullptr, 1);
    else  // several instances: layout() uploaded the table and reset the boxes
      hipLaunchKernelGGL(k_head<false>, dim3(nblk(m)), dim3(TB), 0, s, w.hin[0], (Inst*)nullptr, I,
                         (const int*)mi[cur].as<int>(), box.as<BBox>(), disp.as<int>(), 0, mf[cur].as<float>(), pitch,
                         (float*)nullptr, 0LL, mx[cur].as<float>(), my[cur].as<float>(), m, (const int*)nullptr,
                         sub_bits, w.keys.as<uint32_t>(), status.as<unsigned>(), w.cells.as<int>(), 2 * ncells,
                         (float2*)nullptr, pxyi.as<float4>(), tiled ? ntoff.as<int>() : (int*)nullptr, ninst);
    mark(1);
  } else {
    hipLaunchKernelGGL(k_collide_box, dim3(nblk(m)), dim3(TB), 0, s, mf[cur].as<float>(), pitch, mx[cur].as<float>(),
                       my[cur].as<float>(), mi[cur].as<int>(), m, (const int*)nullptr, I, box.as<BBox>(),
                       status.as<unsigned>(), 1, 0, (float*)nullptr, 0LL);
    mark(1);
    hipLaunchKernelGGL(k_keys, dim3(nblk(m)), dim3(TB), 0, s, mx[cur].as<float>(), my[cur].as<float>(),
                       mi[cur].as<int>(), m, (const int*)nullptr, I, sub_bits, w.keys.as<uint32_t>(),
                       status.as<unsigned>(), dead, w.cells.as<int>(), 2 * ncells,
                       ninst == 1 ? pxy.as<float2>() : (float2*)nullptr);
  }
  if (E > 0)
    hipLaunchKernelGGL(k_bounce_emit, dim3(nblk(m)), dim3(TB), 0, s, mx[cur].as<float>(), my[cur].as<float>(),
                       mi[cur].as<int>(), mf[cur].as<float>(), pitch, m, I, sub_bits, boff.as<int2>(), (int)N,
                       w.keys.as<uint32_t>(), ex.as<float4>(), status.as<unsigned>());
  mark(2);
  DRB_CHECK(sorter.sort_swap(w.keys, w.vals, (int)NT, bits, true, s));
  mark(3);
  // cells (+1: bounced components, which sort last, count into the slot past the grid); k_keys
  // cleared the counts
  hipLaunchKernelGGL(k_mark<kMarkChunks>, dim3(nblk((NT + kMarkChunks - 1) / kMarkChunks)), dim3(TB), 0, s, w.keys.as<uint32_t>(), w.vals.as<uint32_t>(), (int)NT,
                     (const int*)nullptr, sub_bits, mx[cur].as<float>(), my[cur].as<float>(), mf[cur].as<float>(), pitch,
                     ninst == 1 ? (const int*)nullptr : mi[cur].as<int>(), (const float*)nullptr, (const float*)nullptr, (const int*)nullptr, I,
                     w.comp.as<Comp>(), w.cv.as<float>(), w.cells.as<int>(), 1, (int)N,
                     obst ? ex.as<float4>() : (const float4*)nullptr,
                     ninst == 1 ? pxy.as<float2>() : (const float2*)nullptr,
                     head_rec ? (const float4*)pxyi.as<float4>() : (const float4*)nullptr);
  DRB_CHECK(scan.run(w.cells.as<int2>(), w.rp.as<int2>(), ncells, s));
  const int* Ptot = w.rp.as<int>() + 2 * ncells + 1;
  const int* Nlive = obst ? w.rp.as<int>() + 2 * ncells : (const int*)nullptr;  // components in cells
  // k_runs also writes the instances' counts (with obstacles or the spawn they come later)
  hipLaunchKernelGGL(k_runs, dim3(nblk(NT)), dim3(TB), 0, s, w.rp.as<int2>(), ncells, (int)NT, (const int*)nullptr,
                     sub_bits, w.keys.as<uint32_t>(), w.comp.as<Comp>(), w.cp0x.as<float>(), w.cp0y.as<float>(),
                     w.cpq.as<int>(), w.cpi.as<int>(), rec ? w.crun.as<int>() : (int*)nullptr,
                     rec ? w.cpn.as<int>() : (int*)nullptr, rec ? w.cps.as<int>() : (int*)nullptr, bcells.as<int2>(),
                     ncells, I, ninst, (!obst && !spawn) ? dcounts.as<int>() : (int*)nullptr);
  mark(4);
  DRB_LAUNCH_PMGK(k_refine, pmode, g_refine, d_refine, dim3(xblk(gblk(pguess, g_refine))), dim3(TB), 0, s, Ptot, (int)pcap, w.cp0x.as<float>(),
                w.cp0y.as<float>(), w.cpi.as<int>(), w.cpq.as<int>(), I, w.rp.as<int>(), w.comp.as<Comp>(),
                w.cpx.as<float>(), w.cpy.as<float>(), w.kap.as<unsigned char>(), disp.as<int>(),
                rec ? w.cpw.as<float>() : (float*)nullptr, status.as<unsigned>(),
                spawn ? (uint32_t*)nullptr : bk.as<uint32_t>(), spawn ? (int2*)nullptr : bcells.as<int2>(),
                spawn ? (int*)nullptr : bsl.as<int>());
  mark(5);
  // Ptot_e: the points of the resample -- the refined ones, and with the spawn the spawned ones
  // after them (P + S on the device)
  const int* Ptot_e = Ptot;
  if (spawn) {
    DRB_CHECK(spflag.ensure((size_t)(pcap + 1) * 8));
    DRB_CHECK(spoff.ensure((size_t)(pcap + 1) * 8));
    DRB_CHECK(spx.ensure((size_t)pcap * 4));
    DRB_CHECK(spy.ensure((size_t)pcap * 4));
    DRB_CHECK(sptot.ensure(sizeof(int)));
    hipLaunchKernelGGL(k_fill_i32, dim3(std::min<unsigned>(nblk(2 * (pcap + 1)), 4096)), dim3(TB), 0, s,
                       spflag.as<int>(), 2 * (pcap + 1), 0);
    hipLaunchKernelGGL(k_spawn_flag, dim3(nblk(pguess)), dim3(TB), 0, s, Ptot, w.cp0x.as<float>(), w.cp0y.as<float>(),
                       w.cpi.as<int>(), I, w.rp.as<int>(), w.comp.as<Comp>(), w.kap.as<unsigned char>(), kSoftOffset,
                       spflag.as<int2>(), spx.as<float>(), spy.as<float>());
    DRB_CHECK(scan.run(spflag.as<int2>(), spoff.as<int2>(), pcap, s));
    hipLaunchKernelGGL(k_spawn_place, dim3(nblk(pguess)), dim3(TB), 0, s, Ptot, spflag.as<int2>(), spoff.as<int2>(),
                       spx.as<float>(), spy.as<float>(), w.cpx.as<float>(), w.cpy.as<float>(), w.cpi.as<int>(),
                       sptot.as<int>());
    hipLaunchKernelGGL(k_counts, dim3((ninst + 63) / 64), dim3(64), 0, s, w.rp.as<int2>(), I, ninst, dcounts.as<int>(),
                       (const int*)sptot.as<int>());
    Ptot_e = sptot.as<int>();
  }
  DRB_CHECK(bin_control_points(w, ncells, pcap_s, Ptot_e, s, false, !spawn));
  // The next step's boxes, counts and this step's status are final here (with obstacles the
  // counts only after the emit's drop): read them back now, the rest of the step runs meanwhile.
  if (!obst) DRB_CHECK(read_back());
  mark(6);
  if (tiled)
    DRB_CHECK(norm_tiles(w, I, ntiles, NT, s));
  else
    DRB_LAUNCH_PM(k_norm, pmode, dim3(nblk(NT)), dim3(TB), 0, s, (int)NT, Nlive, w.comp.as<Comp>(), w.cv.as<float>(), I,
                  brp.as<int>(), bx.as<float>(), by.as<float>(), (const int*)nullptr, w.swc.as<float>());
  mark(7);
  DRB_LAUNCH_PMGK(k_gather, pmode, g_gather, d_gather, dim3(xblk(gblk(pguess, g_gather))), dim3(TB), 0, s, Ptot_e, (int)pcap, w.cpx.as<float>(),
                w.cpy.as<float>(), w.cpi.as<int>(), I, w.rp.as<int>(), w.comp.as<Comp>(),
                w.cpf.as<float>(), cpcap, w.emask.as<unsigned char>(), bslot.as<int>(), bem.as<unsigned char>());
  if (obst) {
    // control points inside an obstacle are dropped after resampling (they took part in it)
    DRB_CHECK(kc.ensure((size_t)(pcap + 1) * 8));
    DRB_CHECK(koff.ensure((size_t)(pcap + 1) * 8));
    hipLaunchKernelGGL(k_keep, dim3(nblk(pcap)), dim3(TB), 0, s, Ptot, w.cpx.as<float>(), w.cpy.as<float>(),
                       w.cpi.as<int>(), pcap, I, kc.as<int2>());
    DRB_CHECK(scan.run(kc.as<int2>(), koff.as<int2>(), pcap, s));
  }
  mark(8);
  // The gap-fill walks the points' own bins.
  DRB_LAUNCH_PM(k_gapfill_emit, pmode, dim3(emit_blocks(pcap_s)), dim3(TB), 0, s, Ptot_e, (int)pcap, w.cpx.as<float>(),
                     w.cpy.as<float>(), w.cpi.as<int>(), I, brp.as<int>(), w.cpf.as<float>(), cpcap,
                     w.emask.as<unsigned char>(), (const uint32_t*)nullptr, mx[nxt].as<float>(), my[nxt].as<float>(),
                     1LL, mf[nxt].as<float>(), 1LL, mcap[nxt], mi[nxt].as<int>(), status.as<unsigned>(), rec ? w.cpg.as<float>() : (float*)nullptr,
                     obst ? koff.as<int2>() : (const int2*)nullptr, bem.as<unsigned char>(), bx.as<float>(),
                     by.as<float>());
  if (obst)
    hipLaunchKernelGGL(k_counts_keep, dim3((ninst + 63) / 64), dim3(64), 0, s, w.rp.as<int2>(), koff.as<int2>(), I,
                       ninst, dcounts.as<int>());
  mark(9);
  if (rec) {
    DRB_CHECK(w.dsp.ensure(sizeof(int) * ninst));
    DRB_HIP(hipMemcpyAsync(w.dsp.p, disp.p, sizeof(int) * ninst, hipMemcpyDeviceToDevice, s));
    // moment offsets of the instances entering the step (the adjoint's per-instance d tau)
    DRB_CHECK(w.dmoff.ensure(sizeof(long long) * (ninst + 1)));
    w.hmoff.assign(ninst + 1, 0);
    for (int b = 0; b < ninst; ++b) w.hmoff[b + 1] = w.hmoff[b] + w.counts_in[b];
    DRB_HIP(hipMemcpyAsync(w.dmoff.p, w.hmoff.data(), sizeof(long long) * (ninst + 1), hipMemcpyHostToDevice, s));
  }
  if (obst) DRB_CHECK(read_back());
  DRB_HIP(hipGetLastError());
  pend_w = &w;
  pend_nxt = nxt;
  pend_rec = rec;
  updates += M;
  return DRB_OK;
}

// Boxes, counts and status to the pinned mirror, then the read-back event.
int drb_sim::read_back() {
  DRB_HIP(hipMemcpyAsync(hpin, hio.p, (size_t)ninst * (sizeof(BBox) + sizeof(int)) + sizeof(unsigned),
                         hipMemcpyDeviceToHost, s));
  DRB_HIP(hipEventRecord(ev_mid, s));
  return DRB_OK;
}

int drb_sim::stage_read(unsigned par, int i0, int i1) {
  for (int i = i0; i < i1; ++i)
    if (timing & (1u << i)) {
      float ms = 0.f;
      DRB_HIP(hipEventSynchronize(ev[par][i + 1]));  // event 6 follows the read back closely
      DRB_HIP(hipEventElapsedTime(&ms, ev[par][i], ev[par][i + 1]));
      stage_sum[i] += ms;
    }
  return DRB_OK;
}

// Wait for the queued step's read back; take its counts as the moments of the next step.  The
// step's later kernels may still run.  Its stages before the read back (and the previous step's
// later ones, queued before it) are complete: their event times are summed.
int drb_sim::step_finish() {
  if (!pend_w) return DRB_OK;  // no step queued (no moments)
  Work& w = *pend_w;
  pend_w = nullptr;
  DRB_HIP(hipEventSynchronize(ev_mid));
  if (timing) {
    if (back_unread) DRB_CHECK(stage_read(tpar ^ 1, 6, kStages));
    DRB_CHECK(stage_read(tpar, 0, obst ? kStages : 6));
    back_unread = !obst;
    ++stage_n;
  }
  tpar ^= 1;
  if (*hstatus() & ST_NONFINITE) {
    (void)hipStreamSynchronize(s);
    return fail(DRB_ERR_NONFINITE, "non-finite value in the state");
  }
  cur = pend_nxt;
  M = 0;
  for (int b = 0; b < ninst; ++b) {
    counts[b] = hcounts()[b];
    M += counts[b];
  }
  w.P = M;
  w.counts_out = counts;
  last = &w;
  if (pend_rec) ++tape_pos;
  return DRB_OK;
}

int drb_sim::step_sync() {
  DRB_CHECK(step_once());
  DRB_CHECK(step_finish());
  DRB_HIP(hipStreamSynchronize(s));
  return DRB_OK;
}

void drb_sim::graph_drop() {
  for (auto& g : gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  gsig.clear();
}

// Graph mode: capacities fixed at capture, live counts on the device.
//  - Periodic instances have a fixed grid: every thinning cell holds at most one control point
//    (five in two-level mode), so P <= pcap and, after one step, M <= pcap.
//  - Open instances' grids follow their bounding boxes.  The capacity of gncells cells covers the
//    current grids grown for `horizon` steps by the most a step can grow them: a moment moves at
//    most dx in a step (components lie within dx of their moment, and the thinning centroids and
//    refined points are weighted means of components), so its query cell moves at most
//    ceil(dx / qcs) + 1 cells per axis and the box around the points grows by as much per side;
//    one cell more covers the rounding of cell indices.  Each step lays itself out on the device
//    (graph_layout); one that would still exceed the capacity runs empty and step_graph resumes
//    from it with a larger one.  pcap follows from the cells as for periodic instances.
int drb_sim::graph_ready(int horizon) {
  if (obst) return fail(DRB_ERR_INVALID_ARG, "graph mode does not support obstacles");
  if (spawn) return fail(DRB_ERR_INVALID_ARG, "graph mode does not support the soft_outer spawn");
  Work& w = w0;
  DRB_CHECK(layout(w, false));
  const bool open = pmode != 1;
  long long ncells = w.ncells;
  if (open) {
    long long need = 0;
    for (int b = 0; b < ninst; ++b) {
      const Inst& p = w.hin[b];
      if (p.periodic) {
        need += (long long)p.nq * p.nq;
      } else if (counts[b] > 0) {  // an instance without moments gets none
        const long long g =
            2LL * std::max(horizon - 1, 0) * ((long long)std::ceil((double)p.dx / (double)p.qcs) + 2);
        need += ((long long)p.nqx + g) * ((long long)p.nqy + g);
      }
    }
    if (need > gncap) gncap = std::max(need, gncap + gncap / 2);
    ncells = std::max<long long>(gncap, 1);
  }
  if (((unsigned long long)ncells << w.sub_bits) > (1ULL << 32))
    return fail(DRB_ERR_KEY_RANGE, "graph mode: cell grid too large for 32-bit keys");
  // Control points: at most one per thinning cell (five sub-cells of a two-level query cell); with
  // open instances the current moments and a sixteenth more, grown by at least an eighth, instead:
  // the launches of the step are sized by it, and a block past the live count still costs its
  // dispatch (a step beyond the capacity runs empty and the call resumes, k_graph_counts)
  long long pcap = std::max<long long>(1, ncells * (w.sub_bits ? 5 : 1));
  if (open) {
    const long long need = std::min(pcap, M + M / 16 + 256);
    if (need > gpcap_open) gpcap_open = std::max(need, gpcap_open + gpcap_open / 8);
    pcap = std::min(pcap, gpcap_open);
  }
  const long long N = 7 * pcap;
  if (N > INT_MAX / 2) return fail(DRB_ERR_CAPACITY, "graph mode: too many cells");
  DRB_CHECK(ensure_state(0, pcap));
  DRB_CHECK(ensure_state(1, pcap));
  DRB_CHECK(w.keys.ensure((size_t)N * 4));
  DRB_CHECK(w.vals.ensure((size_t)N * 4));
  DRB_CHECK(w.comp.ensure((size_t)N * sizeof(Comp)));
  DRB_CHECK(w.swc.ensure((size_t)N * 4));
  DRB_CHECK(w.cv.ensure((size_t)N * 4));
  DRB_CHECK(pxyi.ensure((size_t)pcap * 16));
  DRB_CHECK(w.cells.ensure((size_t)(ncells + 1) * 8));
  DRB_CHECK(w.rp.ensure((size_t)(ncells + 1) * 8));
  if (pcap > w.cpcap) {
    for (Buf* b : {&w.cp0x, &w.cp0y, &w.cpq, &w.cpi, &w.cpx, &w.cpy}) DRB_CHECK(b->ensure((size_t)pcap * 4));
    DRB_CHECK(w.cpf.ensure((size_t)pcap * 7 * 4));
    DRB_CHECK(w.emask.ensure((size_t)pcap));
    DRB_CHECK(w.kap.ensure((size_t)pcap));
    w.cpcap = pcap;
  }
  DRB_CHECK(ensure_bins(ncells, pcap));
  DRB_CHECK(sorter.reserve((int)N));
  DRB_CHECK(scan.reserve(ncells));
  DRB_CHECK(gcnt.ensure(16));
  DRB_CHECK(gsv.ensure(sizeof(int) * (4 + ninst)));
  // periodic one-level grids: the normaliser's tiles (a fixed count)
  gtiles = 0;
  if (!open && w.sub_bits == 0 && N >= kTileMin) {
    for (const Inst& q : w.hin) gtiles += grid_tiles(q);
    if (gtiles > INT_MAX / 2) return fail(DRB_ERR_CAPACITY, "graph mode: too many cells");
    DRB_CHECK(tile_buffers(gtiles, N));
  }
  if (open) {
    // the second layout table starts as a copy of the first (the parameters; graph_layout fills in the grids)
    DRB_CHECK(ginst1.ensure(sizeof(Inst) * ninst));
    DRB_HIP(hipMemcpyAsync(ginst1.p, w.d_inst.p, sizeof(Inst) * ninst, hipMemcpyDeviceToDevice, s));
  }
  if (!gs) DRB_HIP(hipStreamCreateWithFlags(&gs, hipStreamNonBlocking));
  // Recapture when any buffer moved or a size changed.
  std::vector<const void*> sig = {mx[0].p, my[0].p, mf[0].p, mi[0].p, mx[1].p, my[1].p, mf[1].p, mi[1].p,
                                  w.keys.p, w.vals.p, w.comp.p, w.swc.p, w.cv.p, w.cells.p, w.rp.p,
                                  w.cp0x.p, w.cp0y.p, w.cpq.p, w.cpi.p, w.cpx.p, w.cpy.p, w.cpf.p, w.emask.p,
                                  w.kap.p, w.d_inst.p, sorter.k1.p, sorter.v1.p, sorter.hist.p, sorter.tot.p,
                                  scan.part.p, gcnt.p, gsv.p, ginst1.p, pxyi.p, ntoff.p, tcnt.p, tpos.p, tseg.p, wstart.p,
                                  (const void*)(intptr_t)gtiles, box.p, disp.p, status.p, dcounts.p, bk.p, bcells.p, brp.p,
                                  bx.p, by.p, bslot.p, bem.p, bsl.p, bidx.p, bpb.p, bpq.p,
                                  (const void*)(intptr_t)mcap[0], (const void*)(intptr_t)mcap[1],
                                  (const void*)(intptr_t)w.cpcap, (const void*)(intptr_t)ncells,
                                  (const void*)(intptr_t)w.sub_bits, (const void*)(intptr_t)g_refine,
                                  (const void*)(intptr_t)g_gather, (const void*)(intptr_t)d_refine,
                                  (const void*)(intptr_t)d_gather, (const void*)(intptr_t)pcap};
  gpcap = pcap;
  gncells = ncells;
  if (sig == gsig && gexec[0] && gexec[1]) return DRB_OK;
  graph_drop();
  DRB_HIP(hipStreamSynchronize(s));
  for (int from = 0; from < 2; ++from) {
    hipGraph_t g = nullptr;
    DRB_HIP(hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal));
    const int st = graph_body(from);
    const hipError_t ec = hipStreamEndCapture(gs, &g);
    if (st != DRB_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (ec != hipSuccess) return fail(DRB_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&gexec[from], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(DRB_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  gsig = sig;
  ++gcaptures;
  return DRB_OK;
}

// One step from state buffer `from` to the other, enqueued on gs with capacity-sized launches and
// the live counts read from gcnt on the device.  No allocation, synchronisation or memset.
int drb_sim::graph_body(int from) {
  Work& w = w0;
  const bool open = pmode != 1;
  const int nxt = from ^ 1;
  const long long ncells = gncells, pcap = gpcap, N = 7 * pcap, pg = pcap;
  const int sub_bits = w.sub_bits;
  int bits = 1;
  while (bits < 32 && (1ULL << bits) < ((unsigned long long)ncells << sub_bits)) ++bits;
  const Inst* I = graph_table(from);
  const long long pitch = mcap[from];
  const int* mdev = gcnt.as<int>();
  const int* ndev = gcnt.as<int>() + 1;
  hipStream_t q = gs;
  // The collided f goes to the other buffer's f planes, which the emit overwrites only after k_mark
  // has read them: the state entering the step stays as it was, for a replay resumed from it.
  // Periodic grids reset the boxes here; open ones were laid out (and reset) by the kernel ending
  // the step before (or the replay's start).
  float* fc = mf[nxt].as<float>();
  const long long cpitch = mcap[nxt];
  hipLaunchKernelGGL(k_head<false>, dim3(nblk(pcap)), dim3(TB), 0, q, w.hin[0], (Inst*)nullptr, I,
                     (const int*)mi[from].as<int>(), box.as<BBox>(), disp.as<int>(), open ? 0 : ninst,
                     mf[from].as<float>(), pitch, fc, cpitch, mx[from].as<float>(), my[from].as<float>(), (int)pcap,
                     mdev, sub_bits, w.keys.as<uint32_t>(), status.as<unsigned>(), w.cells.as<int>(), 2 * ncells,
                     (float2*)nullptr, pxyi.as<float4>(), gtiles ? ntoff.as<int>() : (int*)nullptr, ninst);
  // the sorted pairs stay where the sort's last pass wrote them
  uint32_t *sk = nullptr, *sv = nullptr;
  DRB_CHECK(sorter.sort_graph(w.keys.as<uint32_t>(), w.vals.as<uint32_t>(), (int)N, bits, q, ndev, &sk, &sv));
  hipLaunchKernelGGL(k_mark<kMarkChunks>, dim3(nblk((N + kMarkChunks - 1) / kMarkChunks)), dim3(TB), 0, q, sk, sv, (int)N, ndev, sub_bits, mx[from].as<float>(),
                     my[from].as<float>(), fc, cpitch, mi[from].as<int>(), (const float*)nullptr,
                     (const float*)nullptr, (const int*)nullptr, I, w.comp.as<Comp>(), w.cv.as<float>(),
                     w.cells.as<int>(), 1, INT_MAX, (const float4*)nullptr, (const float2*)nullptr,
                     (const float4*)pxyi.as<float4>());
  DRB_CHECK(scan.run(w.cells.as<int2>(), w.rp.as<int2>(), ncells, q));
  const int* Ptot = w.rp.as<int>() + 2 * ncells + 1;
  if (open)
    hipLaunchKernelGGL(k_graph_counts, dim3(1), dim3(TB), 0, q, w.rp.as<int2>(), I, ninst, dcounts.as<int>(),
                       w.rp.as<int>() + 2 * ncells + 1, (int)pcap, gcnt.as<int>(), gsv.as<int>());
  // periodic grids: k_runs writes the instances' counts
  hipLaunchKernelGGL(k_runs, dim3(nblk(N)), dim3(TB), 0, q, w.rp.as<int2>(), ncells, (int)N, ndev, sub_bits, sk,
                     w.comp.as<Comp>(), w.cp0x.as<float>(), w.cp0y.as<float>(), w.cpq.as<int>(), w.cpi.as<int>(),
                     (int*)nullptr, (int*)nullptr, (int*)nullptr, bcells.as<int2>(), ncells, I, ninst,
                     open ? (int*)nullptr : dcounts.as<int>());
  DRB_LAUNCH_PMGK(k_refine, pmode, g_refine, d_refine, dim3(xblk(gblk(pg, g_refine))), dim3(TB), 0, q, Ptot, (int)pcap,
                  w.cp0x.as<float>(), w.cp0y.as<float>(), w.cpi.as<int>(), w.cpq.as<int>(), I, w.rp.as<int>(),
                  w.comp.as<Comp>(), w.cpx.as<float>(), w.cpy.as<float>(), w.kap.as<unsigned char>(), disp.as<int>(),
                  (float*)nullptr, status.as<unsigned>(), bk.as<uint32_t>(), bcells.as<int2>(), bsl.as<int>());
  DRB_CHECK(bin_control_points(w, ncells, pcap, Ptot, q, true, true, I));
  if (gtiles)  // periodic one-level grids: a fixed set of tiles
    DRB_CHECK(norm_tiles(w, I, gtiles, N, q));
  else
    DRB_LAUNCH_PM(k_norm, pmode, dim3(nblk(N)), dim3(TB), 0, q, (int)N, ndev, w.comp.as<Comp>(), w.cv.as<float>(), I,
                  brp.as<int>(), bx.as<float>(), by.as<float>(), (const int*)nullptr, w.swc.as<float>());
  DRB_LAUNCH_PMGK(k_gather, pmode, g_gather, d_gather, dim3(xblk(gblk(pg, g_gather))), dim3(TB), 0, q, Ptot, (int)pcap,
                  w.cpx.as<float>(), w.cpy.as<float>(), w.cpi.as<int>(), I, w.rp.as<int>(), w.comp.as<Comp>(),
                  w.cpf.as<float>(), w.cpcap, w.emask.as<unsigned char>(), bslot.as<int>(), bem.as<unsigned char>());
  DRB_LAUNCH_PM(k_gapfill_emit, pmode, dim3(emit_blocks(pcap)), dim3(TB), 0, q, Ptot, (int)pcap, w.cpx.as<float>(),
                w.cpy.as<float>(), w.cpi.as<int>(), I, brp.as<int>(), w.cpf.as<float>(), w.cpcap,
                w.emask.as<unsigned char>(), (const uint32_t*)nullptr, mx[nxt].as<float>(), my[nxt].as<float>(), 1LL,
                mf[nxt].as<float>(), 1LL, mcap[nxt], mi[nxt].as<int>(), status.as<unsigned>(),
                (float*)nullptr, (const int2*)nullptr, bem.as<unsigned char>(), bx.as<float>(), by.as<float>());
  hipLaunchKernelGGL(k_graph_next, dim3(1), dim3(64), 0, q, Ptot, gcnt.as<int>(), graph_upd(), graph_layout_args(nxt));
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

// Open instances: the steps from state buffer 0 run with the table w0.d_inst, those from buffer 1
// with ginst1, so laying out a step never overwrites the table of the step before (the last step of
// a replay lays out one that does not run).
Inst* drb_sim::graph_table(int parity) { return parity == 0 || pmode == 1 ? w0.d_inst.as<Inst>() : ginst1.as<Inst>(); }

GraphLayout drb_sim::graph_layout_args(int parity) {
  if (pmode == 1) return GraphLayout{nullptr, 0, nullptr, nullptr, nullptr, 0, nullptr, nullptr};
  return GraphLayout{graph_table(parity), ninst, box.as<BBox>(), disp.as<int>(), (const int*)dcounts.as<int>(), gncells,
                     gsv.as<int>(), status.as<unsigned>()};
}

int drb_sim::step_graph(int nsteps) {
  const bool open = pmode != 1;
  while (nsteps > 0) {
    DRB_CHECK(graph_ready(std::min(nsteps, ghorizon)));
    // A loaded state may hold more moments than the capacity: plain steps until it fits.
    if (M > gpcap) {
      DRB_CHECK(step_sync());
      --nsteps;
      continue;
    }
    if (M == 0) return DRB_OK;
    DRB_HIP(hipStreamSynchronize(s));
    if (open) {  // the moments of each instance, which the first step's layout reads
      for (int b = 0; b < ninst; ++b) hcounts()[b] = (int)counts[b];
      DRB_HIP(hipMemcpyAsync(dcounts.p, hcounts(), (size_t)ninst * sizeof(int), hipMemcpyHostToDevice, gs));
    }
    DRB_HIP(hipMemsetAsync(status.p, 0, sizeof(unsigned), gs));
    hipLaunchKernelGGL(k_graph_set, dim3(1), dim3(64), 0, gs, gcnt.as<int>(), (int)M, graph_upd(),
                       graph_layout_args(cur));
    const int cur0 = cur;
    for (int i = 0; i < nsteps; ++i) {
      DRB_HIP(hipGraphLaunch(gexec[cur], gs));
      cur ^= 1;
    }
    DRB_HIP(hipMemcpyAsync(hcounts(), dcounts.p, (size_t)ninst * sizeof(int), hipMemcpyDeviceToHost, gs));
    DRB_HIP(hipMemcpyAsync(hstatus(), status.p, sizeof(unsigned), hipMemcpyDeviceToHost, gs));
    DRB_HIP(hipMemcpyAsync(hm(), gcnt.p, sizeof(int), hipMemcpyDeviceToHost, gs));
    unsigned long long upd = 0;
    DRB_HIP(hipMemcpyAsync(&upd, graph_upd(), sizeof(upd), hipMemcpyDeviceToHost, gs));
    std::vector<int> sv(4 + ninst, 0);
    if (open) DRB_HIP(hipMemcpyAsync(sv.data(), gsv.p, sizeof(int) * (4 + ninst), hipMemcpyDeviceToHost, gs));
    DRB_HIP(hipStreamSynchronize(gs));
    updates += (long long)upd;
    w0.ncells = gncells;  // the cell arrays the replays left (diagnostics)
    if (*hstatus() & ST_NONFINITE) return fail(DRB_ERR_NONFINITE, "non-finite value in the state");
    // (the last step's kernel lays out one more step, which does not run: that one may find its
    // grid beyond the capacity, which matters only to the next call)
    if (open && sv[0] && sv[1] < nsteps) {
      // Step sv[1] of the replay found its grid, or its control points, beyond the capacity and ran
      // empty, as did the rest: the state entering it is in that step's input buffer.  Resume from
      // it with the capacity grown (the cells from the boxes, in graph_ready).
      const int done = sv[1];
      if (sv[3] > 0) {
        gpcap_open = std::max(gpcap_open + gpcap_open / 8, (long long)sv[3] + sv[3] / 16 + 256);
      } else if (done == 0) {
        // the first step's grid is the one the capacity was sized from: it fits unless layout() refuses it
        return fail(DRB_ERR_KEY_RANGE, "graph mode: point cloud spans too many cells");
      }
      cur = cur0 ^ (done & 1);
      M = sv[2];
      for (int b = 0; b < ninst; ++b) counts[b] = sv[4 + b];
      ++gresumes;
      nsteps -= done;
      last = nullptr;  // the empty steps overwrote the step buffers
      // the boxes of that state (laying out the step reset them)
      DRB_CHECK(rebox());
      if (*hstatus() & ST_NONFINITE) return fail(DRB_ERR_NONFINITE, "non-finite value in the state");
      continue;
    }
    M = 0;
    for (int b = 0; b < ninst; ++b) {
      counts[b] = hcounts()[b];
      M += counts[b];
    }
    // the moment count of the step after the last, which the layout zeroes when it does not fit
    const long long mdev = (open && sv[0]) ? (long long)sv[2] : (long long)*hm();
    if (M != mdev) return fail(DRB_ERR_STATE, "graph mode: moment count mismatch");
    w0.P = M;
    w0.counts_out = counts;
    last = &w0;
    if (open) {
      // the table the last step ran with, where diagnostics read it
      if ((cur ^ 1) == 1)
        DRB_HIP(hipMemcpyAsync(w0.d_inst.p, ginst1.p, sizeof(Inst) * ninst, hipMemcpyDeviceToDevice, s));
      // the boxes of the state reached (the last step's kernel laid out a step that did not run,
      // resetting them), for the next call or plain step
      DRB_CHECK(rebox());
    }
    return DRB_OK;
  }
  return DRB_OK;
}

// Autotuner: one ordinary step, then the gather and refine kernels are relaunched on that step's
// buffers with 2, 4 and 8 lanes per point (both kernels are pure functions of their inputs) and
// timed with HIP events; the fastest setting of each is kept for the following steps.  The gather
// goes first, on the records as the step left them (weight slot = val / sw); then the records'
// refine weights exp(val) are restored for refine.  The lane
// count sets the order of each point's partial sums, so a tuned simulation is deterministic but
// may differ from an untuned one in the last bits.
int drb_sim::autotune(int* out) {
  if (tape_pos < tape_n) return fail(DRB_ERR_STATE, "drb_sim_autotune: recording in progress");
  DRB_CHECK(step_sync());
  if (last != &w0 || w0.M == 0) {
    out[0] = g_refine;
    out[1] = g_gather;
    return DRB_OK;
  }
  Work& w = w0;
  const long long N = w.N, ncells = w.ncells;
  const long long pcap = std::max<long long>(1, std::min<long long>(N, ncells * (w.sub_bits ? 5 : 1)));
  const long long pguess = std::min<long long>(pcap, w.M + w.M / 8 + 1024);
  const int* Ptot = w.rp.as<int>() + 2 * ncells + 1;
  const int* Ptot_e = spawn ? sptot.as<int>() : Ptot;  // the gather's points (with the spawned ones)
  const Inst* I = w.d_inst.as<Inst>();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DRB_HIP(hipEventCreate(&e0));
  DRB_HIP(hipEventCreate(&e1));
  // tiles (lanes per point, candidates in flight per lane): short window ranges (fine query cells)
  // favour fewer slots per pass, long ones more lanes
  const int tiles[5][2] = {{2, 4}, {4, 2}, {4, 4}, {8, 2}, {8, 4}};
  float best[2] = {1e30f, 1e30f};
  int pick[2][2] = {{g_refine, d_refine}, {g_gather, d_gather}};
  for (int kind : {1, 0}) {
    if (kind == 0)
      hipLaunchKernelGGL(k_weights, dim3(nblk(N)), dim3(TB), 0, s, w.cv.as<float>(), (int)N, w.comp.as<Comp>());
    for (const auto& t : tiles) {
      const int g = t[0], k = t[1];
      float ms = 0.f;
      for (int rep = 0; rep < 4; ++rep) {
        if (rep == 1) DRB_HIP(hipEventRecord(e0, s));
        if (kind == 0)
          DRB_LAUNCH_PMGK(k_refine, pmode, g, k, dim3(xblk(gblk(pguess, g))), dim3(TB), 0, s, Ptot, (int)pcap,
                         w.cp0x.as<float>(), w.cp0y.as<float>(), w.cpi.as<int>(), w.cpq.as<int>(), I, w.rp.as<int>(),
                         w.comp.as<Comp>(), w.cpx.as<float>(), w.cpy.as<float>(), w.kap.as<unsigned char>(),
                         disp.as<int>(), (float*)nullptr, status.as<unsigned>(), (uint32_t*)nullptr, (int2*)nullptr,
                         (int*)nullptr);
        else
          DRB_LAUNCH_PMGK(k_gather, pmode, g, k, dim3(xblk(gblk(pguess, g))), dim3(TB), 0, s, Ptot_e, (int)pcap,
                         w.cpx.as<float>(), w.cpy.as<float>(), w.cpi.as<int>(), I, w.rp.as<int>(), w.comp.as<Comp>(),
                         w.cpf.as<float>(), w.cpcap, w.emask.as<unsigned char>(), (const int*)nullptr,
                         (unsigned char*)nullptr);
      }
      DRB_HIP(hipEventRecord(e1, s));
      DRB_HIP(hipEventSynchronize(e1));
      DRB_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best[kind]) {
        best[kind] = ms;
        pick[kind][0] = g;
        pick[kind][1] = k;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // Leave the step's buffers as the step wrote them (diagnostics read them): refine, the
  // normaliser (which turns the record weights back into val / sw) and the gather again, with the
  // lanes the step ran with.  All three are pure functions of buffers the timing did not touch.
  hipLaunchKernelGGL(k_weights, dim3(nblk(N)), dim3(TB), 0, s, w.cv.as<float>(), (int)N, w.comp.as<Comp>());
  DRB_LAUNCH_PMGK(k_refine, pmode, g_refine, d_refine, dim3(xblk(gblk(pguess, g_refine))), dim3(TB), 0, s, Ptot, (int)pcap,
                 w.cp0x.as<float>(), w.cp0y.as<float>(), w.cpi.as<int>(), w.cpq.as<int>(), I, w.rp.as<int>(),
                 w.comp.as<Comp>(), w.cpx.as<float>(), w.cpy.as<float>(), w.kap.as<unsigned char>(), disp.as<int>(),
                 (float*)nullptr, status.as<unsigned>(), (uint32_t*)nullptr, (int2*)nullptr, (int*)nullptr);
  DRB_LAUNCH_PM(k_norm, pmode, dim3(nblk(N)), dim3(TB), 0, s, (int)N,
                obst ? w.rp.as<int>() + 2 * ncells : (const int*)nullptr, w.comp.as<Comp>(), w.cv.as<float>(), I,
                brp.as<int>(), bx.as<float>(), by.as<float>(), (const int*)nullptr, w.swc.as<float>());
  DRB_LAUNCH_PMGK(k_gather, pmode, g_gather, d_gather, dim3(xblk(gblk(pguess, g_gather))), dim3(TB), 0, s, Ptot_e, (int)pcap,
                 w.cpx.as<float>(), w.cpy.as<float>(), w.cpi.as<int>(), I, w.rp.as<int>(), w.comp.as<Comp>(),
                 w.cpf.as<float>(), w.cpcap, w.emask.as<unsigned char>(), bslot.as<int>(), bem.as<unsigned char>());
  DRB_HIP(hipGetLastError());
  DRB_HIP(hipStreamSynchronize(s));
  g_refine = pick[0][0];
  d_refine = pick[0][1];
  g_gather = pick[1][0];
  d_gather = pick[1][1];
  graph_drop();
  out[0] = g_refine;
  out[1] = g_gather;
  return DRB_OK;
}

int drb_sim::init(const drb_params* prm, int n, hipStream_t stream) {
  ninst = n;
  s = stream;
  hin.resize(n);
  counts.assign(n, 0);
  for (int b = 0; b < n; ++b) DRB_CHECK(make_inst(prm[b], hin[b]));
  spawn = false;
  for (int b = 0; b < n; ++b) spawn = spawn || hin[b].kappa_soft > 0;
  if (spawn && n > 1) return fail(DRB_ERR_INVALID_ARG, "the soft_outer spawn runs in single simulations (n_inst = 1)");
  int nper = 0;
  for (int b = 0; b < n; ++b) nper += hin[b].periodic;
  pmode = nper == 0 ? 0 : (nper == n ? 1 : 2);
  DRB_HIP(hipHostMalloc(&hpin, pin_need(), hipHostMallocDefault));
  hpin_bytes = pin_need();
  std::memset(hpin, 0, hpin_bytes);
  // boxes, counts and the status word in one device block laid out as the pinned mirror, so the
  // per-step read is one copy
  DRB_CHECK(hio.ensure(pin_need()));
  box.set_view(hio.p);
  dcounts.set_view((char*)hio.p + (size_t)n * sizeof(BBox));
  status.set_view((char*)hio.p + (size_t)n * (sizeof(BBox) + sizeof(int)));
  DRB_CHECK(disp.ensure(sizeof(int) * n));
  DRB_CHECK(moff.ensure(sizeof(long long) * (n + 1)));
  DRB_CHECK(diag.ensure(sizeof(double) * 5 * n));
  DRB_CHECK(d_inst.ensure(sizeof(Inst) * n));
  DRB_HIP(hipMemsetAsync(status.p, 0, sizeof(unsigned), s));
  DRB_HIP(hipMemcpyAsync(d_inst.p, hin.data(), sizeof(Inst) * n, hipMemcpyHostToDevice, s));
  for (auto& set : ev)
    for (auto& e : set) DRB_HIP(hipEventCreate(&e));
  DRB_HIP(hipEventCreateWithFlags(&ev_mid, hipEventDisableTiming));
  return DRB_OK;
}

int drb_sim::set_state(const float* pos, const float* f, long long m, const int64_t* cnt, bool device) {
  long long tot = 0;
  for (int b = 0; b < ninst; ++b) {
    long long c = cnt ? (long long)cnt[b] : m;
    if (c < 0) return fail(DRB_ERR_INVALID_ARG, "negative instance count");
    tot += c;
  }
  if (tot != m) return fail(DRB_ERR_INVALID_ARG, "instance counts do not sum to m");
  if (m > (1LL << 28)) return fail(DRB_ERR_CAPACITY, "too many moments");
  for (int b = 0; b < ninst; ++b) counts[b] = cnt ? (long long)cnt[b] : m;
  M = 0;
  cur = 0;
  DRB_CHECK(ensure_state(0, std::max<long long>(m, 1)));
  M = m;
  last = nullptr;
  std::vector<long long> offs(ninst + 1, 0);
  for (int b = 0; b < ninst; ++b) offs[b + 1] = offs[b] + counts[b];
  DRB_HIP(hipMemcpyAsync(moff.p, offs.data(), sizeof(long long) * (ninst + 1), hipMemcpyHostToDevice, s));
  if (m > 0) {
    Buf tp, tf;
    const float *dp = pos, *df = f;
    if (!device) {
      DRB_CHECK(tp.ensure((size_t)m * 8));
      DRB_CHECK(tf.ensure((size_t)m * 28));
      DRB_HIP(hipMemcpyAsync(tp.p, pos, (size_t)m * 8, hipMemcpyHostToDevice, s));
      DRB_HIP(hipMemcpyAsync(tf.p, f, (size_t)m * 28, hipMemcpyHostToDevice, s));
      dp = tp.as<float>();
      df = tf.as<float>();
    }
    hipLaunchKernelGGL(k_unpack, dim3(nblk(m)), dim3(TB), 0, s, dp, df, (int)m, mx[0].as<float>(), my[0].as<float>(),
                       mf[0].as<float>(), mcap[0]);
    hipLaunchKernelGGL(k_fill_inst, dim3(ninst), dim3(TB), 0, s, mi[0].as<int>(), moff.as<long long>(), ninst);
    DRB_HIP(hipStreamSynchronize(s));
    tp.release();
    tf.release();
  }
  DRB_CHECK(rebox());
  if (*hstatus() & ST_NONFINITE) return fail(DRB_ERR_NONFINITE, "non-finite value in the initial state");
  if (*hstatus() & ST_RANGE) return fail(DRB_ERR_INVALID_ARG, "periodic positions must lie in [0, box)");
  have_state = true;
  return DRB_OK;
}

// Bounding boxes of the current moments (and a finiteness and range check) into box and the
// pinned mirror; the status word is reset first.  Waits for the stream.
int drb_sim::rebox() {
  hipLaunchKernelGGL(k_box_init, dim3((ninst + 63) / 64), dim3(64), 0, s, box.as<BBox>(), disp.as<int>(), ninst);
  DRB_HIP(hipMemsetAsync(status.p, 0, sizeof(unsigned), s));
  if (M > 0)
    hipLaunchKernelGGL(k_collide_box, dim3(nblk(M)), dim3(TB), 0, s, mf[cur].as<float>(), mcap[cur], mx[cur].as<float>(),
                       my[cur].as<float>(), mi[cur].as<int>(), (int)M, (const int*)nullptr, d_inst.as<Inst>(),
                       box.as<BBox>(), status.as<unsigned>(), 0, 1, (float*)nullptr, 0LL);
  DRB_HIP(hipGetLastError());
  DRB_HIP(hipMemcpyAsync(hbox(), box.p, (size_t)ninst * sizeof(BBox), hipMemcpyDeviceToHost, s));
  DRB_HIP(hipMemcpyAsync(hstatus(), status.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  DRB_HIP(hipStreamSynchronize(s));
  return DRB_OK;
}

int drb_sim::get_state(float* pos, float* f, bool device) {
  if (M == 0) return DRB_OK;
  Buf tp, tf;
  float *dp = pos, *df = f;
  if (!device || !pos) {
    DRB_CHECK(tp.ensure((size_t)M * 8));
    dp = tp.as<float>();
  }
  if (!device || !f) {
    DRB_CHECK(tf.ensure((size_t)M * 28));
    df = tf.as<float>();
  }
  hipLaunchKernelGGL(k_pack, dim3(nblk(M)), dim3(TB), 0, s, mx[cur].as<float>(), my[cur].as<float>(),
                     mf[cur].as<float>(), mcap[cur], (int)M, dp, df);
  DRB_HIP(hipGetLastError());
  if (!device) {
    if (pos) DRB_HIP(hipMemcpyAsync(pos, tp.p, (size_t)M * 8, hipMemcpyDeviceToHost, s));
    if (f) DRB_HIP(hipMemcpyAsync(f, tf.p, (size_t)M * 28, hipMemcpyDeviceToHost, s));
  }
  DRB_HIP(hipStreamSynchronize(s));
  tp.release();
  tf.release();
  return DRB_OK;
}

int drb_sim::diagnostics(double* out) {
  std::vector<long long> offs(ninst + 1, 0);
  for (int b = 0; b < ninst; ++b) offs[b + 1] = offs[b] + counts[b];
  DRB_HIP(hipMemcpyAsync(moff.p, offs.data(), sizeof(long long) * (ninst + 1), hipMemcpyHostToDevice, s));
  DRB_CHECK(diagpart.ensure(sizeof(double) * 5 * kDiagChunks * ninst));
  Work* w = last;
  hipLaunchKernelGGL(k_diag, dim3(kDiagChunks, ninst), dim3(TB), 0, s, mf[cur].as<float>(), mcap[cur],
                     moff.as<long long>(), w ? w->cv.as<float>() : (const float*)nullptr,
                     w ? w->swc.as<float>() : (const float*)nullptr, w ? w->rp.as<int2>() : (const int2*)nullptr,
                     w ? w->d_inst.as<Inst>() : d_inst.as<Inst>(), w ? w->cpf.as<float>() : (const float*)nullptr,
                     w ? w->cpcap : 0LL, w ? w->emask.as<unsigned char>() : (const unsigned char*)nullptr,
                     w ? 1 : 0, (w && obst) ? koff.as<int2>() : (const int2*)nullptr,
                     (w && spawn) ? (const int*)sptot.as<int>() : (const int*)nullptr, diagpart.as<double>());
  hipLaunchKernelGGL(k_diag_final, dim3((5 * ninst + 63) / 64), dim3(64), 0, s, diagpart.as<double>(), ninst,
                     diag.as<double>());
  DRB_HIP(hipGetLastError());
  DRB_HIP(hipMemcpyAsync(out, diag.p, sizeof(double) * 5 * ninst, hipMemcpyDeviceToHost, s));
  DRB_HIP(hipStreamSynchronize(s));
  return DRB_OK;
}

extern "C" int drb_sim_create(drb_sim** out, const drb_params* prm, int32_t n_inst, drb_stream_t s) {
  if (!out || !prm || n_inst < 1) return fail(DRB_ERR_INVALID_ARG, "drb_sim_create: bad arguments");
  *out = nullptr;
  drb_sim* sim = new (std::nothrow) drb_sim();
  if (!sim) return fail(DRB_ERR_CAPACITY, "out of host memory");
  int st = sim->init(prm, n_inst, (hipStream_t)s);
  if (st != DRB_OK) {
    delete sim;
    return st;
  }
  *out = sim;
  return DRB_OK;
}

extern "C" int drb_sim_destroy(drb_sim* sim) {
  if (!sim) return DRB_OK;
  (void)hipStreamSynchronize(sim->s);
  delete sim;
  return DRB_OK;
}

extern "C" int drb_sim_set_state(drb_sim* sim, const float* pos, const float* f, int64_t m, const int64_t* counts) {
  if (!sim || m < 0 || (m > 0 && (!pos || !f))) return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_state: bad arguments");
  if (sim->ninst > 1 && !counts) return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_state: counts required");
  return sim->set_state(pos, f, m, counts, false);
}

extern "C" int drb_sim_set_state_device(drb_sim* sim, const float* pos, const float* f, int64_t m,
                                        const int64_t* counts) {
  if (!sim || m < 0 || (m > 0 && (!pos || !f)))
    return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_state_device: bad arguments");
  if (sim->ninst > 1 && !counts) return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_state_device: counts required");
  return sim->set_state(pos, f, m, counts, true);
}

extern "C" int drb_sim_step(drb_sim* sim, int32_t nsteps) {
  if (!sim || nsteps < 0) return fail(DRB_ERR_INVALID_ARG, "drb_sim_step: bad arguments");
  if (!sim->have_state) return fail(DRB_ERR_STATE, "drb_sim_step: no state loaded");
  if (sim->graph_on && !sim->timing && sim->tape_pos >= sim->tape_n) return sim->step_graph(nsteps);
  // Each step's read back is waited for after the step is queued whole, so the host lays out and
  // queues the next step while the GPU finishes the previous one.
  sim->stage_n = 0;
  for (int i = 0; i < kStages; ++i) sim->stage_sum[i] = 0.f;
  sim->back_unread = false;
  for (int i = 0; i < nsteps; ++i) {
    DRB_CHECK(sim->step_once());
    DRB_CHECK(sim->step_finish());
  }
  DRB_HIP(hipStreamSynchronize(sim->s));
  if (sim->timing && sim->back_unread) DRB_CHECK(sim->stage_read(sim->tpar ^ 1, 6, kStages));
  sim->back_unread = false;
  if (sim->timing && sim->stage_n > 0)
    for (int i = 0; i < kStages; ++i) sim->stage_ms[i] = sim->stage_sum[i] / (float)sim->stage_n;
  return DRB_OK;
}

extern "C" int drb_sim_autotune(drb_sim* sim, int32_t* lanes) {
  if (!sim || !lanes) return fail(DRB_ERR_INVALID_ARG, "drb_sim_autotune: bad arguments");
  if (!sim->have_state) return fail(DRB_ERR_STATE, "drb_sim_autotune: no state loaded");
  int out[2];
  DRB_CHECK(sim->autotune(out));
  lanes[0] = out[0];
  lanes[1] = out[1];
  return DRB_OK;
}

extern "C" int drb_sim_set_lanes(drb_sim* sim, int32_t refine, int32_t gather) {
  auto ok = [](int g) { return g == 2 || g == 4 || g == 8; };
  if (!sim || !ok(refine) || !ok(gather)) return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_lanes: lanes must be 2, 4 or 8");
  sim->g_refine = refine;
  sim->g_gather = gather;
  sim->graph_drop();
  return DRB_OK;
}

extern "C" int drb_sim_set_tiles(drb_sim* sim, int32_t refine_lanes, int32_t refine_depth, int32_t gather_lanes,
                                 int32_t gather_depth) {
  auto okg = [](int g) { return g == 2 || g == 4 || g == 8; };
  auto okk = [](int k) { return k == 2 || k == 4; };
  if (!sim || !okg(refine_lanes) || !okg(gather_lanes) || !okk(refine_depth) || !okk(gather_depth))
    return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_tiles: lanes must be 2, 4 or 8 and depths 2 or 4");
  sim->g_refine = refine_lanes;
  sim->d_refine = refine_depth;
  sim->g_gather = gather_lanes;
  sim->d_gather = gather_depth;
  sim->graph_drop();
  return DRB_OK;
}

extern "C" int drb_sim_get_tiles(drb_sim* sim, int32_t* out) {
  if (!sim || !out) return fail(DRB_ERR_INVALID_ARG, "drb_sim_get_tiles: bad arguments");
  out[0] = sim->g_refine;
  out[1] = sim->d_refine;
  out[2] = sim->g_gather;
  out[3] = sim->d_gather;
  return DRB_OK;
}

// Circular obstacle of instance `inst` (fluid outside), or none for r <= 0.  A periodic instance's
// circle must keep dx clear of the box edges.
extern "C" int drb_sim_set_obstacle(drb_sim* sim, int32_t inst, float cx, float cy, float r) {
  if (!sim || inst < 0 || inst >= sim->ninst || !std::isfinite(cx) || !std::isfinite(cy) || !std::isfinite(r))
    return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_obstacle: bad arguments");
  if (sim->graph_on && r > 0.f) return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_obstacle: graph mode is on");
  if (sim->spawn && r > 0.f)
    return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_obstacle: not supported with the soft_outer spawn");
  Inst& p = sim->hin[inst];
  if (r > 0.f && p.periodic) {
    const float lo = p.dx, hi = p.L - p.dx;
    if (!(cx - r >= lo && cx + r <= hi && cy - r >= lo && cy + r <= hi))
      return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_obstacle: the circle must stay dx inside the periodic box");
  }
  p.obstacle = r > 0.f ? 1 : 0;
  p.ocx = cx;
  p.ocy = cy;
  p.oR = r > 0.f ? r : 0.f;
  p.oR2 = p.oR * p.oR;
  sim->obst = false;
  for (const Inst& q : sim->hin) sim->obst = sim->obst || q.obstacle;
  sim->graph_drop();
  DRB_HIP(hipMemcpyAsync(sim->d_inst.p, sim->hin.data(), sizeof(Inst) * sim->ninst, hipMemcpyHostToDevice, sim->s));
  DRB_HIP(hipStreamSynchronize(sim->s));
  return DRB_OK;
}

extern "C" int drb_sim_set_graph(drb_sim* sim, int32_t on) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_graph: null handle");
  if (on && sim->obst) return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_graph: obstacles are not supported in graph mode");
  if (on && sim->spawn)
    return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_graph: the soft_outer spawn is not supported in graph mode");
  sim->graph_on = on != 0;
  if (!on) sim->graph_drop();
  return DRB_OK;
}

extern "C" int drb_sim_set_graph_horizon(drb_sim* sim, int32_t steps) {
  if (!sim || steps < 1) return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_graph_horizon: bad arguments");
  sim->ghorizon = steps;
  return DRB_OK;
}

extern "C" int drb_sim_graph_info(drb_sim* sim, int64_t* out) {
  if (!sim || !out) return fail(DRB_ERR_INVALID_ARG, "drb_sim_graph_info: bad arguments");
  out[0] = sim->gcaptures;
  out[1] = sim->gresumes;
  out[2] = sim->gncells;
  out[3] = sim->gpcap;
  return DRB_OK;
}

extern "C" int drb_sim_count(drb_sim* sim, int64_t* m, int64_t* counts) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim_count: null handle");
  if (m) *m = sim->M;
  if (counts)
    for (int b = 0; b < sim->ninst; ++b) counts[b] = sim->counts[b];
  return DRB_OK;
}

extern "C" int drb_sim_get_state(drb_sim* sim, float* pos, float* f) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim_get_state: null handle");
  return sim->get_state(pos, f, false);
}

extern "C" int drb_sim_get_state_device(drb_sim* sim, float* pos, float* f) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim_get_state_device: null handle");
  return sim->get_state(pos, f, true);
}

extern "C" int drb_sim_tape(drb_sim* sim, int32_t nslots) {
  if (!sim || nslots < 0) return fail(DRB_ERR_INVALID_ARG, "drb_sim_tape: bad arguments");
  if (sim->last != &sim->w0) sim->last = nullptr;  // tape slots may move or be released below
  if (nslots == 0) {
    for (auto& w : sim->tape) w.release();
    sim->tape.clear();
  } else if ((int)sim->tape.size() < nslots) {
    sim->tape.resize(nslots);  // slots are kept (and reused) when a later recording is shorter
  }
  for (auto& w : sim->tape) w.taped = false;
  sim->tape_n = nslots;
  sim->tape_pos = 0;
  return DRB_OK;
}

extern "C" int drb_sim_tape_info(drb_sim* sim, int32_t slot, int64_t* m_in, int64_t* p_out) {
  if (!sim || slot < 0 || slot >= sim->tape_n) return fail(DRB_ERR_INVALID_ARG, "drb_sim_tape_info: bad slot");
  const Work& w = sim->tape[slot];
  if (!w.taped) return fail(DRB_ERR_STATE, "drb_sim_tape_info: slot not recorded");
  if (m_in) *m_in = w.M;
  if (p_out) *p_out = w.P;
  return DRB_OK;
}

extern "C" int drb_sim_diagnostics(drb_sim* sim, double* out) {
  if (!sim || !out) return fail(DRB_ERR_INVALID_ARG, "drb_sim_diagnostics: bad arguments");
  return sim->diagnostics(out);
}

extern "C" int drb_sim_set_timing(drb_sim* sim, int32_t on) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_timing: null handle");
  if (on < 0 || (on > 1 && ((on & ~0x3ff) != 0 || !(on & 0x200))))
    return fail(DRB_ERR_INVALID_ARG, "drb_sim_set_timing: 0, 1 or 0x200 | stage mask");
  sim->timing = on == 1 ? 0x1ffu : (unsigned)on & 0x1ffu;
  return DRB_OK;
}

extern "C" int drb_sim_updates(drb_sim* sim, int64_t* out) {
  if (!sim || !out) return fail(DRB_ERR_INVALID_ARG, "drb_sim_updates: bad arguments");
  *out = sim->updates;
  return DRB_OK;
}

extern "C" int drb_sim_stage_ms(drb_sim* sim, float* ms9) {
  if (!sim || !ms9) return fail(DRB_ERR_INVALID_ARG, "drb_sim_stage_ms: bad arguments");
  for (int i = 0; i < kStages; ++i) ms9[i] = sim->stage_ms[i];
  return DRB_OK;
}

// =================================================================================== stages
extern "C" int drb_stage_create(drb_stage** out) {
  if (!out) return fail(DRB_ERR_INVALID_ARG, "drb_stage_create: null");
  *out = nullptr;
  drb_stage* st = new (std::nothrow) drb_stage();
  if (!st) return fail(DRB_ERR_CAPACITY, "out of host memory");
  if (hipHostMalloc(&st->hpin, 64, hipHostMallocDefault) != hipSuccess) {
    delete st;
    return fail(DRB_ERR_HIP, "hipHostMalloc failed");
  }
  *out = st;
  return DRB_OK;
}

extern "C" int drb_stage_destroy(drb_stage* st) {
  delete st;
  return DRB_OK;
}

extern "C" int drb_radix_sort_pairs(drb_stage* st, uint32_t* keys, uint32_t* vals, int64_t n, int32_t bits,
                                    drb_stream_t stream) {
  if (!st || n < 0 || (n > 0 && (!keys || !vals)) || bits < 1 || bits > 32 || n > INT_MAX / 2)
    return fail(DRB_ERR_INVALID_ARG, "drb_radix_sort_pairs: bad arguments");
  if (n == 0) return DRB_OK;
  DRB_CHECK(st->sorter.reserve(1));  // histograms sized for n below
  DRB_CHECK(st->sorter.hist.ensure(radix_hist_words((int)n) * 4));
  return radix_sort_into(keys, vals, (int)n, bits, false, keys + n, vals + n, st->sorter.hist.as<uint32_t>(),
                         st->sorter.tot.as<uint32_t>(), (hipStream_t)stream);
}

namespace drb {
// Grid over free points: bounding box of (a, na) and (b, nb), cell size `cell`, then keys and a
// stable sort of `a` (and of `b` when given).  Fills p (the single instance) and uploads it.
// Directions of `a` (when given) must lie in 0..6: one beyond would address an eighth sum.
static int stage_grid(drb_stage* st, Inst& p, float cell, const float* a, long long na, const float* b, long long nb,
                      hipStream_t s, long long& ncells, int& bits, const int32_t* dirs = nullptr) {
  std::memset(&p, 0, sizeof(p));
  p.cs = cell;
  p.qcs = cell;
  p.iq = 1.f / cell;
  p.margin = 1e-3f * cell;
  DRB_CHECK(st->d_inst.ensure(sizeof(Inst)));
  DRB_CHECK(st->box.ensure(sizeof(BBox)));
  DRB_CHECK(st->status.ensure(sizeof(unsigned)));
  DRB_HIP(hipMemcpyAsync(st->d_inst.p, &p, sizeof(Inst), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_box_init, dim3(1), dim3(64), 0, s, st->box.as<BBox>(), (int*)st->status.p, 1);
  DRB_HIP(hipMemsetAsync(st->status.p, 0, sizeof(unsigned), s));
  if (na > 0)
    hipLaunchKernelGGL(k_box_points, dim3(nblk(na)), dim3(TB), 0, s, a, (int)na, st->d_inst.as<Inst>(),
                       st->box.as<BBox>(), st->status.as<unsigned>());
  if (nb > 0)
    hipLaunchKernelGGL(k_box_points, dim3(nblk(nb)), dim3(TB), 0, s, b, (int)nb, st->d_inst.as<Inst>(),
                       st->box.as<BBox>(), st->status.as<unsigned>());
  if (dirs && na > 0)
    hipLaunchKernelGGL(k_check_dirs, dim3(nblk(na)), dim3(TB), 0, s, dirs, (int)na, st->status.as<unsigned>());
  DRB_HIP(hipGetLastError());
  DRB_HIP(hipMemcpyAsync(st->hpin, st->box.p, sizeof(BBox), hipMemcpyDeviceToHost, s));
  DRB_HIP(hipMemcpyAsync((char*)st->hpin + sizeof(BBox), st->status.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  DRB_HIP(hipStreamSynchronize(s));
  const BBox bb = *(BBox*)st->hpin;
  if (*(unsigned*)((char*)st->hpin + sizeof(BBox)) & ST_NONFINITE) return fail(DRB_ERR_NONFINITE, "non-finite position");
  if (*(unsigned*)((char*)st->hpin + sizeof(BBox)) & ST_RANGE) return fail(DRB_ERR_INVALID_ARG, "direction outside 0..6");
  if (bb.qxmin > bb.qxmax) {
    p.nqx = p.nqy = 0;
  } else {
    long long nx = (long long)bb.qxmax - bb.qxmin + 1, ny = (long long)bb.qymax - bb.qymin + 1;
    if (nx > (1LL << 28) || ny > (1LL << 28) || nx * ny > (1LL << 31)) return fail(DRB_ERR_KEY_RANGE, "points span too many cells");
    p.qx0 = bb.qxmin;
    p.qy0 = bb.qymin;
    p.nqx = (int)nx;
    p.nqy = (int)ny;
    const double far = std::max({std::fabs((double)p.qx0), std::fabs((double)p.qx0 + p.nqx),
                                 std::fabs((double)p.qy0), std::fabs((double)p.qy0 + p.nqy)});
    p.margin = (float)((double)cell * std::max(1e-3, 16.0 * far * 0x1p-24));
  }
  ncells = (long long)p.nqx * p.nqy;
  bits = 1;
  while (bits < 32 && (1LL << bits) < ncells) ++bits;
  DRB_HIP(hipMemcpyAsync(st->d_inst.p, &p, sizeof(Inst), hipMemcpyHostToDevice, s));
  DRB_CHECK(st->cells.ensure((size_t)(ncells + 1) * 8));
  DRB_CHECK(st->rp.ensure((size_t)(ncells + 1) * 8));
  DRB_HIP(hipMemsetAsync(st->cells.p, 0, (size_t)(ncells + 1) * 8, s));
  if (na > 0) {
    DRB_CHECK(st->keys.ensure((size_t)na * 4));
    DRB_CHECK(st->vals.ensure((size_t)na * 4));
    hipLaunchKernelGGL(k_keys_points, dim3(nblk(na)), dim3(TB), 0, s, a, (int)na, st->d_inst.as<Inst>(),
                       st->keys.as<uint32_t>(), st->status.as<unsigned>());
    DRB_CHECK(st->sorter.sort(st->keys.as<uint32_t>(), st->vals.as<uint32_t>(), (int)na, bits, true, s));
  }
  if (nb > 0) {
    DRB_CHECK(st->ckeys.ensure((size_t)nb * 4));
    DRB_CHECK(st->cvals.ensure((size_t)nb * 4));
    hipLaunchKernelGGL(k_keys_points, dim3(nblk(nb)), dim3(TB), 0, s, b, (int)nb, st->d_inst.as<Inst>(),
                       st->ckeys.as<uint32_t>(), st->status.as<unsigned>());
    DRB_CHECK(st->sorter.sort(st->ckeys.as<uint32_t>(), st->cvals.as<uint32_t>(), (int)nb, bits, true, s));
  }
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}
}  // namespace drb

extern "C" int drb_create_control_points(drb_stage* st, const float* pos_c, int64_t nc, float cs, float* cp_out,
                                         int64_t cp_cap, int64_t* p_out, drb_stream_t stream) {
  if (!st || nc < 0 || (nc > 0 && (!pos_c || !cp_out)) || !p_out || !(cs > 0.f) || nc > INT_MAX / 2)
    return fail(DRB_ERR_INVALID_ARG, "drb_create_control_points: bad arguments");
  *p_out = 0;
  if (nc == 0) return DRB_OK;
  hipStream_t s = (hipStream_t)stream;
  Inst p;
  long long ncells;
  int bits;
  DRB_CHECK(stage_grid(st, p, cs, pos_c, nc, nullptr, 0, s, ncells, bits));
  DRB_CHECK(st->comp.ensure((size_t)nc * sizeof(Comp)));
  DRB_CHECK(st->cv.ensure((size_t)nc * 4));
  hipLaunchKernelGGL(k_mark<1>, dim3(nblk(nc)), dim3(TB), 0, s, st->keys.as<uint32_t>(), st->vals.as<uint32_t>(), (int)nc,
                     (const int*)nullptr, 0, (const float*)nullptr, (const float*)nullptr, (const float*)nullptr, 0LL, (const int*)nullptr,
                     pos_c, (const float*)nullptr, (const int*)nullptr, st->d_inst.as<Inst>(), st->comp.as<Comp>(),
                     st->cv.as<float>(), st->cells.as<int>(), 1, INT_MAX, (const float4*)nullptr, (const float2*)nullptr, (const float4*)nullptr);
  DRB_CHECK(st->scan.run(st->cells.as<int2>(), st->rp.as<int2>(), ncells, s));
  DRB_HIP(hipMemcpyAsync(st->hpin, st->rp.as<int2>() + ncells, sizeof(int2), hipMemcpyDeviceToHost, s));
  DRB_HIP(hipStreamSynchronize(s));
  const int P = ((int2*)st->hpin)->y;
  *p_out = P;
  if (P > cp_cap) return fail(DRB_ERR_CAPACITY, "drb_create_control_points: cp_cap too small");
  DRB_CHECK(st->cpx.ensure((size_t)P * 4 + 4));
  DRB_CHECK(st->cpy.ensure((size_t)P * 4 + 4));
  hipLaunchKernelGGL(k_runs, dim3(nblk(nc)), dim3(TB), 0, s, st->rp.as<int2>(), ncells, (int)nc, (const int*)nullptr,
                     0, st->keys.as<uint32_t>(), st->comp.as<Comp>(), st->cpx.as<float>(), st->cpy.as<float>(),
                     (int*)nullptr, (int*)nullptr, (int*)nullptr, (int*)nullptr, (int*)nullptr, (int2*)nullptr, 0LL,
                     (const Inst*)nullptr, 0, (int*)nullptr);
  DRB_HIP(hipGetLastError());
  DRB_HIP(hipMemcpy2DAsync(cp_out, 8, st->cpx.p, 4, 4, P, hipMemcpyDeviceToDevice, s));
  DRB_HIP(hipMemcpy2DAsync(cp_out + 1, 8, st->cpy.p, 4, 4, P, hipMemcpyDeviceToDevice, s));
  return DRB_OK;
}

extern "C" int drb_refine_control_points(drb_stage* st, const float* cp, int64_t np, const float* pos_c,
                                         const float* val_c, const int32_t* dir_c, int64_t nc, float dx,
                                         int kappa_hard, float* new_pos, drb_stream_t stream) {
  if (!st || np < 0 || nc < 0 || (np > 0 && (!cp || !new_pos)) || (nc > 0 && (!pos_c || !val_c || !dir_c)) ||
      !(dx > 0.f) || nc > INT_MAX / 2 || np > INT_MAX / 2)
    return fail(DRB_ERR_INVALID_ARG, "drb_refine_control_points: bad arguments");
  if (np == 0) return DRB_OK;
  hipStream_t s = (hipStream_t)stream;
  Inst p;
  long long ncells;
  int bits;
  DRB_CHECK(stage_grid(st, p, dx * 0.25f, pos_c, nc, nullptr, 0, s, ncells, bits, dir_c));
  p.dx = dx;
  p.t_dx = sqrt_threshold(dx);
  p.kappa_hard = kappa_hard;
  DRB_HIP(hipMemcpyAsync(st->d_inst.p, &p, sizeof(Inst), hipMemcpyHostToDevice, s));
  DRB_CHECK(st->comp.ensure((size_t)std::max<long long>(nc, 1) * sizeof(Comp)));
  DRB_CHECK(st->cv.ensure((size_t)std::max<long long>(nc, 1) * 4));
  if (nc > 0)
    hipLaunchKernelGGL(k_mark<1>, dim3(nblk(nc)), dim3(TB), 0, s, st->keys.as<uint32_t>(), st->vals.as<uint32_t>(),
                       (int)nc, (const int*)nullptr, 0, (const float*)nullptr, (const float*)nullptr,
                       (const float*)nullptr, 0LL,
                       (const int*)nullptr, pos_c, val_c, dir_c, st->d_inst.as<Inst>(), st->comp.as<Comp>(),
                       st->cv.as<float>(), st->cells.as<int>(), 0, INT_MAX, (const float4*)nullptr, (const float2*)nullptr, (const float4*)nullptr);
  DRB_CHECK(st->scan.run(st->cells.as<int2>(), st->rp.as<int2>(), ncells, s));
  DRB_CHECK(st->cpx.ensure((size_t)np * 4));
  DRB_CHECK(st->cpy.ensure((size_t)np * 4));
  DRB_HIP(hipMemcpy2DAsync(st->cpx.p, 4, cp, 8, 4, np, hipMemcpyDeviceToDevice, s));
  DRB_HIP(hipMemcpy2DAsync(st->cpy.p, 4, cp + 1, 8, 4, np, hipMemcpyDeviceToDevice, s));
  DRB_CHECK(st->cpsx.ensure((size_t)np * 4));
  DRB_CHECK(st->cpsy.ensure((size_t)np * 4));
  DRB_LAUNCH_PMGK(k_refine, 0, kGroup, kInflight, dim3(xblk(gblk(np))), dim3(TB), 0, s, (const int*)nullptr, (int)np, st->cpx.as<float>(),
                     st->cpy.as<float>(), (const int*)nullptr, (const int*)nullptr, st->d_inst.as<Inst>(),
                     st->rp.as<int>(), st->comp.as<Comp>(), st->cpsx.as<float>(), st->cpsy.as<float>(),
                     (unsigned char*)nullptr, (int*)nullptr, (float*)nullptr, st->status.as<unsigned>(),
                     (uint32_t*)nullptr, (int2*)nullptr, (int*)nullptr);
  DRB_HIP(hipGetLastError());
  DRB_HIP(hipMemcpy2DAsync(new_pos, 8, st->cpsx.p, 4, 4, np, hipMemcpyDeviceToDevice, s));
  DRB_HIP(hipMemcpy2DAsync(new_pos + 1, 8, st->cpsy.p, 4, 4, np, hipMemcpyDeviceToDevice, s));
  return DRB_OK;
}

extern "C" int drb_resample(drb_stage* st, const float* pos_c, const float* val_c, const int32_t* dir_c, int64_t nc,
                            const float* cp, int64_t np, float dx, const float* f_rest, float* out_pos, float* out_f,
                            drb_stream_t stream) {
  if (!st || np < 0 || nc < 0 || (np > 0 && (!cp || !out_pos || !out_f || !f_rest)) ||
      (nc > 0 && (!pos_c || !val_c || !dir_c)) || !(dx > 0.f) || nc > INT_MAX / 2 || np > INT_MAX / 2)
    return fail(DRB_ERR_INVALID_ARG, "drb_resample: bad arguments");
  if (np == 0) return DRB_OK;
  hipStream_t s = (hipStream_t)stream;
  Inst p;
  long long ncells;
  int bits;
  DRB_CHECK(stage_grid(st, p, dx * 0.25f, pos_c, nc, cp, np, s, ncells, bits, dir_c));
  p.dx = dx;
  p.rr = dx;
  p.t_dx = p.t_rr = sqrt_threshold(dx);
  p.eps = 1e-6f;
  float fr[7];
  DRB_HIP(hipMemcpyAsync(fr, f_rest, sizeof(fr), hipMemcpyDeviceToHost, s));
  DRB_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < 7; ++k) p.frest[k] = fr[k];
  DRB_HIP(hipMemcpyAsync(st->d_inst.p, &p, sizeof(Inst), hipMemcpyHostToDevice, s));
  DRB_CHECK(st->comp.ensure((size_t)std::max<long long>(nc, 1) * sizeof(Comp)));
  DRB_CHECK(st->cv.ensure((size_t)std::max<long long>(nc, 1) * 4));
  DRB_CHECK(st->swc.ensure((size_t)std::max<long long>(nc, 1) * 4));
  if (nc > 0)
    hipLaunchKernelGGL(k_mark<1>, dim3(nblk(nc)), dim3(TB), 0, s, st->keys.as<uint32_t>(), st->vals.as<uint32_t>(),
                       (int)nc, (const int*)nullptr, 0, (const float*)nullptr, (const float*)nullptr,
                       (const float*)nullptr, 0LL,
                       (const int*)nullptr, pos_c, val_c, dir_c, st->d_inst.as<Inst>(), st->comp.as<Comp>(),
                       st->cv.as<float>(), st->cells.as<int>(), 0, INT_MAX, (const float4*)nullptr, (const float2*)nullptr, (const float4*)nullptr);
  DRB_CHECK(st->cpsx.ensure((size_t)np * 4));
  DRB_CHECK(st->cpsy.ensure((size_t)np * 4));
  hipLaunchKernelGGL(k_mark_cp, dim3(nblk(np)), dim3(TB), 0, s, st->ckeys.as<uint32_t>(), st->cvals.as<uint32_t>(),
                     (int)np, cp, st->cpsx.as<float>(), st->cpsy.as<float>(), st->cells.as<int>());
  DRB_CHECK(st->scan.run(st->cells.as<int2>(), st->rp.as<int2>(), ncells, s));
  DRB_CHECK(st->cpf.ensure((size_t)np * 7 * 4));
  DRB_CHECK(st->emask.ensure((size_t)np));
  if (nc > 0)
    DRB_LAUNCH_PM(k_norm, 0, dim3(nblk(nc)), dim3(TB), 0, s, (int)nc, (const int*)nullptr, st->comp.as<Comp>(), st->cv.as<float>(),
                       st->d_inst.as<Inst>(), st->rp.as<int>(), st->cpsx.as<float>(), st->cpsy.as<float>(),
                       (const int*)nullptr, st->swc.as<float>());
  DRB_LAUNCH_PMGK(k_gather, 0, kGroup, kInflight, dim3(xblk(gblk(np))), dim3(TB), 0, s, (const int*)nullptr, (int)np, st->cpsx.as<float>(),
                     st->cpsy.as<float>(), (const int*)nullptr, st->d_inst.as<Inst>(), st->rp.as<int>(),
                     st->comp.as<Comp>(), st->cpf.as<float>(), (long long)np,
                     st->emask.as<unsigned char>(), (const int*)nullptr, (unsigned char*)nullptr);
  // the control points are in cell order here: their own arrays are the gap-fill's lists
  DRB_LAUNCH_PM(k_gapfill_emit, 0, dim3(emit_blocks(np)), dim3(TB), 0, s, (const int*)nullptr, (int)np, st->cpsx.as<float>(),
                     st->cpsy.as<float>(), (const int*)nullptr, st->d_inst.as<Inst>(), st->rp.as<int>(),
                     st->cpf.as<float>(), (long long)np, st->emask.as<unsigned char>(),
                     st->cvals.as<uint32_t>(), out_pos, out_pos + 1, 2LL, out_f, 7LL, 1LL, (int*)nullptr,
                     (unsigned*)nullptr, (float*)nullptr, (const int2*)nullptr,
                     st->emask.as<unsigned char>(), st->cpsx.as<float>(), st->cpsy.as<float>());
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

// ==================================================================== allocation-free thinning
// drb_create_control_points_noalloc: the thinning stage of drb_create_control_points with every
// intermediate in one static device array, so a call allocates no device memory.  The array is
// carved per call: the segment-table entry, the bounding box and status word, the sort's key and
// value buffers (twice, ping-pong) and its histograms, the component records and values, the
// centroids, and the cell counts, prefix and scan partials of the components' bounding grid.
namespace drb {
constexpr size_t kNoAllocBytes = (size_t)DRB_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[kMaxDevices] = {};

// Device address of this GPU's copy of `scratch` (each GPU holds its own, at its own address),
// looked up once per device with that device current.  Host code uses only this address: the
// array's own name is a host-side stand-in that no kernel may be handed.
static int noalloc_scratch(unsigned char** out) {
  int dev = 0;
  DRB_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(DRB_ERR_INVALID_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(DRB_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = (unsigned char*)g_scratch[dev];
  return DRB_OK;
}

// Bytes of the array a call over nc components and a grid of ncells cells uses (each region
// 256-byte aligned); the layout below follows the same order.
static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
static size_t noalloc_need(long long nc, long long ncells) {
  const size_t n = (size_t)std::max<long long>(nc, 1);
  return al256(sizeof(Inst)) + al256(sizeof(BBox) + sizeof(unsigned)) + 4 * al256(n * 4) +
         al256(radix_hist_words((int)n) * 4) + al256(256 * 4) + al256(n * sizeof(Comp)) + 3 * al256(n * 4) +
         2 * al256((size_t)(ncells + 1) * 8) + al256(scan2_part_elems(ncells) * sizeof(int2));
}

// Interleave the centroids into the caller's (P, 2) array.
__global__ __launch_bounds__(TB) void k_pack_xy(const float* __restrict__ x, const float* __restrict__ y, int n,
                                                float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    out[2 * (size_t)i] = x[i];
    out[2 * (size_t)i + 1] = y[i];
  }
}
}  // namespace drb

extern "C" int drb_create_control_points_noalloc(const float* pos_c, int64_t nc, float cs, float* cp_out,
                                                 int64_t cp_cap, int64_t* p_out, drb_stream_t stream) {
  if (nc < 0 || (nc > 0 && (!pos_c || !cp_out)) || !p_out || !(cs > 0.f) || nc > INT_MAX / 2)
    return fail(DRB_ERR_INVALID_ARG, "drb_create_control_points_noalloc: bad arguments");
  *p_out = 0;
  if (nc == 0) return DRB_OK;
  if (noalloc_need(nc, 0) > kNoAllocBytes)
    return fail(DRB_ERR_CAPACITY, "drb_create_control_points_noalloc: " + std::to_string(nc) +
                                      " components exceed the " + std::to_string(kNoAllocBytes >> 20) +
                                      " MiB scratch array (44 bytes per component + 16 per grid cell)");
  hipStream_t s = (hipStream_t)stream;
  unsigned char* base = nullptr;
  DRB_CHECK(noalloc_scratch(&base));
  // fixed head: segment-table entry, bounding box + status
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += al256(bytes);
    return p;
  };
  Inst* d_inst = (Inst*)take(sizeof(Inst));
  BBox* box = (BBox*)take(sizeof(BBox) + sizeof(unsigned));
  unsigned* status = (unsigned*)(box + 1);
  // Bounding box of the components' cells (as stage_grid, with a host stack mirror).
  Inst p;
  std::memset(&p, 0, sizeof(p));
  p.cs = p.qcs = cs;
  p.iq = 1.f / cs;
  p.margin = 1e-3f * cs;
  DRB_HIP(hipMemcpyAsync(d_inst, &p, sizeof(Inst), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_box_init, dim3(1), dim3(64), 0, s, box, (int*)status, 1);  // box and status word
  hipLaunchKernelGGL(k_box_points, dim3(nblk(nc)), dim3(TB), 0, s, pos_c, (int)nc, (const Inst*)d_inst, box, status);
  DRB_HIP(hipGetLastError());
  struct {
    BBox b;
    unsigned st;
  } h;
  DRB_HIP(hipMemcpyAsync(&h, box, sizeof(h), hipMemcpyDeviceToHost, s));
  DRB_HIP(hipStreamSynchronize(s));
  if (h.st & ST_NONFINITE) return fail(DRB_ERR_NONFINITE, "non-finite position");
  const long long nx = (long long)h.b.qxmax - h.b.qxmin + 1, ny = (long long)h.b.qymax - h.b.qymin + 1;
  if (nx > (1LL << 28) || ny > (1LL << 28) || nx * ny > (1LL << 31))
    return fail(DRB_ERR_KEY_RANGE, "points span too many cells");
  const long long ncells = nx * ny;
  if (noalloc_need(nc, ncells) > kNoAllocBytes)
    return fail(DRB_ERR_CAPACITY, "drb_create_control_points_noalloc: " + std::to_string(nc) + " components over " +
                                      std::to_string(ncells) + " cells exceed the " +
                                      std::to_string(kNoAllocBytes >> 20) +
                                      " MiB scratch array (44 bytes per component + 16 per grid cell)");
  p.qx0 = h.b.qxmin;
  p.qy0 = h.b.qymin;
  p.nqx = (int)nx;
  p.nqy = (int)ny;
  const double far = std::max({std::fabs((double)p.qx0), std::fabs((double)p.qx0 + p.nqx), std::fabs((double)p.qy0),
                               std::fabs((double)p.qy0 + p.nqy)});
  p.margin = (float)((double)cs * std::max(1e-3, 16.0 * far * 0x1p-24));
  int bits = 1;
  while (bits < 32 && (1LL << bits) < ncells) ++bits;
  const size_t n4 = (size_t)nc * 4;
  uint32_t* keys = (uint32_t*)take(n4);
  uint32_t* vals = (uint32_t*)take(n4);
  uint32_t* k1 = (uint32_t*)take(n4);
  uint32_t* v1 = (uint32_t*)take(n4);
  uint32_t* hist = (uint32_t*)take(radix_hist_words((int)nc) * 4);
  uint32_t* tot = (uint32_t*)take(256 * 4);
  Comp* comp = (Comp*)take((size_t)nc * sizeof(Comp));
  float* cv = (float*)take(n4);
  float* cpx = (float*)take(n4);
  float* cpy = (float*)take(n4);
  int2* cells = (int2*)take((size_t)(ncells + 1) * 8);
  int2* rp = (int2*)take((size_t)(ncells + 1) * 8);
  int2* part = (int2*)take(scan2_part_elems(ncells) * sizeof(int2));
  if (off > kNoAllocBytes) return fail(DRB_ERR_CAPACITY, "drb_create_control_points_noalloc: scratch layout overflow");
  DRB_HIP(hipMemcpyAsync(d_inst, &p, sizeof(Inst), hipMemcpyHostToDevice, s));
  // the array keeps the previous call's contents: the cell counts must start at zero
  hipLaunchKernelGGL(k_fill_i32, dim3(std::min<unsigned>(nblk(2 * (ncells + 1)), 4096)), dim3(TB), 0, s, (int*)cells,
                     2 * (ncells + 1), 0);
  hipLaunchKernelGGL(k_keys_points, dim3(nblk(nc)), dim3(TB), 0, s, pos_c, (int)nc, (const Inst*)d_inst, keys, status);
  DRB_CHECK(radix_sort_into(keys, vals, (int)nc, bits, true, k1, v1, hist, tot, s));
  hipLaunchKernelGGL(k_mark<1>, dim3(nblk(nc)), dim3(TB), 0, s, keys, vals, (int)nc, (const int*)nullptr, 0,
                     (const float*)nullptr, (const float*)nullptr, (const float*)nullptr, 0LL, (const int*)nullptr,
                     pos_c, (const float*)nullptr, (const int*)nullptr, (const Inst*)d_inst, comp, cv, (int*)cells, 1,
                     INT_MAX, (const float4*)nullptr, (const float2*)nullptr, (const float4*)nullptr);
  DRB_CHECK(scan2_into(cells, rp, ncells, part, s));
  int2 tot2 = make_int2(0, 0);
  DRB_HIP(hipMemcpyAsync(&tot2, rp + ncells, sizeof(int2), hipMemcpyDeviceToHost, s));
  DRB_HIP(hipStreamSynchronize(s));
  const int P = tot2.y;
  *p_out = P;
  if (P > cp_cap) return fail(DRB_ERR_CAPACITY, "drb_create_control_points_noalloc: cp_cap too small");
  hipLaunchKernelGGL(k_runs, dim3(nblk(nc)), dim3(TB), 0, s, rp, ncells, (int)nc, (const int*)nullptr, 0, keys, comp,
                     cpx, cpy, (int*)nullptr, (int*)nullptr, (int*)nullptr, (int*)nullptr, (int*)nullptr,
                     (int2*)nullptr, 0LL, (const Inst*)nullptr, 0, (int*)nullptr);
  hipLaunchKernelGGL(k_pack_xy, dim3(nblk(P)), dim3(TB), 0, s, cpx, cpy, P, cp_out);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}
