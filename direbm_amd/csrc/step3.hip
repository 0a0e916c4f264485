// The DiReBM step in 3D on the D3Q13 icosahedral lattice, one GPU.
//
// The reference runs 3D only in its per-object CPU solver (direbm/reference/simulator.py:30-198
// with lattice=D3Q13, direbm/lattices.py:44-61); its notes plan the GPU step for 3D as the 2D step
// with vec3 positions, 13 directions and cs^2 = 1/5 (HANDOFF.md:135-172).  This engine is that step
// with the same arithmetic contract and structure as the 2D engine (step.hip): components stay
// implicit (c = 13 i + k), 32-bit dense cell keys (x-major, then y, then z) sorted stably with the
// component index, one scan for the per-cell (components, control points) prefix, and every
// neighbourhood stage a walk of the (x, y) columns of a ball, each column's z-range one contiguous
// run of the sorted arrays.  No atomics on values: results are bit-for-bit reproducible.
//
// Scope: one simulation per handle, open domain or periodic cube, single-level thinning.
#include <type_traits>

#include "engine.h"

namespace drb {

constexpr int Q3 = 13;

// Parameters and dense-grid layout of a 3D simulation (device copy read by every kernel).
struct alignas(16) Inst3 {
  float tau, dx, cs, rr;
  float eps, L, margin, iq;  // iq = 1 / cs
  float t_dx, t_rr;          // radius predicates: s <= t_r  <=>  sqrt_rn(s) <= r
  float frest[Q3];
  float C[Q3][3];            // fp32 velocity set (direbm/lattices.py:44-61, uploaded as fp32)
  float W[Q3];
  int kappa_hard, periodic, n;  // n: cells per period (periodic)
  int x0, y0, z0, nx, ny, nz;   // dense grid origin and extent in cells
  float wlim;                   // periodic: L/2 - 2 cs, the reach below which image shifts are per range
  int sphere;                   // 1: spherical obstacle, fluid outside
  float sc[3], sR, sR2;         // its centre, radius and fp32 R * R
  double Cd[Q3][3];             // FP64 velocity set (bounce split)
};

struct BBox3 {
  int lo[3], hi[3];
};

__device__ __forceinline__ float wrap3(float x, const Inst3& p) {
  if (x < 0.f) x = fadd(x, p.L);
  if (x >= p.L) x = fsub(x, p.L);
  if (floorf(fdiv(x, p.cs)) >= (float)p.n) x = 0.f;
  return x;
}

// Minimum image on the periodic cube (PM 1), the plain difference in the open domain (PM 0); PM 2
// decides at run time.  The neighbourhood kernels are instantiated on PM so that the selects
// vanish from the open-domain loops.
template <int PM = 2>
__device__ __forceinline__ float mi3(float d, const Inst3& p) {
  if (PM == 0) return d;
  return min_image(d, p.L, PM == 1 ? 1 : p.periodic);
}

// fp32 squared distance ((dx^2 + dy^2) + dz^2), one rounding per operation.
template <int PM = 2>
__device__ __forceinline__ float d2_3(float ax, float ay, float az, float bx, float by, float bz, const Inst3& p) {
  const float dx = mi3<PM>(fsub(ax, bx), p), dy = mi3<PM>(fsub(ay, by), p), dz = mi3<PM>(fsub(az, bz), p);
  return fadd(fadd(fmul(dx, dx), fmul(dy, dy)), fmul(dz, dz));
}

__device__ __forceinline__ void comp_pos3(float mx, float my, float mz, int k, const Inst3& p, float& x, float& y,
                                          float& z) {
  x = fadd(mx, fmul(p.C[k][0], p.dx));
  y = fadd(my, fmul(p.C[k][1], p.dx));
  z = fadd(mz, fmul(p.C[k][2], p.dx));
  if (p.periodic) {
    x = wrap3(x, p);
    y = wrap3(y, p);
    z = wrap3(z, p);
  }
}

// Minimum image of a candidate-minus-query difference from its range's wrap code (two bits per
// axis, x in bits 0-1): 0 the plain difference, 1 subtract L, 2 add L, 3 the general minimum image.
// A candidate in a window cell lies within reach = span + rm + cs of every query point the window
// serves (span: the extent of a box window's box).  With reach < L/2, every candidate of an
// unwrapped range has |d| < L/2 (min_image leaves d alone), and every candidate of a range wrapped
// below 0 (above n) has d > L/2 (d < -L/2), where min_image applies exactly the fp32 subtraction
// (addition) of L that the code names.  So codes 0-2 are the minimum image bit for bit, and a
// range skips the per-candidate selects.  Windows with a longer reach use code 3.
template <int PM>
__device__ __forceinline__ float wrapd(float d, int code, const Inst3& p) {
  if (PM == 0 || code == 0) return d;
  if (code == 1) return fsub(d, p.L);
  if (code == 2) return fadd(d, p.L);
  return min_image(d, p.L, 1);
}

// fp32 squared distance of candidate a from query b with the range's wrap code (PLAIN: code 0).
template <int PM, bool PLAIN>
__device__ __forceinline__ float d2w(float ax, float ay, float az, float bx, float by, float bz, int code,
                                     const Inst3& p) {
  float dx = fsub(ax, bx), dy = fsub(ay, by), dz = fsub(az, bz);
  if (!PLAIN) {
    dx = wrapd<PM>(dx, code & 3, p);
    dy = wrapd<PM>(dy, (code >> 2) & 3, p);
    dz = wrapd<PM>(dz, code >> 4, p);
  }
  return fadd(fadd(fmul(dx, dx), fmul(dy, dy)), fmul(dz, dz));
}

// Strictly inside the obstacle, fp32 ((dx^2 + dy^2) + dz^2) < R * R (direbm/reference/boundary.py:24-26).
__device__ __forceinline__ bool inside3(float x, float y, float z, const Inst3& p) {
  const float dx = fsub(x, p.sc[0]), dy = fsub(y, p.sc[1]), dz = fsub(z, p.sc[2]);
  return fadd(fadd(fmul(dx, dx), fmul(dy, dy)), fmul(dz, dz)) < p.sR2;
}

// Specular bounce of component k, which would land at x1 inside the obstacle, from its moment at
// x0 (direbm/reference/simulator.py:105-117): FP64 first crossing of the segment with the sphere
// and the outward normal there (boundary.py:28-47), c_k reflected about it (:56-60) and split over
// the twelve non-rest directions with weights max(d.c_j, 0)^2 / sum, parts above 1e-9 kept
// (:73-85).  Dot products are ((a0 b0 + a1 b1) + a2 b2) and the weight sum runs j = 1..12 in order.
// Writes the parts' directions and weights; returns their number (0: no crossing, the component
// is lost, as in the reference).
__device__ int bounce3(const float x0[3], const float x1[3], int k, const Inst3& p, int pj[12], double pw[12]) {
  double d[3], f[3];
  for (int a = 0; a < 3; ++a) {
    d[a] = (double)x1[a] - (double)x0[a];
    f[a] = (double)x0[a] - (double)p.sc[a];
  }
  const double R = (double)p.sR;
  const double A = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
  const double B = 2.0 * ((f[0] * d[0] + f[1] * d[1]) + f[2] * d[2]);
  const double Cc = ((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2]) - R * R;
  const double disc = B * B - 4.0 * A * Cc;
  if (A == 0.0 || disc < 0.0) return 0;
  const double sq = sqrt(disc);
  const double t0 = (-B - sq) / (2.0 * A), t1 = (-B + sq) / (2.0 * A);
  double t;
  if (0.0 <= t0 && t0 <= 1.0) t = t0;
  else if (0.0 <= t1 && t1 <= 1.0) t = t1;
  else return 0;
  double n[3];
  for (int a = 0; a < 3; ++a) n[a] = (((double)x0[a] + t * d[a]) - (double)p.sc[a]) / R;
  const double* ck = p.Cd[k];
  const double cn = (ck[0] * n[0] + ck[1] * n[1]) + ck[2] * n[2];
  double r[3];
  for (int a = 0; a < 3; ++a) r[a] = ck[a] - 2.0 * cn * n[a];
  double w[12], dots[12], sum = 0.0;
  for (int j = 0; j < 12; ++j) {
    const double* cj = p.Cd[j + 1];
    dots[j] = (cj[0] * r[0] + cj[1] * r[1]) + cj[2] * r[2];
    const double m = fmax(dots[j], 0.0);
    w[j] = m * m;
    sum += w[j];
  }
  if (sum <= 0.0) {
    int best = 0;
    for (int j = 1; j < 12; ++j)
      if (dots[j] > dots[best]) best = j;
    pj[0] = best + 1;
    pw[0] = 1.0;
    return 1;
  }
  int np = 0;
  for (int j = 0; j < 12; ++j) {
    const double q = w[j] / sum;
    if (q > 1e-9) {
      pj[np] = j + 1;
      pw[np] = q;
      ++np;
    }
  }
  return np;
}

__device__ __forceinline__ void cell_of3(float x, float y, float z, const Inst3& p, int& cx, int& cy, int& cz) {
  cx = cell_int(floorf(fdiv(x, p.cs)));
  cy = cell_int(floorf(fdiv(y, p.cs)));
  cz = cell_int(floorf(fdiv(z, p.cs)));
  if (p.periodic) {
    if (cx >= p.n) cx -= p.n;
    if (cy >= p.n) cy -= p.n;
    if (cz >= p.n) cz -= p.n;
  }
}

// Keep a cell inside the dense grid (the periodic cube, or the bounding box of every finite point).
// Only a non-finite coordinate can fall outside; it is flagged and the step fails.
__device__ __forceinline__ void clamp_cell3(const Inst3& p, int& cx, int& cy, int& cz, unsigned* status) {
  const int ax = min(max(cx, p.x0), p.x0 + p.nx - 1), ay = min(max(cy, p.y0), p.y0 + p.ny - 1),
            az = min(max(cz, p.z0), p.z0 + p.nz - 1);
  if (ax != cx || ay != cy || az != cz) atomicOr(status, ST_NONFINITE);
  cx = ax;
  cy = ay;
  cz = az;
}

__device__ __forceinline__ long long cindex3(const Inst3& p, int cx, int cy, int cz) {
  return ((long long)(cx - p.x0) * p.ny + (cy - p.y0)) * p.nz + (cz - p.z0);
}

// Visit every contiguous range of sorted items in the cells that may hold a point within r of
// (x, y, z), widened by `ext` cells: the (x, y) columns of the ball in ascending order, each
// column's z-range (unwrapped, ascending) as one range per periodic segment.  rp2 is one component
// of the int2 cell prefix, addressed with stride 2.  The bounds use the hardware square root and
// 1/cs; the margin (several ulps of the largest cell coordinate) keeps the window a superset.
template <class F>
__device__ __forceinline__ void window3(const Inst3& p, const int* __restrict__ rp2, float x, float y, float z,
                                        float r, int ext, F&& fn) {
  const float q = p.cs, iq = p.iq;
  const float rm = r + p.margin + (float)ext * q;
  const float rm2 = rm * rm;
  const bool fast = rm < p.wlim;
  int lox = (int)floorf((x - rm) * iq), hix = (int)floorf((x + rm) * iq);
  bool fullx = false;
  if (p.periodic) {
    if (hix - lox + 1 >= p.n) { lox = 0; hix = p.n - 1; fullx = true; }
  } else {
    lox = max(lox, p.x0);
    hix = min(hix, p.x0 + p.nx - 1);
  }
  for (int cx = lox; cx <= hix; ++cx) {
    float gx = 0.f;
    if (!fullx) gx = fmaxf(fmaxf((float)cx * q - x, x - (float)(cx + 1) * q), 0.f);
    if (gx > rm) continue;
    const int ccx = !p.periodic ? 0 : (fullx || !fast) ? 3 : cx < 0 ? 1 : cx >= p.n ? 2 : 0;
    const float ry = fast_sqrt(fmaxf(rm2 - gx * gx, 0.f));
    int loy = (int)floorf((y - ry) * iq), hiy = (int)floorf((y + ry) * iq);
    bool fully = false;
    if (p.periodic) {
      if (hiy - loy + 1 >= p.n) { loy = 0; hiy = p.n - 1; fully = true; }
    } else {
      loy = max(loy, p.y0);
      hiy = min(hiy, p.y0 + p.ny - 1);
    }
    int wx = cx;
    if (p.periodic) {
      wx = cx % p.n;
      if (wx < 0) wx += p.n;
    }
    for (int cy = loy; cy <= hiy; ++cy) {
      float gy = 0.f;
      if (!fully) gy = fmaxf(fmaxf((float)cy * q - y, y - (float)(cy + 1) * q), 0.f);
      const float rz2 = rm2 - gx * gx - gy * gy;
      if (rz2 < 0.f) continue;
      const float rz = fast_sqrt(rz2);
      int loz = (int)floorf((z - rz) * iq), hiz = (int)floorf((z + rz) * iq);
      if (p.periodic) {
        const int n = p.n;
        int wy = cy % n;
        if (wy < 0) wy += n;
        const int ccy = (fully || !fast) ? 3 : cy < 0 ? 1 : cy >= n ? 2 : 0;
        const int cxy = ccx | (ccy << 2);
        const bool fullz = hiz - loz + 1 >= n;
        if (fullz) { loz = 0; hiz = n - 1; }
        const int gen = (fullz || !fast) ? 3 : 0;
        const long long base = ((long long)wx * n + wy) * n;
        if (loz < 0) fn(rp2[2 * (base + loz + n)], rp2[2 * (base + min(hiz, -1) + n + 1)], cxy | ((gen | 1) << 4));
        const int a = max(loz, 0), b = min(hiz, n - 1);
        if (a <= b) fn(rp2[2 * (base + a)], rp2[2 * (base + b + 1)], cxy | (gen << 4));
        if (hiz >= n) fn(rp2[2 * (base + max(loz, n) - n)], rp2[2 * (base + hiz - n + 1)], cxy | ((gen | 2) << 4));
      } else {
        loz = max(loz, p.z0);
        hiz = min(hiz, p.z0 + p.nz - 1);
        if (loz > hiz) continue;
        const long long base = ((long long)(wx - p.x0) * p.ny + (cy - p.y0)) * p.nz - p.z0;
        fn(rp2[2 * (base + loz)], rp2[2 * (base + hiz + 1)], 0);
      }
    }
  }
}

// The window of an axis-aligned box [x0, x1] x [y0, y1] x [z0, z1] (the positions of a wave's
// lanes): every range of cells that may hold a point within r of some point of the box, widened by
// `ext` cells, in the order window3 visits them.  With wave-uniform arguments every bound is
// wave-uniform.  Periodic: the caller guarantees that no axis of the widened box covers the whole
// period (box_window_fits3), so every lane sees its candidates in the order of its own window3.
__device__ __forceinline__ bool box_window_fits3(const Inst3& p, float x0, float x1, float y0, float y1, float z0,
                                                 float z1, float r, int ext) {
  if (!p.periodic) return true;
  const float rm = r + p.margin + (float)ext * p.cs;
  return (int)floorf((x1 + rm) * p.iq) - (int)floorf((x0 - rm) * p.iq) + 1 < p.n &&
         (int)floorf((y1 + rm) * p.iq) - (int)floorf((y0 - rm) * p.iq) + 1 < p.n &&
         (int)floorf((z1 + rm) * p.iq) - (int)floorf((z0 - rm) * p.iq) + 1 < p.n;
}

template <class F>
__device__ __forceinline__ void box_window3(const Inst3& p, const int* __restrict__ rp2, float x0, float x1, float y0,
                                            float y1, float z0, float z1, float r, int ext, F&& fn) {
  const float q = p.cs, iq = p.iq;
  const float rm = r + p.margin + (float)ext * q;
  const float rm2 = rm * rm;
  const bool fast = fmaxf(fmaxf(x1 - x0, y1 - y0), z1 - z0) + rm < p.wlim;
  int lox = (int)floorf((x0 - rm) * iq), hix = (int)floorf((x1 + rm) * iq);
  if (!p.periodic) {
    lox = max(lox, p.x0);
    hix = min(hix, p.x0 + p.nx - 1);
  }
  for (int cx = lox; cx <= hix; ++cx) {
    const float gx = fmaxf(fmaxf((float)cx * q - x1, x0 - (float)(cx + 1) * q), 0.f);
    if (gx > rm) continue;
    const int ccx = !p.periodic ? 0 : !fast ? 3 : cx < 0 ? 1 : cx >= p.n ? 2 : 0;
    const float ry = fast_sqrt(fmaxf(rm2 - gx * gx, 0.f));
    int loy = (int)floorf((y0 - ry) * iq), hiy = (int)floorf((y1 + ry) * iq);
    if (!p.periodic) {
      loy = max(loy, p.y0);
      hiy = min(hiy, p.y0 + p.ny - 1);
    }
    int wx = cx;
    if (p.periodic) {
      wx = cx % p.n;
      if (wx < 0) wx += p.n;
    }
    for (int cy = loy; cy <= hiy; ++cy) {
      const float gy = fmaxf(fmaxf((float)cy * q - y1, y0 - (float)(cy + 1) * q), 0.f);
      const float rz2 = rm2 - gx * gx - gy * gy;
      if (rz2 < 0.f) continue;
      const float rz = fast_sqrt(rz2);
      int loz = (int)floorf((z0 - rz) * iq), hiz = (int)floorf((z1 + rz) * iq);
      if (p.periodic) {
        const int n = p.n;
        int wy = cy % n;
        if (wy < 0) wy += n;
        const int gen = !fast ? 3 : 0;
        const int cxy = ccx | ((gen ? 3 : cy < 0 ? 1 : cy >= n ? 2 : 0) << 2);
        const long long base = ((long long)wx * n + wy) * n;
        if (loz < 0) fn(ldu(rp2, 2 * (base + loz + n)), ldu(rp2, 2 * (base + min(hiz, -1) + n + 1)), cxy | ((gen | 1) << 4));
        const int a = max(loz, 0), b = min(hiz, n - 1);
        if (a <= b) fn(ldu(rp2, 2 * (base + a)), ldu(rp2, 2 * (base + b + 1)), cxy | (gen << 4));
        if (hiz >= n) fn(ldu(rp2, 2 * (base + max(loz, n) - n)), ldu(rp2, 2 * (base + hiz - n + 1)), cxy | ((gen | 2) << 4));
      } else {
        loz = max(loz, p.z0);
        hiz = min(hiz, p.z0 + p.nz - 1);
        if (loz > hiz) continue;
        const long long base = ((long long)(wx - p.x0) * p.ny + (cy - p.y0)) * p.nz - p.z0;
        fn(ldu(rp2, 2 * (base + loz)), ldu(rp2, 2 * (base + hiz + 1)), 0);
      }
    }
  }
}

// Bounding box of the active lanes' points, reduced over the wave (every lane gets the result).
__device__ __forceinline__ bool wave_box3(bool active, float x, float y, float z, float b[6]) {
  b[0] = active ? x : INFINITY;
  b[1] = active ? x : -INFINITY;
  b[2] = active ? y : INFINITY;
  b[3] = active ? y : -INFINITY;
  b[4] = active ? z : INFINITY;
  b[5] = active ? z : -INFINITY;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int a = 0; a < 6; a += 2) {
      b[a] = fminf(b[a], __shfl_xor(b[a], o, 64));
      b[a + 1] = fmaxf(b[a + 1], __shfl_xor(b[a + 1], o, 64));
    }
  }
#pragma unroll
  for (int a = 0; a < 6; ++a) b[a] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(b[a])));
  return b[1] >= b[0];
}

// Workgroups are dealt to the 8 XCDs round-robin.  The 3D neighbourhood kernels take their
// work from the XCD-major index instead, so that each XCD's L2 serves one contiguous run of
// points, whose windows overlap (grids are launched as a multiple of 8 blocks).
__device__ __forceinline__ int xcd_major_block() {
  return (int)(blockIdx.x & 7) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
}

struct alignas(16) Comp3 {
  float x, y, z, w;  // position; w = refine weight exp(value), then val / sw after k3_norm
};

// ---------------------------------------------------------------------------------- kernels
__device__ __forceinline__ void collide13(float f[Q3], const Inst3& p, unsigned& st) {
  float rho = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
#pragma unroll
  for (int k = 0; k < Q3; ++k) {
    rho = fadd(rho, f[k]);
    jx = fadd(jx, fmul(f[k], p.C[k][0]));
    jy = fadd(jy, fmul(f[k], p.C[k][1]));
    jz = fadd(jz, fmul(f[k], p.C[k][2]));
  }
  float ux = 0.f, uy = 0.f, uz = 0.f;
  if (rho != 0.f) {
    ux = fdiv(jx, rho);
    uy = fdiv(jy, rho);
    uz = fdiv(jz, rho);
  } else {
    st |= ST_EMPTY_RHO;
  }
  const float inv = 5.0f;  // 1 / cs^2, cs^2 = 1/5
  const float usq = fadd(fadd(fmul(ux, ux), fmul(uy, uy)), fmul(uz, uz));
  // the rest equilibrium takes the remainder rho - sum_{k>0} f_eq_k: the fp32 weights sum to
  // 1 + 1.5e-8, and the remainder keeps each collide's mass change to unbiased rounding (as collide7)
  float fe[Q3], moving = 0.f;
#pragma unroll
  for (int k = 1; k < Q3; ++k) {
    const float cu = fadd(fadd(fmul(p.C[k][0], ux), fmul(p.C[k][1], uy)), fmul(p.C[k][2], uz));
    const float a = fadd(1.0f, fmul(inv, cu));
    const float b = fmul(fmul(fmul(fmul(0.5f, inv), inv), cu), cu);
    const float c = fmul(fmul(0.5f, inv), usq);
    fe[k] = fmul(fmul(rho, p.W[k]), fsub(fadd(a, b), c));
    moving = fadd(moving, fe[k]);
  }
  fe[0] = fsub(rho, moving);
#pragma unroll
  for (int k = 0; k < Q3; ++k) f[k] = fadd(f[k], fdiv(fsub(fe[k], f[k]), p.tau));
}

// Merge the block's boxes into box (every thread of the block calls it): wave reductions, then
// the waves' boxes in LDS, then atomics only where the box grows.  Blocks over cells in x-major
// order widen hi[0] one after the other, and atomics on one address serialise.
__device__ __forceinline__ void box3_flush_block(int lo[3], int hi[3], BBox3* box) {
  __shared__ int sl[TB / 64][3], sh[TB / 64][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lo[d] = min(lo[d], __shfl_xor(lo[d], o, 64));
      hi[d] = max(hi[d], __shfl_xor(hi[d], o, 64));
    }
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 3; ++d) {
      sl[wv][d] = lo[d];
      sh[wv][d] = hi[d];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q)
      for (int d = 0; d < 3; ++d) {
        lo[d] = min(lo[d], sl[q][d]);
        hi[d] = max(hi[d], sh[q][d]);
      }
    if (lo[0] != INT_MAX) {
      const BBox3 cur = *box;
      for (int d = 0; d < 3; ++d) {
        if (lo[d] < cur.lo[d]) atomicMin(&box->lo[d], lo[d]);
        if (hi[d] > cur.hi[d]) atomicMax(&box->hi[d], hi[d]);
      }
    }
  }
}

__global__ void k3_box_init(BBox3* box, int* disp) {
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; ++d) {
      box->lo[d] = INT_MAX;
      box->hi[d] = INT_MIN;
    }
    *disp = 0;
  }
}

// Collide in place (do_collide) and/or widen the bounding box of cells by the moments (do_box).
__global__ __launch_bounds__(TB) void k3_collide_box(float* __restrict__ f, long long pitch, const float* __restrict__ mx,
                                                     const float* __restrict__ my, const float* __restrict__ mz, int m,
                                                     const Inst3* __restrict__ ins, BBox3* box, unsigned* status,
                                                     int do_collide, int do_box) {
  const Inst3& p = *ins;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned st = 0;
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  if (i < m) {
    if (do_collide) {
      float v[Q3];
#pragma unroll
      for (int k = 0; k < Q3; ++k) v[k] = f[k * pitch + i];
      collide13(v, p, st);
#pragma unroll
      for (int k = 0; k < Q3; ++k) {
        if (!isfinite(v[k])) st |= ST_NONFINITE;
        f[k * pitch + i] = v[k];
      }
    }
    const float x = mx[i], y = my[i], z = mz[i];
    if (!isfinite(x) || !isfinite(y) || !isfinite(z)) st |= ST_NONFINITE;
    else if (do_box && p.periodic && !(x >= 0.f && x < p.L && y >= 0.f && y < p.L && z >= 0.f && z < p.L))
      st |= ST_RANGE;
    else if (do_box && !p.periodic) {
      int c[3];
      cell_of3(x, y, z, p, c[0], c[1], c[2]);
      const int e = (int)ceilf(p.dx / p.cs) + 2;  // components lie within dx of their moment
      for (int d = 0; d < 3; ++d) {
        lo[d] = c[d] - e;
        hi[d] = c[d] + e;
      }
    }
  }
  if (st) atomicOr(status, st);
  if (do_box) box3_flush_block(lo, hi, box);  // do_box is uniform: every thread calls it
}

__global__ __launch_bounds__(TB) void k3_keys(const float* __restrict__ mx, const float* __restrict__ my,
                                              const float* __restrict__ mz, int m, const Inst3* __restrict__ ins,
                                              uint32_t* __restrict__ keys, unsigned* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const Inst3& p = *ins;
  for (int k = 0; k < Q3; ++k) {
    float x, y, z;
    comp_pos3(mx[i], my[i], mz[i], k, p, x, y, z);
    if (p.sphere && k && inside3(x, y, z, p)) {  // bounced: its parts are keyed by k3_bounce_emit
      keys[(size_t)Q3 * i + k] = (uint32_t)((long long)p.nx * p.ny * p.nz);
      continue;
    }
    int cx, cy, cz;
    cell_of3(x, y, z, p, cx, cy, cz);
    clamp_cell3(p, cx, cy, cz, status);
    keys[(size_t)Q3 * i + k] = (uint32_t)cindex3(p, cx, cy, cz);
  }
}

// Number of bounce parts of each moment (obstacle only), as (count, 0) for the int2 scan.
__global__ __launch_bounds__(TB) void k3_bounce_count(const float* __restrict__ mx, const float* __restrict__ my,
                                                      const float* __restrict__ mz, int m,
                                                      const Inst3* __restrict__ ins, int2* __restrict__ bcnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const Inst3& p = *ins;
  const float x0[3] = {mx[i], my[i], mz[i]};
  int cnt = 0;
  for (int k = 1; k < Q3; ++k) {
    float x1[3];
    comp_pos3(x0[0], x0[1], x0[2], k, p, x1[0], x1[1], x1[2]);
    if (!inside3(x1[0], x1[1], x1[2], p)) continue;
    int pj[12];
    double pw[12];
    cnt += bounce3(x0, x1, k, p, pj, pw);
  }
  bcnt[i] = make_int2(cnt, 0);
}

// Bounce parts as explicit components after the N implicit ones: record (position, value),
// direction, and key; component index N + offset, so that within a cell the parts follow the
// regular components in (moment, direction, part) order.  A part whose position is inside the
// obstacle stays at its moment's position (direbm/reference/simulator.py:113-115).
__global__ __launch_bounds__(TB) void k3_bounce_emit(const float* __restrict__ mx, const float* __restrict__ my,
                                                     const float* __restrict__ mz, const float* __restrict__ f,
                                                     long long pitch, int m, const Inst3* __restrict__ ins,
                                                     const int2* __restrict__ boff, int N, uint32_t* __restrict__ keys,
                                                     float4* __restrict__ ex, unsigned char* __restrict__ exdir,
                                                     unsigned* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int o = boff[i].x;
  if (boff[i + 1].x == o) return;
  const Inst3& p = *ins;
  const float x0[3] = {mx[i], my[i], mz[i]};
  for (int k = 1; k < Q3; ++k) {
    float x1[3];
    comp_pos3(x0[0], x0[1], x0[2], k, p, x1[0], x1[1], x1[2]);
    if (!inside3(x1[0], x1[1], x1[2], p)) continue;
    int pj[12];
    double pw[12];
    const int np = bounce3(x0, x1, k, p, pj, pw);
    const double v = (double)f[k * pitch + i];
    for (int a = 0; a < np; ++a) {
      float q[3];
      comp_pos3(x0[0], x0[1], x0[2], pj[a], p, q[0], q[1], q[2]);
      if (inside3(q[0], q[1], q[2], p)) {
        q[0] = x0[0];
        q[1] = x0[1];
        q[2] = x0[2];
        if (p.periodic) {
          q[0] = wrap3(q[0], p);
          q[1] = wrap3(q[1], p);
          q[2] = wrap3(q[2], p);
        }
      }
      int cx, cy, cz;
      cell_of3(q[0], q[1], q[2], p, cx, cy, cz);
      if (!p.periodic) {
        const int ax = min(max(cx, p.x0), p.x0 + p.nx - 1), ay = min(max(cy, p.y0), p.y0 + p.ny - 1),
                  az = min(max(cz, p.z0), p.z0 + p.nz - 1);
        if (ax != cx || ay != cy || az != cz) atomicOr(status, ST_NONFINITE);
        cx = ax;
        cy = ay;
        cz = az;
      }
      keys[(size_t)N + o] = (uint32_t)cindex3(p, cx, cy, cz);
      ex[o] = make_float4(q[0], q[1], q[2], (float)(v * pw[a]));
      exdir[o] = (unsigned char)pj[a];
      ++o;
    }
  }
}

// Records of the sorted components; the first component of each cell counts the cell's components
// (and its one control point).
__global__ __launch_bounds__(TB) void k3_mark(const uint32_t* __restrict__ K, const uint32_t* __restrict__ V, int n,
                                              const float* __restrict__ mx, const float* __restrict__ my,
                                              const float* __restrict__ mz, const float* __restrict__ f,
                                              long long pitch, const Inst3* __restrict__ ins, Comp3* __restrict__ comp,
                                              unsigned char* __restrict__ cdir, float* __restrict__ cv,
                                              int* __restrict__ cnt2, int N, const float4* __restrict__ ex,
                                              const unsigned char* __restrict__ exdir, uint32_t ncells) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, base = s - lane;
  if (base >= n) return;  // the whole wave is past the end
  // bounced components (key ncells) sort after every cell and belong to none
  const uint32_t q = s < n ? K[s] : 0xffffffffu;
  const bool active = q < ncells;
  // Cell counts from ballots, as in 2D k_mark: a cell ends at the next cell start within the
  // wave; the wave's last cell is followed through the next 64-key chunks by the whole wave.
  uint32_t prev = __shfl_up(q, 1, 64);
  if (lane == 0) prev = base > 0 ? K[base - 1] : ~q;
  const bool first = active && q != prev;
  const unsigned long long fm = __ballot(first), live = __ballot(active);  // wave-uniform
  if (fm) {
    const int lf = 63 - __clzll((long long)fm);
    int tail = 0;
    if (base + 64 < n) {
      const uint32_t tq = __shfl(q, lf, 64);
      for (int t0 = base + 64; t0 < n; t0 += 64) {
        const int t = t0 + lane;
        const int len = __popcll(__ballot(t < n && K[t] == tq));
        tail += len;
        if (len < 64) break;
      }
    }
    if (first) {
      const unsigned long long above = lane == 63 ? 0ull : fm & ~((2ull << lane) - 1ull);
      int e = above ? __ffsll((long long)above) - 1 : 64;
      // the cell ends before the first bounced key or the end of the components
      const unsigned long long from = ~((1ull << lane) - 1ull);
      const unsigned long long dead_after = ~live & from;
      if (dead_after) e = min(e, __ffsll((long long)dead_after) - 1);
      int nc = e - lane;
      if (lane == lf && e == 64) nc += tail;
      cnt2[2 * (size_t)q] = nc;
      cnt2[2 * (size_t)q + 1] = 1;
    }
  }
  if (!active) return;
  const Inst3& p = *ins;
  const uint32_t c = V[s];
  Comp3 r;
  float v;
  int k;
  if (c < (uint32_t)N) {
    const uint32_t i = c / Q3;
    k = (int)(c - Q3 * i);
    comp_pos3(mx[i], my[i], mz[i], k, p, r.x, r.y, r.z);
    v = f[k * pitch + i];
  } else {
    const float4 e = ex[c - N];
    r.x = e.x;
    r.y = e.y;
    r.z = e.z;
    v = e.w;
    k = exdir[c - N];
  }
  r.w = expf(v);  // refine weight exp(f) (direbm/warp/propagation.py:100)
  comp[s] = r;
  cdir[s] = (unsigned char)k;
  cv[s] = v;
}

// Control point of every occupied cell: the fp32 centroid of its members in member order.
__global__ __launch_bounds__(TB) void k3_centroid(const int2* __restrict__ rp, long long ncells,
                                                  const Comp3* __restrict__ comp, float* __restrict__ cx,
                                                  float* __restrict__ cy, float* __restrict__ cz,
                                                  int* __restrict__ cpq) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncells) return;
  const int s0 = rp[q].x, s1 = rp[q + 1].x;
  if (s0 == s1) return;
  float ax = 0.f, ay = 0.f, az = 0.f;
  for (int s = s0; s < s1; ++s) {
    const Comp3 c = comp[s];
    ax = fadd(ax, c.x);
    ay = fadd(ay, c.y);
    az = fadd(az, c.z);
  }
  const float n = (float)(s1 - s0);
  const int r = rp[q].y;
  cx[r] = fdiv(ax, n);
  cy[r] = fdiv(ay, n);
  cz[r] = fdiv(az, n);
  cpq[r] = (int)q;
}

// Refine (direbm/warp/propagation.py:88-112 in 3D): kappa from the 13-bit direction mask of the
// components within dx; kappa <= kappa_hard keeps the point, else the exp(val)-weighted centroid.
// A group of G lanes per point strides through each window range; a fixed butterfly combines the
// lanes.  Also the Chebyshev cell displacement bound for the control-point windows.
template <int PM, int G>
__global__ __launch_bounds__(TB) void k3_refine(const int* __restrict__ Ptot, const float* __restrict__ x0s,
                                                const float* __restrict__ y0s, const float* __restrict__ z0s,
                                                const int* __restrict__ cpq, const Inst3* __restrict__ ins,
                                                const int* __restrict__ rp2, const Comp3* __restrict__ comp,
                                                const unsigned char* __restrict__ cdir, float* __restrict__ ox,
                                                float* __restrict__ oy, float* __restrict__ oz,
                                                unsigned char* __restrict__ kap, int* __restrict__ disp) {
  const Inst3& p = *ins;
  const int lane = threadIdx.x & (G - 1);
  const int P = *Ptot;
  const int ngroups = gridDim.x * (blockDim.x / G);
  const int rlast = ((P + ngroups - 1) / ngroups) * ngroups;
  const int per = PM == 2 ? p.periodic : PM;
  for (int r = (xcd_major_block() * blockDim.x + threadIdx.x) / G; r < rlast; r += ngroups) {
    const bool active = r < P;
    const float x0 = active ? x0s[r] : 0.f, y0 = active ? y0s[r] : 0.f, z0 = active ? z0s[r] : 0.f;
    unsigned mask = 0;
    float sw = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
    if (active)
      window3(p, rp2, x0, y0, z0, p.dx, 0, [&](int a, int e, int wc) {
        auto run = [&](auto plain) {
          auto take = [&](const Comp3& c, int k, bool in) {
            float dx = fsub(c.x, x0), dy = fsub(c.y, y0), dz = fsub(c.z, z0);
            if (!decltype(plain)::value) {
              dx = wrapd<PM>(dx, wc & 3, p);
              dy = wrapd<PM>(dy, (wc >> 2) & 3, p);
              dz = wrapd<PM>(dz, wc >> 4, p);
            }
            // branch-free: a rejected candidate adds weight +0, which leaves the sums' bits as
            // they are (no sum is ever -0)
            const bool ok = in && fadd(fadd(fmul(dx, dx), fmul(dy, dy)), fmul(dz, dz)) <= p.t_dx;
            const float w = ok ? c.w : 0.f;
            mask |= (ok ? 1u : 0u) << (k & 31);  // k < 13: one select and one shift-or
            // the weighted sums are continuous quantities: fused multiply-adds (one rounding each)
            sw = fadd(sw, w);
            sx = __builtin_fmaf(w, per ? dx : c.x, sx);
            sy = __builtin_fmaf(w, per ? dy : c.y, sy);
            sz = __builtin_fmaf(w, per ? dz : c.z, sz);
          };
          // kInflight candidates per lane in flight: their records are loaded before any is used, and
          // they are added in ascending index order as one at a time would be.  Full passes need no
          // range test or clamp; the last, partial pass clamps its loads and masks the candidates
          // past the range's end.
          int s0 = a + lane;
          for (; s0 + (kInflight - 1) * G < e; s0 += kInflight * G) {
            Comp3 cc[kInflight];
            int kk[kInflight];
#pragma unroll
            for (int u = 0; u < kInflight; ++u) {
              cc[u] = comp[s0 + u * G];
              kk[u] = cdir[s0 + u * G];
            }
#pragma unroll
            for (int u = 0; u < kInflight; ++u) take(cc[u], kk[u], true);
          }
          if (s0 < e) {
            Comp3 cc[kInflight];
            int kk[kInflight];
#pragma unroll
            for (int u = 0; u < kInflight; ++u) {
              const int s = min(s0 + u * G, e - 1);
              cc[u] = comp[s];
              kk[u] = cdir[s];
            }
#pragma unroll
            for (int u = 0; u < kInflight; ++u) take(cc[u], kk[u], s0 + u * G < e);
          }
        };
        if (wc == 0) run(std::true_type{});
        else run(std::false_type{});
      });
    mask = group_or<G>(mask);
    sw = group_sum<G>(sw);
    sx = group_sum<G>(sx);
    sy = group_sum<G>(sy);
    sz = group_sum<G>(sz);
    if (!active || lane != 0) continue;
    const int kappa = __popc(mask);
    float nx = x0, ny = y0, nz = z0;
    if (kappa > p.kappa_hard) {
      if (per) {
        nx = wrap3(fadd(x0, fdiv(sx, sw)), p);
        ny = wrap3(fadd(y0, fdiv(sy, sw)), p);
        nz = wrap3(fadd(z0, fdiv(sz, sw)), p);
      } else {
        nx = fdiv(sx, sw);
        ny = fdiv(sy, sw);
        nz = fdiv(sz, sw);
      }
    }
    ox[r] = nx;
    oy[r] = ny;
    oz[r] = nz;
    kap[r] = (unsigned char)kappa;
    // Chebyshev distance in cells between the refined position and the registered cell
    long long rel = cpq[r];
    int rc[3];
    rc[2] = (int)(rel % p.nz) + p.z0;
    rel /= p.nz;
    rc[1] = (int)(rel % p.ny) + p.y0;
    rc[0] = (int)(rel / p.ny) + p.x0;
    int c[3];
    cell_of3(nx, ny, nz, p, c[0], c[1], c[2]);
    int dd = 0;
    for (int d = 0; d < 3; ++d) {
      int v = abs(c[d] - rc[d]);
      if (per) v = min(v, p.n - v);
      dd = max(dd, v);
    }
    // a refined point moves at most dx (a weighted mean of components within dx); a non-finite
    // mean (overflowing weights) is left to k3_cpkeys, which flags it
    dd = min(dd, (int)(p.dx / p.cs) + 2);
    if (isfinite(nx) && isfinite(ny) && isfinite(nz) && dd > 0 && dd > *disp) atomicMax(disp, dd);
  }
}

// Scatter normaliser per component: sw = sum over control points within r of (r - d); q = v / sw.
// One lane per component.  When the wave's components span a compact box, the wave walks the
// window of that box with wave-uniform bounds (scalar loads of each candidate, no divergence) and
// every lane tests every candidate; otherwise each lane walks its own window.  Either way a lane
// adds its candidates in the order of its own window.
// The refined control points are binned by the cell of their refined position (k3_cpkeys, sort,
// k3_cpbins): spos holds them in bin order as (x, y, z, index), rpb is the bins' prefix (stride 2),
// and a window of radius r around a point holds every control point within r of it.
// q = val / sw replaces the record's refine weight (comp[s].w; refine has run), so the gather
// reads position and q in one 16-byte load.
template <int PM>
__global__ __launch_bounds__(TB) void k3_norm(int n, Comp3* __restrict__ comp, const float* __restrict__ cv,
                                              const Inst3* __restrict__ ins, const int* __restrict__ rpb,
                                              const float4* __restrict__ spos, const int* __restrict__ nvalid,
                                              float* __restrict__ swc) {
  n = min(n, *nvalid);  // bounced components sort last and are not part of any cell
  const int s = xcd_major_block() * blockDim.x + threadIdx.x;
  const bool active = s < n;
  const Inst3& p = *ins;
  Comp3 c{0.f, 0.f, 0.f, 0.f};
  if (active) c = comp[s];
  float b[6];
  if (!wave_box3(active, c.x, c.y, c.z, b)) return;  // the whole wave is past the end
  const float rr = p.rr, t2 = p.t_rr;
  const float lim = 3.f * (rr + p.cs);
  float sw = 0.f;
  if (b[1] - b[0] <= lim && b[3] - b[2] <= lim && b[5] - b[4] <= lim &&
      box_window_fits3(p, b[0], b[1], b[2], b[3], b[4], b[5], rr, 0)) {
    box_window3(p, rpb, b[0], b[1], b[2], b[3], b[4], b[5], rr, 0, [&](int a, int e, int wc) {
      const int ua = __builtin_amdgcn_readfirstlane(a), ue = __builtin_amdgcn_readfirstlane(e);
      const int uw = __builtin_amdgcn_readfirstlane(wc);
      auto run = [&](auto plain) {
#pragma unroll 4
        for (int j = ua; j < ue; ++j) {
          const float4 q = spos[j];
          const float s2 = d2w<PM, decltype(plain)::value>(q.x, q.y, q.z, c.x, c.y, c.z, uw, p);
          if (s2 <= t2) sw = fadd(sw, fsub(rr, fast_sqrt(s2)));
        }
      };
      if (uw == 0) run(std::true_type{});
      else run(std::false_type{});
    });
  } else if (active) {
    window3(p, rpb, c.x, c.y, c.z, rr, 0, [&](int a, int e, int wc) {
      auto run = [&](auto plain) {
        for (int j = a; j < e; ++j) {
          const float4 q = spos[j];
          const float s2 = d2w<PM, decltype(plain)::value>(q.x, q.y, q.z, c.x, c.y, c.z, wc, p);
          if (s2 <= t2) sw = fadd(sw, fsub(rr, fast_sqrt(s2)));
        }
      };
      if (wc == 0) run(std::true_type{});
      else run(std::false_type{});
    });
  }
  if (!active) return;
  swc[s] = sw;
  comp[s].w = sw > 0.f ? fdiv(cv[s], sw) : 0.f;
}

// Scatter as a gather: each control point sums (r - d) q over components within r, per direction
// (q = val / sw in the record's weight slot, written by k3_norm).
template <int PM, int G>
__global__ __launch_bounds__(TB) void k3_gather(const int* __restrict__ Ptot, const float* __restrict__ cpx,
                                                const float* __restrict__ cpy, const float* __restrict__ cpz,
                                                const Inst3* __restrict__ ins, const int* __restrict__ rp2,
                                                const Comp3* __restrict__ comp, const unsigned char* __restrict__ cdir,
                                                float* __restrict__ cpf, long long pitch,
                                                unsigned short* __restrict__ emask) {
  const Inst3& p = *ins;
  const int lane = threadIdx.x & (G - 1);
  const int P = *Ptot;
  const int ngroups = gridDim.x * (blockDim.x / G);
  const int rlast = ((P + ngroups - 1) / ngroups) * ngroups;
  const float rr = p.rr;
  // Each lane's thirteen direction sums live in LDS (slot k * TB + thread: bank-conflict free
  // whatever the directions), so a candidate costs one read-add-write of its own direction.
  __shared__ float sacc[Q3 * TB];
  float* my = sacc + threadIdx.x;
  for (int r = (xcd_major_block() * blockDim.x + threadIdx.x) / G; r < rlast; r += ngroups) {
    const bool active = r < P;
    const float x = active ? cpx[r] : 0.f, y = active ? cpy[r] : 0.f, z = active ? cpz[r] : 0.f;
#pragma unroll
    for (int k = 0; k < Q3; ++k) my[k * TB] = 0.f;
    if (active)
      window3(p, rp2, x, y, z, rr, 0, [&](int a, int e, int wc) {
        auto run = [&](auto plain) {
          auto take = [&](const Comp3& c, int k, bool in) {
            const float s2 = d2w<PM, decltype(plain)::value>(c.x, c.y, c.z, x, y, z, wc, p);
            // branch-free (measured faster here than a branch): a rejected candidate adds a zero
            // product (r - d) * 0 to its direction's slot in one fused multiply-add, which leaves
            // the slot's bits as they are (no slot is ever -0)
            const bool ok = in && s2 <= p.t_rr;
            float* slot = my + k * TB;
            *slot = __builtin_fmaf(fsub(rr, fast_sqrt(s2)), ok ? c.w : 0.f, *slot);
          };
          // as in k3_refine: full passes unclamped, the last partial pass clamped and masked
          int s0 = a + lane;
          for (; s0 + (kInflight - 1) * G < e; s0 += kInflight * G) {
            Comp3 cc[kInflight];
            int kk[kInflight];
#pragma unroll
            for (int u = 0; u < kInflight; ++u) {
              cc[u] = comp[s0 + u * G];
              kk[u] = cdir[s0 + u * G];
            }
#pragma unroll
            for (int u = 0; u < kInflight; ++u) take(cc[u], kk[u], true);
          }
          if (s0 < e) {
            Comp3 cc[kInflight];
            int kk[kInflight];
#pragma unroll
            for (int u = 0; u < kInflight; ++u) {
              const int s = min(s0 + u * G, e - 1);
              cc[u] = comp[s];
              kk[u] = cdir[s];
            }
#pragma unroll
            for (int u = 0; u < kInflight; ++u) take(cc[u], kk[u], s0 + u * G < e);
          }
        };
        if (wc == 0) run(std::true_type{});
        else run(std::false_type{});
      });
    float acc[Q3];
#pragma unroll
    for (int k = 0; k < Q3; ++k) acc[k] = group_sum<G>(my[k * TB]);
    if (!active || lane != 0) continue;
    unsigned em = 0;
#pragma unroll
    for (int k = 0; k < Q3; ++k) {
      cpf[k * pitch + r] = acc[k];
      if (acc[k] < p.eps) em |= 1u << k;
    }
    emask[r] = (unsigned short)em;
  }
}

// Cell of each refined control point (dense grid index); entries past the live count are not read.
__global__ __launch_bounds__(TB) void k3_cpkeys(const int* __restrict__ Ptot, const float* __restrict__ cpx,
                                                const float* __restrict__ cpy, const float* __restrict__ cpz,
                                                const Inst3* __restrict__ ins, uint32_t* __restrict__ keys,
                                                unsigned* status) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= *Ptot) return;
  const Inst3& p = *ins;
  int cx, cy, cz;
  cell_of3(cpx[r], cpy[r], cpz[r], p, cx, cy, cz);
  // a non-finite refined mean (overflowing weights) gets a cell of the grid; the step then fails
  if (!isfinite(cpx[r]) || !isfinite(cpy[r]) || !isfinite(cpz[r])) atomicOr(status, ST_NONFINITE);
  clamp_cell3(p, cx, cy, cz, status);
  keys[r] = (uint32_t)cindex3(p, cx, cy, cz);
}

// Sorted refined control points: (x, y, z, index) records in bin order, and per-cell counts
// (the first entry of each run counts the run) as (count, 0) for the int2 scan.
__global__ __launch_bounds__(TB) void k3_cpbins(const int* __restrict__ Ptot, const uint32_t* __restrict__ K,
                                                const uint32_t* __restrict__ V, const float* __restrict__ cpx,
                                                const float* __restrict__ cpy, const float* __restrict__ cpz,
                                                float4* __restrict__ spos, int2* __restrict__ cnt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int P = *Ptot;
  if (s >= P) return;
  const uint32_t r = V[s];
  spos[s] = make_float4(cpx[r], cpy[r], cpz[r], __int_as_float((int)r));
  const uint32_t q = K[s];
  if (s == 0 || K[s - 1] != q) {
    int n = 0;
    for (int t = s; t < P && K[t] == q; ++t) ++n;
    cnt[q] = make_int2(n, 0);
  }
}

// Empty-direction masks of the control points in bin order.
__global__ __launch_bounds__(TB) void k3_sem(const int* __restrict__ Ptot, const float4* __restrict__ spos,
                                             const unsigned short* __restrict__ emask,
                                             unsigned short* __restrict__ sem) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= *Ptot) return;
  sem[s] = emask[__float_as_int(spos[s].w)];
}

// Control points that stay after resampling: those outside the obstacle (direbm/reference/
// simulator.py:79-80), as (flag, 0) over the whole capacity for the int2 scan.
__global__ __launch_bounds__(TB) void k3_keep(const int* __restrict__ Ptot, const float* __restrict__ cpx,
                                              const float* __restrict__ cpy, const float* __restrict__ cpz,
                                              long long pcap, const Inst3* __restrict__ ins, int2* __restrict__ kc) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= pcap) return;
  const int P = *Ptot;
  kc[r] = make_int2(r < P && !inside3(cpx[r], cpy[r], cpz[r], *ins) ? 1 : 0, 0);
}

// Gap-fill (empty direction i becomes (r / sw) f_rest[i] over also-empty control points within r,
// self included) and emission of the new moments (SoA) and the next bounding box.
template <int PM>
__global__ __launch_bounds__(TB) void k3_gapfill_emit(const int* __restrict__ Ptot, const float* __restrict__ cpx,
                                                      const float* __restrict__ cpy, const float* __restrict__ cpz,
                                                      const Inst3* __restrict__ ins, const int* __restrict__ rpb,
                                                      const float4* __restrict__ spos,
                                                      const unsigned short* __restrict__ sem,
                                                      const float* __restrict__ cpf,
                                                      long long pitch, const unsigned short* __restrict__ emask,
                                                      float* __restrict__ ox, float* __restrict__ oy,
                                                      float* __restrict__ oz, float* __restrict__ of, long long opitch,
                                                      const int2* __restrict__ koff, BBox3* box, unsigned* status) {
  const Inst3& p = *ins;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int P = *Ptot;
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  // With an obstacle, the kept points are emitted in order at their prefix index.
  int o = r;
  bool kept = r < P;
  if (koff && kept) {
    o = koff[r].x;
    kept = koff[r + 1].x > o;
  }
  // The gap-fill sums of the points with an empty direction: over the wave's box window when
  // those points span a compact box (wave-uniform walk), else over each point's own window.
  const float x = kept ? cpx[r] : 0.f, y = kept ? cpy[r] : 0.f, z = kept ? cpz[r] : 0.f, rr = p.rr;
  const unsigned e = kept ? emask[r] : 0u;
  float sw[Q3];
#pragma unroll
  for (int k = 0; k < Q3; ++k) sw[k] = 0.f;
  auto visit = [&](int j, int wc) {
    const unsigned m = e & sem[j];
    if (m) {
      const float4 q = spos[j];
      const float s2 = d2w<PM, false>(q.x, q.y, q.z, x, y, z, wc, p);
      if (s2 <= p.t_rr) {
        const float w = fsub(rr, fast_sqrt(s2));
#pragma unroll
        for (int k = 0; k < Q3; ++k)
          if ((m >> k) & 1u) sw[k] = fadd(sw[k], w);
      }
    }
  };
  float bb[6];
  if (wave_box3(e != 0u, x, y, z, bb)) {
    const float lim = 3.f * (rr + p.cs);
    if (bb[1] - bb[0] <= lim && bb[3] - bb[2] <= lim && bb[5] - bb[4] <= lim &&
        box_window_fits3(p, bb[0], bb[1], bb[2], bb[3], bb[4], bb[5], rr, 0)) {
      box_window3(p, rpb, bb[0], bb[1], bb[2], bb[3], bb[4], bb[5], rr, 0, [&](int a, int en, int wc) {
        const int ua = __builtin_amdgcn_readfirstlane(a), ue = __builtin_amdgcn_readfirstlane(en);
        const int uw = __builtin_amdgcn_readfirstlane(wc);
        for (int j = ua; j < ue; ++j) visit(j, uw);
      });
    } else if (e) {
      window3(p, rpb, x, y, z, rr, 0, [&](int a, int en, int wc) {
        for (int j = a; j < en; ++j) visit(j, wc);
      });
    }
  }
  if (kept) {
    float out[Q3];
#pragma unroll
    for (int k = 0; k < Q3; ++k) out[k] = cpf[k * pitch + r];
#pragma unroll
    for (int k = 0; k < Q3; ++k)
      if (((e >> k) & 1u) && sw[k] > 0.f) out[k] = fmul(fdiv(rr, sw[k]), p.frest[k]);
    ox[o] = x;
    oy[o] = y;
    oz[o] = z;
#pragma unroll
    for (int k = 0; k < Q3; ++k) of[k * opitch + o] = out[k];
    if (!isfinite(x) || !isfinite(y) || !isfinite(z)) {
      atomicOr(status, ST_NONFINITE);
    } else if (!p.periodic) {
      int c[3];
      cell_of3(x, y, z, p, c[0], c[1], c[2]);
      const int e2 = (int)ceilf(p.dx / p.cs) + 2;
      for (int d = 0; d < 3; ++d) {
        lo[d] = c[d] - e2;
        hi[d] = c[d] + e2;
      }
    }
  }
  box3_flush_block(lo, hi, box);
}

__global__ __launch_bounds__(TB) void k3_pack(const float* __restrict__ mx, const float* __restrict__ my,
                                              const float* __restrict__ mz, const float* __restrict__ f, long long pitch,
                                              int m, float* __restrict__ pos, float* __restrict__ fo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  pos[3 * (size_t)i] = mx[i];
  pos[3 * (size_t)i + 1] = my[i];
  pos[3 * (size_t)i + 2] = mz[i];
  for (int k = 0; k < Q3; ++k) fo[Q3 * (size_t)i + k] = f[k * pitch + i];
}

__global__ __launch_bounds__(TB) void k3_unpack(const float* __restrict__ pos, const float* __restrict__ fi, int m,
                                                float* __restrict__ mx, float* __restrict__ my, float* __restrict__ mz,
                                                float* __restrict__ f, long long pitch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  mx[i] = pos[3 * (size_t)i];
  my[i] = pos[3 * (size_t)i + 1];
  mz[i] = pos[3 * (size_t)i + 2];
  for (int k = 0; k < Q3; ++k) f[k * pitch + i] = fi[Q3 * (size_t)i + k];
}

// FP64 mass and momentum (fixed-order block tree, then the blocks in order).
constexpr int kDiag3Blocks = 256;
__global__ __launch_bounds__(TB) void k3_diag(const float* __restrict__ f, long long pitch, int m,
                                              const Inst3* __restrict__ ins, double* __restrict__ part) {
  __shared__ double sh[4][TB];
  const Inst3& p = *ins;
  const int t = threadIdx.x;
  double a[4] = {0, 0, 0, 0};
  for (long long i = (long long)blockIdx.x * TB + t; i < m; i += (long long)gridDim.x * TB)
    for (int k = 0; k < Q3; ++k) {
      const double v = f[k * pitch + i];
      a[0] += v;
      a[1] += v * (double)p.C[k][0];
      a[2] += v * (double)p.C[k][1];
      a[3] += v * (double)p.C[k][2];
    }
  for (int q = 0; q < 4; ++q) sh[q][t] = a[q];
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int q = 0; q < 4; ++q) sh[q][t] += sh[q][t + w];
    __syncthreads();
  }
  if (t == 0)
    for (int q = 0; q < 4; ++q) part[(size_t)blockIdx.x * 4 + q] = sh[q][0];
}

// The D3Q13 icosahedral set in FP64 (direbm/lattices.py:44-61): rest, then the twelve
// normalised (0, +-1, +-phi) cyclic vertices in the reference's order.
static void d3q13_fp64(double C[Q3][3], double W[Q3]) {
  const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
  const double v[12][3] = {{0, 1, phi},  {0, 1, -phi},  {0, -1, phi},  {0, -1, -phi}, {1, phi, 0},  {1, -phi, 0},
                           {-1, phi, 0}, {-1, -phi, 0}, {phi, 0, 1},  {phi, 0, -1},  {-phi, 0, 1}, {-phi, 0, -1}};
  const double norm = std::sqrt(0.0 * 0.0 + 1.0 * 1.0 + phi * phi);
  for (int d = 0; d < 3; ++d) C[0][d] = 0.0;
  for (int i = 0; i < 12; ++i)
    for (int d = 0; d < 3; ++d) C[i + 1][d] = v[i][d] / norm;
  W[0] = 2.0 / 5.0;
  for (int i = 1; i < Q3; ++i) W[i] = 1.0 / 20.0;
}

}  // namespace drb

using namespace drb;

// Launch a 3D neighbourhood kernel instantiated on the periodic mode (0 open, 1 periodic cube) and,
// for the point-group kernels, on the lanes per point (4, 8 or 16).
#define DRB3_LAUNCH_PM(kern, pm, grid, block, shm, stream, ...)                          \
  do {                                                                                 \
    if (pm) hipLaunchKernelGGL((kern<1>), grid, block, shm, stream, __VA_ARGS__);      \
    else hipLaunchKernelGGL((kern<0>), grid, block, shm, stream, __VA_ARGS__);         \
  } while (0)
#define DRB3_LAUNCH_G(kern, pm, G, grid, block, shm, stream, ...)                        \
  do {                                                                                 \
    if (pm) hipLaunchKernelGGL((kern<1, G>), grid, block, shm, stream, __VA_ARGS__);   \
    else hipLaunchKernelGGL((kern<0, G>), grid, block, shm, stream, __VA_ARGS__);      \
  } while (0)
#define DRB3_LAUNCH_PMG(kern, pm, g, grid, block, shm, stream, ...)                         \
  do {                                                                                    \
    if ((g) == 8) DRB3_LAUNCH_G(kern, pm, 8, grid, block, shm, stream, __VA_ARGS__);       \
    else if ((g) == 16) DRB3_LAUNCH_G(kern, pm, 16, grid, block, shm, stream, __VA_ARGS__); \
    else DRB3_LAUNCH_G(kern, pm, 4, grid, block, shm, stream, __VA_ARGS__);                \
  } while (0)

struct drb_sim3 {
  hipStream_t s = nullptr;
  Inst3 h{};
  Buf d_inst, mx[2], my[2], mz[2], mf[2];
  long long mcap[2] = {0, 0};
  int cur = 0;
  long long M = 0;
  bool have_state = false;
  long long ncells = 0, cpcap = 0;
  Buf keys, vals, comp, cdir, cv, swc, cells, rp, cp0x, cp0y, cp0z, cpq, cpx, cpy, cpz, cpf, emask, kap;
  Buf box, disp, status, diag;
  Buf bcnt, boff, ex, exdir, kc, koff;  // obstacle: bounce parts and the kept control points
  Buf keys2, vals2, cellsb, rpb, spos, sem;  // refined control points binned by cell
  RadixSort sorter;
  Scan2 scan, scan_b, scan_k;
  int g_refine = 4, g_gather = 4;  // lanes per control point of refine and gather
  void* hpin = nullptr;  // box (24 B), status (4 B), P (4 B)

  ~drb_sim3() {
    for (Buf* b : {&d_inst, &mx[0], &mx[1], &my[0], &my[1], &mz[0], &mz[1], &mf[0], &mf[1], &keys, &vals, &comp, &cdir,
                   &cv, &swc, &cells, &rp, &cp0x, &cp0y, &cp0z, &cpq, &cpx, &cpy, &cpz, &cpf, &emask, &kap, &box,
                   &disp, &status, &diag, &bcnt, &boff, &ex, &exdir, &kc, &koff, &keys2, &vals2, &cellsb, &rpb,
                   &spos, &sem})
      b->release();
    sorter.release();
    scan.release();
    scan_b.release();
    scan_k.release();
    if (hpin) (void)hipHostFree(hpin);
  }
  int alloc_fixed() {
    DRB_HIP(hipHostMalloc(&hpin, 64, hipHostMallocDefault));
    DRB_CHECK(d_inst.ensure(sizeof(Inst3)));
    DRB_CHECK(box.ensure(sizeof(BBox3) + 16));
    DRB_CHECK(disp.ensure(16));
    DRB_CHECK(status.ensure(16));
    return DRB_OK;
  }
  BBox3* hbox() const { return (BBox3*)hpin; }
  unsigned* hstatus() const { return (unsigned*)((char*)hpin + sizeof(BBox3)); }
  int* hP() const { return (int*)((char*)hpin + sizeof(BBox3) + 4); }
  int* hE() const { return (int*)((char*)hpin + sizeof(BBox3) + 8); }

  int ensure_state(int which, long long cap) {
    if (cap <= mcap[which]) return DRB_OK;
    const long long c = std::max<long long>(cap + cap / 4, 1024);
    Buf nx, ny, nz, nf;
    DRB_CHECK(nx.ensure((size_t)c * 4));
    DRB_CHECK(ny.ensure((size_t)c * 4));
    DRB_CHECK(nz.ensure((size_t)c * 4));
    DRB_CHECK(nf.ensure((size_t)c * Q3 * 4));
    if (which == cur && M > 0 && mcap[which] > 0) {
      DRB_HIP(hipMemcpyAsync(nx.p, mx[which].p, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
      DRB_HIP(hipMemcpyAsync(ny.p, my[which].p, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
      DRB_HIP(hipMemcpyAsync(nz.p, mz[which].p, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
      DRB_HIP(hipMemcpy2DAsync(nf.p, (size_t)c * 4, mf[which].p, (size_t)mcap[which] * 4, (size_t)M * 4, Q3,
                               hipMemcpyDeviceToDevice, s));
      DRB_HIP(hipStreamSynchronize(s));
    }
    mx[which].release(); my[which].release(); mz[which].release(); mf[which].release();
    mx[which] = nx; my[which] = ny; mz[which] = nz; mf[which] = nf;
    mcap[which] = c;
    return DRB_OK;
  }

  int layout() {
    Inst3& p = h;
    if (p.periodic) {
      p.x0 = p.y0 = p.z0 = 0;
      p.nx = p.ny = p.nz = p.n;
    } else {
      const BBox3& b = *hbox();
      if (b.lo[0] > b.hi[0]) return fail(DRB_ERR_NONFINITE, "no finite position");
      long long n[3];
      for (int d = 0; d < 3; ++d) n[d] = (long long)b.hi[d] - b.lo[d] + 1;
      if (n[0] * n[1] * n[2] > (1LL << 31)) return fail(DRB_ERR_KEY_RANGE, "point cloud spans too many cells");
      p.x0 = b.lo[0];
      p.y0 = b.lo[1];
      p.z0 = b.lo[2];
      p.nx = (int)n[0];
      p.ny = (int)n[1];
      p.nz = (int)n[2];
    }
    ncells = (long long)p.nx * p.ny * p.nz;
    const double far = std::max({std::fabs((double)p.x0), std::fabs((double)p.x0 + p.nx), std::fabs((double)p.y0),
                                 std::fabs((double)p.y0 + p.ny), std::fabs((double)p.z0),
                                 std::fabs((double)p.z0 + p.nz), (double)p.n});
    p.margin = (float)((double)p.cs * std::max(1e-3, 16.0 * far * 0x1p-24));
    // Reach limit of the per-range image shifts (wrapd): L/2 less one cell, and one more cell
    // against the fp32 rounding of the window bounds.
    p.wlim = p.periodic ? (float)(0.5 * (double)p.L - 2.0 * (double)p.cs) : -1.f;
    DRB_HIP(hipMemcpyAsync(d_inst.p, &h, sizeof(Inst3), hipMemcpyHostToDevice, s));
    return DRB_OK;
  }

  int step_once() {
    if (M == 0) return DRB_OK;
    const long long N = (long long)Q3 * M;
    if (N > INT_MAX / 2) return fail(DRB_ERR_CAPACITY, "too many components");
    DRB_CHECK(layout());
    // Keys run over [0, ncells]; ncells marks a bounced component.
    int bits = 1;
    while (bits < 32 && (1ULL << bits) <= (unsigned long long)ncells) ++bits;
    const Inst3* I = d_inst.as<Inst3>();
    const int m = (int)M;
    const long long pitch = mcap[cur];
    const int nxt = cur ^ 1;
    hipLaunchKernelGGL(k3_box_init, dim3(1), dim3(64), 0, s, box.as<BBox3>(), disp.as<int>());
    hipLaunchKernelGGL(k3_collide_box, dim3(nblk(m)), dim3(TB), 0, s, mf[cur].as<float>(), pitch, mx[cur].as<float>(),
                       my[cur].as<float>(), mz[cur].as<float>(), m, I, box.as<BBox3>(), status.as<unsigned>(), 1, 0);
    // Obstacle: the number of bounce parts, read by the host to size the component arrays.
    long long E = 0;
    if (h.sphere) {
      DRB_CHECK(bcnt.ensure((size_t)(M + 1) * 8));
      DRB_CHECK(boff.ensure((size_t)(M + 1) * 8));
      hipLaunchKernelGGL(k3_bounce_count, dim3(nblk(m)), dim3(TB), 0, s, mx[cur].as<float>(), my[cur].as<float>(),
                         mz[cur].as<float>(), m, I, bcnt.as<int2>());
      DRB_CHECK(scan_b.run(bcnt.as<int2>(), boff.as<int2>(), M, s));
      DRB_HIP(hipMemcpyAsync(hE(), boff.as<int2>() + M, 4, hipMemcpyDeviceToHost, s));
      DRB_HIP(hipStreamSynchronize(s));
      E = *hE();
      if (N + E > INT_MAX / 2) return fail(DRB_ERR_CAPACITY, "too many components");
      DRB_CHECK(ex.ensure((size_t)std::max<long long>(E, 1) * sizeof(float4)));
      DRB_CHECK(exdir.ensure((size_t)std::max<long long>(E, 1)));
    }
    const long long NT = N + E;
    const long long pcap = std::max<long long>(1, std::min<long long>(NT, ncells));
    DRB_CHECK(ensure_state(nxt, pcap));
    DRB_CHECK(keys.ensure((size_t)NT * 4));
    DRB_CHECK(vals.ensure((size_t)NT * 4));
    DRB_CHECK(comp.ensure((size_t)NT * sizeof(Comp3)));
    DRB_CHECK(cdir.ensure((size_t)NT));
    DRB_CHECK(cv.ensure((size_t)NT * 4));
    DRB_CHECK(swc.ensure((size_t)NT * 4));
    DRB_CHECK(cells.ensure((size_t)(ncells + 1) * 8));
    DRB_CHECK(rp.ensure((size_t)(ncells + 1) * 8));
    DRB_CHECK(cellsb.ensure((size_t)(ncells + 1) * 8));
    DRB_CHECK(rpb.ensure((size_t)(ncells + 1) * 8));
    DRB_CHECK(keys2.ensure((size_t)pcap * 4));
    DRB_CHECK(vals2.ensure((size_t)pcap * 4));
    DRB_CHECK(spos.ensure((size_t)pcap * 16));
    DRB_CHECK(sem.ensure((size_t)pcap * 2));
    if (h.sphere) {
      DRB_CHECK(kc.ensure((size_t)(pcap + 1) * 8));
      DRB_CHECK(koff.ensure((size_t)(pcap + 1) * 8));
    }
    if (pcap > cpcap) {
      const long long c = pcap + pcap / 4;
      for (Buf* b : {&cp0x, &cp0y, &cp0z, &cpq, &cpx, &cpy, &cpz}) DRB_CHECK(b->ensure((size_t)c * 4));
      DRB_CHECK(cpf.ensure((size_t)c * Q3 * 4));
      DRB_CHECK(emask.ensure((size_t)c * 2));
      DRB_CHECK(kap.ensure((size_t)c));
      cpcap = c;
    }
    const long long pguess = std::min<long long>(pcap, M + M / 8 + 1024);
    hipLaunchKernelGGL(k3_keys, dim3(nblk(m)), dim3(TB), 0, s, mx[cur].as<float>(), my[cur].as<float>(),
                       mz[cur].as<float>(), m, I, keys.as<uint32_t>(), status.as<unsigned>());
    if (E > 0)
      hipLaunchKernelGGL(k3_bounce_emit, dim3(nblk(m)), dim3(TB), 0, s, mx[cur].as<float>(), my[cur].as<float>(),
                         mz[cur].as<float>(), mf[cur].as<float>(), pitch, m, I, boff.as<int2>(), (int)N,
                         keys.as<uint32_t>(), ex.as<float4>(), exdir.as<unsigned char>(), status.as<unsigned>());
    DRB_CHECK(sorter.sort(keys.as<uint32_t>(), vals.as<uint32_t>(), (int)NT, bits, true, s));
    DRB_HIP(hipMemsetAsync(cells.p, 0, (size_t)ncells * 8, s));
    hipLaunchKernelGGL(k3_mark, dim3(nblk(NT)), dim3(TB), 0, s, keys.as<uint32_t>(), vals.as<uint32_t>(), (int)NT,
                       mx[cur].as<float>(), my[cur].as<float>(), mz[cur].as<float>(), mf[cur].as<float>(), pitch, I,
                       comp.as<Comp3>(), cdir.as<unsigned char>(), cv.as<float>(), cells.as<int>(), (int)N,
                       ex.as<float4>(), exdir.as<unsigned char>(), (uint32_t)ncells);
    DRB_CHECK(scan.run(cells.as<int2>(), rp.as<int2>(), ncells, s));
    const int* Ptot = rp.as<int>() + 2 * ncells + 1;
    hipLaunchKernelGGL(k3_centroid, dim3(nblk(ncells)), dim3(TB), 0, s, rp.as<int2>(), ncells, comp.as<Comp3>(),
                       cp0x.as<float>(), cp0y.as<float>(), cp0z.as<float>(), cpq.as<int>());
    const int pm = h.periodic;
    DRB3_LAUNCH_PMG(k3_refine, pm, g_refine, dim3(xblk(gblk(pguess, g_refine))), dim3(TB), 0, s, Ptot, cp0x.as<float>(),
                    cp0y.as<float>(), cp0z.as<float>(), cpq.as<int>(), I, rp.as<int>(), comp.as<Comp3>(),
                    cdir.as<unsigned char>(), cpx.as<float>(), cpy.as<float>(), cpz.as<float>(),
                    kap.as<unsigned char>(), disp.as<int>());
    // bins of the refined control points (their registered cells may be up to disp cells away)
    hipLaunchKernelGGL(k3_cpkeys, dim3(nblk(pcap)), dim3(TB), 0, s, Ptot, cpx.as<float>(), cpy.as<float>(),
                       cpz.as<float>(), I, keys2.as<uint32_t>(), status.as<unsigned>());
    DRB_CHECK(sorter.sort(keys2.as<uint32_t>(), vals2.as<uint32_t>(), (int)pcap, bits, true, s, nullptr, Ptot));
    DRB_HIP(hipMemsetAsync(cellsb.p, 0, (size_t)ncells * 8, s));
    hipLaunchKernelGGL(k3_cpbins, dim3(nblk(pcap)), dim3(TB), 0, s, Ptot, keys2.as<uint32_t>(), vals2.as<uint32_t>(),
                       cpx.as<float>(), cpy.as<float>(), cpz.as<float>(), spos.as<float4>(), cellsb.as<int2>());
    DRB_CHECK(scan.run(cellsb.as<int2>(), rpb.as<int2>(), ncells, s));
    DRB3_LAUNCH_PM(k3_norm, pm, dim3(xblk(nblk(NT))), dim3(TB), 0, s, (int)NT, comp.as<Comp3>(), cv.as<float>(), I,
                   rpb.as<int>(), spos.as<float4>(), rp.as<int>() + 2 * ncells, swc.as<float>());
    DRB3_LAUNCH_PMG(k3_gather, pm, g_gather, dim3(xblk(gblk(pguess, g_gather))), dim3(TB), 0, s, Ptot, cpx.as<float>(),
                    cpy.as<float>(), cpz.as<float>(), I, rp.as<int>(), comp.as<Comp3>(), cdir.as<unsigned char>(),
                    cpf.as<float>(), cpcap, emask.as<unsigned short>());
    hipLaunchKernelGGL(k3_sem, dim3(nblk(pcap)), dim3(TB), 0, s, Ptot, spos.as<float4>(), emask.as<unsigned short>(),
                       sem.as<unsigned short>());
    const int* Mnext = Ptot;
    if (h.sphere) {
      hipLaunchKernelGGL(k3_keep, dim3(nblk(pcap)), dim3(TB), 0, s, Ptot, cpx.as<float>(), cpy.as<float>(),
                         cpz.as<float>(), pcap, I, kc.as<int2>());
      DRB_CHECK(scan_k.run(kc.as<int2>(), koff.as<int2>(), pcap, s));
      Mnext = koff.as<int>() + 2 * pcap;
    }
    DRB3_LAUNCH_PM(k3_gapfill_emit, pm, dim3(nblk(pcap)), dim3(TB), 0, s, Ptot, cpx.as<float>(), cpy.as<float>(),
                   cpz.as<float>(), I, rpb.as<int>(), spos.as<float4>(), sem.as<unsigned short>(), cpf.as<float>(),
                   cpcap, emask.as<unsigned short>(), mx[nxt].as<float>(), my[nxt].as<float>(), mz[nxt].as<float>(),
                   mf[nxt].as<float>(), mcap[nxt], h.sphere ? koff.as<int2>() : (const int2*)nullptr,
                   box.as<BBox3>(), status.as<unsigned>());
    DRB_HIP(hipGetLastError());
    DRB_HIP(hipMemcpyAsync(hbox(), box.p, sizeof(BBox3), hipMemcpyDeviceToHost, s));
    DRB_HIP(hipMemcpyAsync(hstatus(), status.p, 4, hipMemcpyDeviceToHost, s));
    DRB_HIP(hipMemcpyAsync(hP(), Mnext, 4, hipMemcpyDeviceToHost, s));
    DRB_HIP(hipStreamSynchronize(s));
    if (*hstatus() & ST_NONFINITE) return fail(DRB_ERR_NONFINITE, "non-finite value in the state");
    cur = nxt;
    M = *hP();
    return DRB_OK;
  }
};

extern "C" int drb_d3q13(float* C39, float* W13) {
  if (!C39 || !W13) return fail(DRB_ERR_INVALID_ARG, "drb_d3q13: null");
  double C[Q3][3], W[Q3];
  d3q13_fp64(C, W);
  for (int k = 0; k < Q3; ++k) {
    for (int d = 0; d < 3; ++d) C39[3 * k + d] = (float)C[k][d];
    W13[k] = (float)W[k];
  }
  return DRB_OK;
}

extern "C" int drb_sim3_create(drb_sim3** out, const drb_params* prm, drb_stream_t stream) {
  if (!out || !prm) return fail(DRB_ERR_INVALID_ARG, "drb_sim3_create: bad arguments");
  *out = nullptr;
  if (!(prm->tau > 0.f) || !(prm->dx > 0.f) || !(prm->alpha > 0.f))
    return fail(DRB_ERR_INVALID_ARG, "tau, dx and alpha must be > 0");
  if (prm->two_level) return fail(DRB_ERR_INVALID_ARG, "3D: two-level thinning is not supported");
  if (prm->kappa_soft > prm->kappa_hard)  // the reference's spawn is 2D only (simulator.py:7-9)
    return fail(DRB_ERR_INVALID_ARG, "3D: the soft_outer spawn is 2D only");
  drb_sim3* sim = new (std::nothrow) drb_sim3();
  if (!sim) return fail(DRB_ERR_CAPACITY, "out of host memory");
  sim->s = (hipStream_t)stream;
  Inst3& p = sim->h;
  double C[Q3][3], W[Q3];
  d3q13_fp64(C, W);
  for (int k = 0; k < Q3; ++k) {
    for (int d = 0; d < 3; ++d) {
      p.C[k][d] = (float)C[k][d];
      p.Cd[k][d] = C[k][d];
    }
    p.W[k] = (float)W[k];
    p.frest[k] = (float)((double)prm->rho_rest * W[k]);
  }
  p.tau = prm->tau;
  p.dx = prm->dx;
  p.cs = (float)((double)prm->dx / (double)prm->alpha);
  p.iq = 1.f / p.cs;
  p.rr = prm->r_res > 0.f ? prm->r_res : prm->dx;
  p.eps = prm->eps > 0.f ? prm->eps : 1e-6f;
  p.t_dx = sqrt_threshold(p.dx);
  p.t_rr = sqrt_threshold(p.rr);
  p.kappa_hard = prm->kappa_hard;
  p.periodic = prm->periodic ? 1 : 0;
  p.margin = 1e-3f * p.cs;
  int st = DRB_OK;
  if (p.periodic) {
    p.L = prm->box;
    const double n = std::round((double)prm->box / (double)p.cs);
    if (n < 1 || std::fabs(n * p.cs - prm->box) > 1e-5 * prm->box)
      st = fail(DRB_ERR_INVALID_ARG, "periodic box must be a whole number of thinning cells");
    p.n = (int)n;
    if (st == DRB_OK && !(std::max(p.rr, p.dx) < 0.5f * p.L)) st = fail(DRB_ERR_INVALID_ARG, "radius must be < box / 2");
  }
  if (st == DRB_OK) st = sim->alloc_fixed();
  if (st != DRB_OK) {
    delete sim;
    return st;
  }
  *out = sim;
  return DRB_OK;
}

extern "C" int drb_sim3_destroy(drb_sim3* sim) {
  if (!sim) return DRB_OK;
  (void)hipStreamSynchronize(sim->s);
  delete sim;
  return DRB_OK;
}

extern "C" int drb_sim3_set_sphere(drb_sim3* sim, float cx, float cy, float cz, float r) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim3_set_sphere: null handle");
  Inst3& p = sim->h;
  if (!(r > 0.f)) {
    p.sphere = 0;
    return DRB_OK;
  }
  if (!std::isfinite(cx) || !std::isfinite(cy) || !std::isfinite(cz) || !std::isfinite(r))
    return fail(DRB_ERR_INVALID_ARG, "drb_sim3_set_sphere: non-finite argument");
  if (p.periodic) {
    // Keep the obstacle and every component that can strike it away from the periodic seams, so
    // that plain differences are the minimum image wherever the obstacle is tested.
    const float c[3] = {cx, cy, cz};
    const double g = (double)r + 2.0 * std::max(p.dx, p.rr);
    for (int d = 0; d < 3; ++d)
      if (!((double)c[d] - g > 0.0 && (double)c[d] + g < (double)p.L))
        return fail(DRB_ERR_INVALID_ARG, "the obstacle must lie 2 dx inside the periodic box");
  }
  p.sphere = 1;
  p.sc[0] = cx;
  p.sc[1] = cy;
  p.sc[2] = cz;
  p.sR = r;
  p.sR2 = r * r;
  return DRB_OK;
}

extern "C" int drb_sim3_set_state(drb_sim3* sim, const float* pos, const float* f, int64_t m) {
  if (!sim || m < 0 || (m > 0 && (!pos || !f))) return fail(DRB_ERR_INVALID_ARG, "drb_sim3_set_state: bad arguments");
  if (m > (1LL << 26)) return fail(DRB_ERR_CAPACITY, "too many moments");
  hipStream_t s = sim->s;
  sim->M = 0;
  sim->cur = 0;
  DRB_CHECK(sim->ensure_state(0, std::max<long long>(m, 1)));
  sim->M = m;
  DRB_HIP(hipMemcpyAsync(sim->d_inst.p, &sim->h, sizeof(Inst3), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k3_box_init, dim3(1), dim3(64), 0, s, sim->box.as<BBox3>(), sim->disp.as<int>());
  DRB_HIP(hipMemsetAsync(sim->status.p, 0, 4, s));
  if (m > 0) {
    Buf tp, tf;
    DRB_CHECK(tp.ensure((size_t)m * 12));
    DRB_CHECK(tf.ensure((size_t)m * Q3 * 4));
    DRB_HIP(hipMemcpyAsync(tp.p, pos, (size_t)m * 12, hipMemcpyHostToDevice, s));
    DRB_HIP(hipMemcpyAsync(tf.p, f, (size_t)m * Q3 * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k3_unpack, dim3(nblk(m)), dim3(TB), 0, s, tp.as<float>(), tf.as<float>(), (int)m,
                       sim->mx[0].as<float>(), sim->my[0].as<float>(), sim->mz[0].as<float>(), sim->mf[0].as<float>(),
                       sim->mcap[0]);
    hipLaunchKernelGGL(k3_collide_box, dim3(nblk(m)), dim3(TB), 0, s, sim->mf[0].as<float>(), sim->mcap[0],
                       sim->mx[0].as<float>(), sim->my[0].as<float>(), sim->mz[0].as<float>(), (int)m,
                       sim->d_inst.as<Inst3>(), sim->box.as<BBox3>(), sim->status.as<unsigned>(), 0, 1);
    DRB_HIP(hipGetLastError());
    DRB_HIP(hipMemcpyAsync(sim->hbox(), sim->box.p, sizeof(BBox3), hipMemcpyDeviceToHost, s));
    DRB_HIP(hipMemcpyAsync(sim->hstatus(), sim->status.p, 4, hipMemcpyDeviceToHost, s));
    DRB_HIP(hipStreamSynchronize(s));
    tp.release();
    tf.release();
    if (*sim->hstatus() & ST_NONFINITE) return fail(DRB_ERR_NONFINITE, "non-finite value in the initial state");
    if (*sim->hstatus() & ST_RANGE) return fail(DRB_ERR_INVALID_ARG, "periodic positions must lie in [0, box)");
  }
  sim->have_state = true;
  return DRB_OK;
}

extern "C" int drb_sim3_step(drb_sim3* sim, int32_t nsteps) {
  if (!sim || nsteps < 0) return fail(DRB_ERR_INVALID_ARG, "drb_sim3_step: bad arguments");
  if (!sim->have_state) return fail(DRB_ERR_STATE, "drb_sim3_step: no state loaded");
  for (int i = 0; i < nsteps; ++i) DRB_CHECK(sim->step_once());
  return DRB_OK;
}

extern "C" int drb_sim3_set_lanes(drb_sim3* sim, int32_t refine_lanes, int32_t gather_lanes) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim3_set_lanes: null handle");
  for (int g : {refine_lanes, gather_lanes})
    if (g != 4 && g != 8 && g != 16) return fail(DRB_ERR_INVALID_ARG, "lanes per point must be 4, 8 or 16");
  sim->g_refine = refine_lanes;
  sim->g_gather = gather_lanes;
  return DRB_OK;
}

extern "C" int drb_sim3_count(drb_sim3* sim, int64_t* m) {
  if (!sim || !m) return fail(DRB_ERR_INVALID_ARG, "drb_sim3_count: bad arguments");
  *m = sim->M;
  return DRB_OK;
}

extern "C" int drb_sim3_get_state(drb_sim3* sim, float* pos, float* f) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim3_get_state: null handle");
  if (sim->M == 0) return DRB_OK;
  hipStream_t s = sim->s;
  Buf tp, tf;
  DRB_CHECK(tp.ensure((size_t)sim->M * 12));
  DRB_CHECK(tf.ensure((size_t)sim->M * Q3 * 4));
  hipLaunchKernelGGL(k3_pack, dim3(nblk(sim->M)), dim3(TB), 0, s, sim->mx[sim->cur].as<float>(),
                     sim->my[sim->cur].as<float>(), sim->mz[sim->cur].as<float>(), sim->mf[sim->cur].as<float>(),
                     sim->mcap[sim->cur], (int)sim->M, tp.as<float>(), tf.as<float>());
  DRB_HIP(hipGetLastError());
  if (pos) DRB_HIP(hipMemcpyAsync(pos, tp.p, (size_t)sim->M * 12, hipMemcpyDeviceToHost, s));
  if (f) DRB_HIP(hipMemcpyAsync(f, tf.p, (size_t)sim->M * Q3 * 4, hipMemcpyDeviceToHost, s));
  DRB_HIP(hipStreamSynchronize(s));
  tp.release();
  tf.release();
  return DRB_OK;
}

extern "C" int drb_sim3_diagnostics(drb_sim3* sim, double* out4) {
  if (!sim || !out4) return fail(DRB_ERR_INVALID_ARG, "drb_sim3_diagnostics: bad arguments");
  hipStream_t s = sim->s;
  DRB_CHECK(sim->diag.ensure(sizeof(double) * 4 * kDiag3Blocks));
  hipLaunchKernelGGL(k3_diag, dim3(kDiag3Blocks), dim3(TB), 0, s, sim->mf[sim->cur].as<float>(), sim->mcap[sim->cur],
                     (int)sim->M, sim->d_inst.as<Inst3>(), sim->diag.as<double>());
  DRB_HIP(hipGetLastError());
  std::vector<double> part(4 * kDiag3Blocks);
  DRB_HIP(hipMemcpyAsync(part.data(), sim->diag.p, sizeof(double) * 4 * kDiag3Blocks, hipMemcpyDeviceToHost, s));
  DRB_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < 4; ++q) {
    double a = 0.0;
    for (int b = 0; b < kDiag3Blocks; ++b) a += part[4 * b + q];
    out4[q] = a;
  }
  return DRB_OK;
}
