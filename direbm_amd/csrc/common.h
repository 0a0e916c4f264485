// Shared definitions of the DiReBM HIP engine (gfx950).
//
// Arithmetic contract: every translation unit is compiled with -ffp-contract=off and with
// IEEE-correct fp32 division and square root, so cell keys and radius predicates are evaluated
// exactly as the reference kernels define them (direbm/warp/propagation.py:34,48-49,98,175,213):
// one rounding per +, -, *, /, sqrt.
#pragma once

#include <type_traits>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/direbm_hip.h"

namespace drb {

// ------------------------------------------------------------------------------------ errors
struct Error {
  int code;
  std::string msg;
};
void set_error(int code, const std::string& msg);
int fail(int code, const std::string& msg);

#define DRB_HIP(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::drb::fail(DRB_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define DRB_CHECK(expr)       \
  do {                        \
    int s_ = (expr);          \
    if (s_ != DRB_OK) return s_; \
  } while (0)

// ------------------------------------------------------------------------------- constants
// D2Q7 velocity set rounded to fp32, as the reference uploads it (direbm/warp/physics.py:25-28).
#define DRB_S32 0.866025388240814208984375f  // float32(sqrt(3)/2)
__device__ __constant__ static const float kCx[7] = {0.f, 1.f, 0.5f, -0.5f, -1.f, -0.5f, 0.5f};
__device__ __constant__ static const float kCy[7] = {0.f, 0.f, DRB_S32, DRB_S32, 0.f, -DRB_S32, -DRB_S32};
__device__ __constant__ static const float kW[7] = {0.5f, 0.083333335816860198974609375f,
                                                    0.083333335816860198974609375f,
                                                    0.083333335816860198974609375f,
                                                    0.083333335816860198974609375f,
                                                    0.083333335816860198974609375f,
                                                    0.083333335816860198974609375f};
static const float hCx[7] = {0.f, 1.f, 0.5f, -0.5f, -1.f, -0.5f, 0.5f};
static const float hCy[7] = {0.f, 0.f, DRB_S32, DRB_S32, 0.f, -DRB_S32, -DRB_S32};

// ------------------------------------------------------------------ per-instance parameters
// One entry of the segment table.  A plain simulation is a table with one entry; the batched
// form concatenates many.  Read-only on the device except the fields marked (device).
struct alignas(16) Inst {
  float tau, dx, cs, rr;      // relaxation time, dispersion length, thinning cell, resample radius
  float eps, qcs, L, margin;  // gap-fill threshold, query-cell size, periodic box, window margin
  float frest[7];             // rest equilibrium rho_rest * W
  float cx, cy, rin2;         // two-level thinning centre and squared inner radius (fp32)
  int kappa_hard, periodic, two_level, nfine;  // nfine: fine cells per period (periodic)
  int nq;                     // query cells per period (periodic)
  int qx0, qy0, nqx, nqy;     // dense query grid origin and extent (in query cells)
  long long cell_off;         // offset of this instance's grid in the global dense grid
  int m_off, m_cnt;           // moments of this instance (host-side bookkeeping)
  float t_dx, t_rr;           // radius predicates: sqrt_rn(s) <= r  <=>  s <= t_r (sqrt_threshold)
  float iq;                   // 1 / qcs (window bounds)
  float wlim;                 // periodic: L/2 - 2 qcs, the reach below which image shifts are per range
  int obstacle;               // 1: circular obstacle (fluid outside) at (ocx, ocy), radius oR
  int kappa_soft;             // > kappa_hard: soft_outer spawn for kappa_hard < kappa <= kappa_soft
  float ocx, ocy, oR, oR2;    // its centre, radius and fp32 oR * oR
};

// Per-instance bounding box of query cells, accumulated with atomics (device).
struct BBox {
  int qxmin, qxmax, qymin, qymax;
};

// Status bits, OR-ed on the device and read by the host.
enum : unsigned {
  ST_NONFINITE = 1u,  // non-finite f or position met
  ST_EMPTY_RHO = 2u,  // rho == 0 met in collide (u = 0 used, as direbm/physics.py:24)
  ST_RANGE = 4u,      // an input lies outside its range: a position outside the periodic box [0, L), or a
                      // stage call's direction outside 0..6
};

// D2Q7 velocity set in FP64 (direbm/constants.py:16-27), for the obstacle's FP64 bounce geometry.
#define DRB_S64 0.86602540378443860  // sqrt(3) / 2
__device__ __constant__ static const double kCd[7][2] = {{0.0, 0.0},  {1.0, 0.0},      {0.5, DRB_S64}, {-0.5, DRB_S64},
                                                          {-1.0, 0.0}, {-0.5, -DRB_S64}, {0.5, -DRB_S64}};

// ----------------------------------------------------------------------- device helpers
// Plain operators are single-rounded under -ffp-contract=off; '/' and sqrtf are the IEEE
// correctly rounded forms under HIP's default -fhip-fp32-correctly-rounded-divide-sqrt.
// (__fsqrt_rn is deliberately not used: without OCML_BASIC_ROUNDED_OPERATIONS it maps to the
// approximate native square root.)
__device__ __forceinline__ float fsub(float a, float b) { return a - b; }
__device__ __forceinline__ float fadd(float a, float b) { return a + b; }
__device__ __forceinline__ float fmul(float a, float b) { return a * b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a / b; }
__device__ __forceinline__ float fsqrt(float a) { return __builtin_sqrtf(a); }

// Wrap an fp32 coordinate into [0, L) so that floor(x / cs) < n (periodic box).
__device__ __forceinline__ float wrap_coord(float x, float L, float cs, int n) {
  if (x < 0.f) x = fadd(x, L);
  if (x >= L) x = fsub(x, L);
  if (floorf(fdiv(x, cs)) >= (float)n) x = 0.f;
  return x;
}

// Minimum-image difference in a periodic box (d itself in an open domain).  Written with selects
// so that it compiles to v_cndmask rather than divergent branches in the neighbourhood loops.
__device__ __forceinline__ float min_image(float d, float L, int periodic) {
  const float h = fmul(0.5f, L);
  float m = d > h ? fsub(d, L) : d;
  m = d < -h ? fadd(d, L) : m;
  return periodic ? m : d;
}

// Squared fp32 distance of the reference's wp.length(a - b) before its square root: one
// rounding per operation, no FMA.
__device__ __forceinline__ float dist2_32(float ax, float ay, float bx, float by, float L, int periodic) {
  float ddx = min_image(fsub(ax, bx), L, periodic);
  float ddy = min_image(fsub(ay, by), L, periodic);
  return fadd(fmul(ddx, ddx), fmul(ddy, ddy));
}

// The reference predicate sqrt_rn(s) <= r is monotone in s, so it equals s <= t_r with t_r the
// largest fp32 value whose correctly rounded square root is <= r (computed on the host by
// sqrt_threshold).  The square root is then only needed for the weights of accepted pairs,
// where the hardware square root (<= 1 ulp) is used.
__device__ __forceinline__ float fast_sqrt(float s) { return __builtin_amdgcn_sqrtf(s); }
// Hardware reciprocal and reciprocal square root (about 1 ulp), for the adjoint's continuous
// terms, which carry no bit-level contract.
__device__ __forceinline__ float fast_rcp(float s) { return __builtin_amdgcn_rcpf(s); }
__device__ __forceinline__ float fast_rsq(float s) { return __builtin_amdgcn_rsqf(s); }

// Reference distance wp.length(a - b): one rounding per operation, no FMA.
__device__ __forceinline__ float dist32(float ax, float ay, float bx, float by, float L, int periodic) {
  float ddx = min_image(fsub(ax, bx), L, periodic);
  float ddy = min_image(fsub(ay, by), L, periodic);
  return fsqrt(fadd(fmul(ddx, ddx), fmul(ddy, ddy)));
}

// Fine thinning cell floor(x / cs) with an IEEE fp32 division (propagation.py:48-49).
// The float is bounded before the conversion, which is undefined for NaN, infinities and values
// past the int range: fmaxf / fminf return the bound for NaN, so such a coordinate gets a cell far
// outside every grid, which the callers clamp into the grid and flag.
__device__ __forceinline__ int cell_int(float v) { return (int)fminf(fmaxf(v, -1.0e9f), 1.0e9f); }
__device__ __forceinline__ int fine_cell(float x, float cs) { return cell_int(floorf(fdiv(x, cs))); }

// Component position pos_m + C[k]*dx (propagation.py:34), wrapped in a periodic box.
__device__ __forceinline__ void comp_pos(float mx, float my, int k, const Inst& p, float& x, float& y) {
  x = fadd(mx, fmul(kCx[k], p.dx));
  y = fadd(my, fmul(kCy[k], p.dx));
  if (p.periodic) {
    x = wrap_coord(x, p.L, p.cs, p.nfine);
    y = wrap_coord(y, p.L, p.cs, p.nfine);
  }
}

// Strictly inside the instance's circular obstacle, fp32 (dx^2 + dy^2) < R * R
// (direbm/reference/boundary.py:24-26).
__device__ __forceinline__ bool inside2(float x, float y, const Inst& p) {
  const float dx = fsub(x, p.ocx), dy = fsub(y, p.ocy);
  return fadd(fmul(dx, dx), fmul(dy, dy)) < p.oR2;
}

// Specular bounce of component k, which would land at (x1, y1) inside the obstacle, from its moment
// at (x0, y0) (direbm/reference/simulator.py:105-117): FP64 first crossing of the segment with the
// circle and the outward normal there (boundary.py:28-47), c_k reflected about it (:56-60) and split
// over the two adjacent lattice directions by angle with linear weights (:63-70).  Writes the two
// parts' directions and weights; returns their number (0: no crossing, the component is lost, as
// in the reference).
__device__ inline int bounce2(float x0, float y0, float x1, float y1, int k, const Inst& p, int pj[2], double pw[2]) {
  const double dx = (double)x1 - (double)x0, dy = (double)y1 - (double)y0;
  const double fx = (double)x0 - (double)p.ocx, fy = (double)y0 - (double)p.ocy;
  const double R = (double)p.oR;
  const double A = dx * dx + dy * dy;
  const double B = 2.0 * (fx * dx + fy * dy);
  const double Cc = (fx * fx + fy * fy) - R * R;
  const double disc = B * B - 4.0 * A * Cc;
  if (A == 0.0 || disc < 0.0) return 0;
  const double sq = sqrt(disc);
  const double t0 = (-B - sq) / (2.0 * A), t1 = (-B + sq) / (2.0 * A);
  double t;
  if (0.0 <= t0 && t0 <= 1.0) t = t0;
  else if (0.0 <= t1 && t1 <= 1.0) t = t1;
  else return 0;
  const double nx = (((double)x0 + t * dx) - (double)p.ocx) / R, ny = (((double)y0 + t * dy) - (double)p.ocy) / R;
  const double cn = kCd[k][0] * nx + kCd[k][1] * ny;
  const double rx = kCd[k][0] - 2.0 * cn * nx, ry = kCd[k][1] - 2.0 * cn * ny;
  const double two_pi = 2.0 * 3.141592653589793;
  double theta = fmod(atan2(ry, rx), two_pi);
  if (theta < 0.0) theta += two_pi;
  const double seg = theta / (3.141592653589793 / 3.0);
  const double fk = floor(seg);
  const double frac = seg - fk;
  const int a = (int)fk % 6;
  pj[0] = a + 1;
  pw[0] = 1.0 - frac;
  pj[1] = (a + 1) % 6 + 1;
  pw[1] = frac;
  return 2;
}

// Query cell and thinning sub-cell of a component position.
__device__ __forceinline__ void qcell_of(float x, float y, const Inst& p, int& qx, int& qy, int& sub) {
  int fx = fine_cell(x, p.cs), fy = fine_cell(y, p.cs);
  if (p.periodic) {
    if (fx >= p.nfine) fx -= p.nfine;
    if (fy >= p.nfine) fy -= p.nfine;
  }
  if (!p.two_level) {
    qx = fx; qy = fy; sub = 0;
  } else {
    float ex = fsub(x, p.cx), ey = fsub(y, p.cy);
    bool inner = fadd(fmul(ex, ex), fmul(ey, ey)) < p.rin2;
    qx = fx >> 1; qy = fy >> 1;
    sub = inner ? (((fx & 1) << 1) | (fy & 1)) : 4;
  }
}

// Global dense cell index of a query cell (x-major, matching the reference key order).
__device__ __forceinline__ long long qindex(const Inst& p, int qx, int qy) {
  return p.cell_off + (long long)(qx - p.qx0) * p.nqy + (qy - p.qy0);
}

// Columns of query cells that may hold a point within radius r of (x, y), widened by `ext`
// query cells: col(cx, loy, hiy) is called for each column cx in ascending order, in the query
// point's own frame, with the row range [loy, hiy] unwrapped (periodic) or clamped to the grid
// (open).  The margin covers fp32 rounding of cell indices.
//
// The bounds use the hardware square root and a multiplication by 1/q: both err by a few ulps,
// far inside the margin (1e-3 of a cell), so the window still covers the disk and the points it
// yields within r are exactly those of an exact window.
// Minimum image of a candidate-minus-query difference from its range's wrap code (two bits per
// axis, x in bits 0-1): 0 the plain difference, 1 subtract L, 2 add L, 3 the general minimum image.
// A candidate in a window cell lies within reach = span + rm + qcs of every query point the window
// serves (span: the extent of a box window's box).  With reach < L/2, every candidate of an
// unwrapped range has |d| < L/2 (min_image leaves d alone), and every candidate of a range wrapped
// below 0 (above n) has d > L/2 (d < -L/2), where min_image applies exactly the fp32 subtraction
// (addition) of L that the code names.  So codes 0-2 are the minimum image bit for bit and a range
// can skip the per-candidate selects; windows with a longer reach, or covering a whole period,
// use code 3.  Open instances always get code 0.  This needs every candidate to lie in the cell
// it is listed under: it holds for components (keyed by their own position), not for refined
// control points listed under their pre-refine cells.
__device__ __forceinline__ float wrap_shift(float d, int code, float L) {
  if (code == 0) return d;
  if (code == 1) return fsub(d, L);
  if (code == 2) return fadd(d, L);
  return min_image(d, L, 1);
}

// Squared fp32 distance of candidate a from query b with the range's wrap code; PLAIN: code 0.
template <bool PLAIN>
__device__ __forceinline__ float dist2_w(float ax, float ay, float bx, float by, int code, float L) {
  float dx = fsub(ax, bx), dy = fsub(ay, by);
  if (!PLAIN) {
    dx = wrap_shift(dx, code & 3, L);
    dy = wrap_shift(dy, code >> 2, L);
  }
  return fadd(fmul(dx, dx), fmul(dy, dy));
}

// Call fn(a, e, code) when it takes the wrap code, else fn(a, e).
template <class F>
__device__ __forceinline__ void call_range(F& fn, int a, int e, int code) {
  if constexpr (std::is_invocable_v<F&, int, int, int>) fn(a, e, code);
  else fn(a, e);
}

template <class F>
__device__ __forceinline__ void window_cols(const Inst& p, float x, float y, float r, int ext, F&& col) {
  const float q = p.qcs, iq = p.iq;
  const float rm = r + p.margin + (float)ext * q;
  const bool fast = rm < p.wlim;
  int lox = (int)floorf((x - rm) * iq), hix = (int)floorf((x + rm) * iq);
  bool fullx = false;
  if (p.periodic) {
    if (hix - lox + 1 >= p.nq) { lox = 0; hix = p.nq - 1; fullx = true; }
  } else {
    lox = max(lox, p.qx0);
    hix = min(hix, p.qx0 + p.nqx - 1);
  }
  for (int cx = lox; cx <= hix; ++cx) {
    float gx = 0.f;
    if (!fullx) {
      const float x0 = (float)cx * q, x1 = x0 + q;
      gx = fmaxf(fmaxf(x0 - x, x - x1), 0.f);
      if (gx > rm) continue;
    }
    const float hy = fast_sqrt(fmaxf(rm * rm - gx * gx, 0.f));
    int loy = (int)floorf((y - hy) * iq), hiy = (int)floorf((y + hy) * iq);
    int code = 0;  // x bits of the wrap code, plus 3 << 2 when the rows cover a whole period
    if (p.periodic) {
      code = (fullx || !fast) ? 3 : cx < 0 ? 1 : cx >= p.nq ? 2 : 0;
      if (hiy - loy + 1 >= p.nq) { loy = 0; hiy = p.nq - 1; code |= 3 << 2; }
      else if (!fast) code |= 3 << 2;
    } else {
      loy = max(loy, p.qy0);
      hiy = min(hiy, p.qy0 + p.nqy - 1);
      if (loy > hiy) continue;
    }
    col(cx, loy, hiy, code);
  }
}

// Visit every contiguous range of sorted items (global memory) in the window of (x, y): rp2 is
// one component (x: components, y: control points) of the dense int2 exclusive prefix (ncells + 1
// entries), addressed with stride 2.  fn(begin, end) is called per column segment; the visiting
// order (columns ascending, rows ascending in the query's frame) defines the summation order.
template <class F>
__device__ __forceinline__ void for_window(const Inst& p, const int* __restrict__ rp2, float x, float y,
                                           float r, int ext, F&& fn) {
  window_cols(p, x, y, r, ext, [&](int cx, int loy, int hiy, int code) {
    if (p.periodic) {
      const int n = p.nq;
      int col = cx % n;
      if (col < 0) col += n;
      const long long base = p.cell_off + (long long)col * n;
      const int gy = (code >> 2) == 3 ? 3 << 2 : 0;  // the y bits of a whole-period or long-reach column
      const int cxb = code & 3;
      if (loy < 0)
        call_range(fn, rp2[2 * (base + loy + n)], rp2[2 * (base + min(hiy, -1) + n + 1)], cxb | (gy ? gy : 1 << 2));
      const int a = max(loy, 0), b = min(hiy, n - 1);
      if (a <= b) call_range(fn, rp2[2 * (base + a)], rp2[2 * (base + b + 1)], cxb | gy);
      if (hiy >= n)
        call_range(fn, rp2[2 * (base + max(loy, n) - n)], rp2[2 * (base + hiy - n + 1)], cxb | (gy ? gy : 2 << 2));
    } else {
      const long long base = p.cell_off + (long long)(cx - p.qx0) * p.nqy - p.qy0;
      call_range(fn, rp2[2 * (base + loy)], rp2[2 * (base + hiy + 1)], 0);
    }
  });
}

// Load p[i] for a wave-uniform index through the scalar cache (the index is moved to a scalar
// register first, so the load is a scalar load rather than a vector load of 64 equal addresses).
__device__ __forceinline__ int ldu(const int* __restrict__ p, long long i) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(i & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane((int)(i >> 32));
  return p[((long long)hi << 32) | (unsigned)lo];
}

// Visit the sorted ranges (global memory) of every cell that may hold a point within r of some
// point of the box [x0, x1] x [y0, y1], widened by `ext` cells, walked columns ascending, rows
// ascending.  For a single point (x0 = x1, y0 = y1) this is exactly for_window; for a box, every
// point's own hits appear in the same order as in its for_window.
template <class F>
__device__ __forceinline__ void for_box_window(const Inst& p, const int* __restrict__ rp2, float x0, float x1, float y0,
                                               float y1, float r, int ext, F&& fn) {
  const float q = p.qcs, iq = p.iq;
  const float rm = r + p.margin + (float)ext * q;
  const bool fast = fmaxf(x1 - x0, y1 - y0) + rm < p.wlim;
  int lox = (int)floorf((x0 - rm) * iq), hix = (int)floorf((x1 + rm) * iq);
  bool fullx = false;
  if (p.periodic) {
    if (hix - lox + 1 >= p.nq) { lox = 0; hix = p.nq - 1; fullx = true; }
  } else {
    lox = max(lox, p.qx0);
    hix = min(hix, p.qx0 + p.nqx - 1);
  }
  for (int cx = lox; cx <= hix; ++cx) {
    float gx = 0.f;
    if (!fullx) {
      const float c0 = (float)cx * q, c1 = c0 + q;
      gx = fmaxf(fmaxf(c0 - x1, x0 - c1), 0.f);
      if (gx > rm) continue;
    }
    const float hy = fast_sqrt(fmaxf(rm * rm - gx * gx, 0.f));
    int loy = (int)floorf((y0 - hy) * iq), hiy = (int)floorf((y1 + hy) * iq);
    if (p.periodic) {
      int gy = (fullx || !fast) ? 3 << 2 : 0;
      if (hiy - loy + 1 >= p.nq) { loy = 0; hiy = p.nq - 1; gy = 3 << 2; }
      const int cxb = (fullx || !fast) ? 3 : cx < 0 ? 1 : cx >= p.nq ? 2 : 0;
      const int n = p.nq;
      int col = cx % n;
      if (col < 0) col += n;
      const long long base = p.cell_off + (long long)col * n;
      if (loy < 0)
        call_range(fn, ldu(rp2, 2 * (base + loy + n)), ldu(rp2, 2 * (base + min(hiy, -1) + n + 1)), cxb | (gy ? gy : 1 << 2));
      const int a = max(loy, 0), b = min(hiy, n - 1);
      if (a <= b) call_range(fn, ldu(rp2, 2 * (base + a)), ldu(rp2, 2 * (base + b + 1)), cxb | gy);
      if (hiy >= n)
        call_range(fn, ldu(rp2, 2 * (base + max(loy, n) - n)), ldu(rp2, 2 * (base + hiy - n + 1)), cxb | (gy ? gy : 2 << 2));
    } else {
      loy = max(loy, p.qy0);
      hiy = min(hiy, p.qy0 + p.nqy - 1);
      if (loy > hiy) continue;
      const long long base = p.cell_off + (long long)(cx - p.qx0) * p.nqy - p.qy0;
      call_range(fn, ldu(rp2, 2 * (base + loy)), ldu(rp2, 2 * (base + hiy + 1)), 0);
    }
  }
}

// Component record in sorted order: position, w, and (direction | instance << 3).  w is the refine
// weight exp(value) until the normaliser replaces it with value / sw for the gather.  The value
// itself lives in a separate array.
struct alignas(16) Comp {
  float x, y, w;
  int bits;
};

// Largest fp32 s with sqrt_rn(s) <= r (host, IEEE sqrtf).
float sqrt_threshold(float r);

// ------------------------------------------------------------------- host-side primitives
// Device buffer owned by the engine (grow-only).
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  bool view = false;  // p points into another Buf's memory: never grown or freed through this one
  int ensure(size_t want);
  void release();
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
  void set_view(void* q) {
    p = q;
    bytes = 0;
    view = true;
  }
};

// Stable LSD radix sort of (key, value) pairs, 8-bit digits, `bits` live key bits.
// Scratch is grown as needed.  Result lands in (keys, vals); the alternate buffers are scratch.
// sort_swap: the same sort on Buf-owned arrays; after an odd number of passes the result is in the
// alternate buffers and the Bufs are exchanged with them instead of copying the result back.
struct RadixSort {
  Buf k1, v1, hist, tot;
  int sort_swap(Buf& keys, Buf& vals, int n, int bits, bool vals_are_iota, hipStream_t s);
  // ndev (optional): the device holds the live count (<= n, the capacity launched for), so the
  // sort can be captured in a graph without knowing the count on the host.
  int sort(uint32_t* keys, uint32_t* vals, int n, int bits, bool vals_are_iota, hipStream_t s,
           float* ms_out = nullptr, const int* ndev = nullptr);
  // For graph capture: the live count on the device (<= n), values the iota, and the result left
  // where the last pass wrote it (*kres, *vres: keys or k1) instead of copied back.  Call reserve(n)
  // before capturing.
  int sort_graph(uint32_t* keys, uint32_t* vals, int n, int bits, hipStream_t s, const int* ndev, uint32_t** kres,
                 uint32_t** vres);
  int reserve(int n);
  void release() { k1.release(); v1.release(); hist.release(); tot.release(); }
};

// The same sort on caller-provided scratch (allocates nothing): k1, v1 hold n words each, hist
// radix_hist_words(n) words, tot 256 words.
size_t radix_hist_words(int n);
// kres / vres (optional): the result is left where the last pass wrote it (keys or k1) and its
// buffers are returned there, instead of copied back into keys / vals.
int radix_sort_into(uint32_t* keys, uint32_t* vals, int n, int bits, bool vals_are_iota, uint32_t* k1, uint32_t* v1,
                    uint32_t* hist, uint32_t* tot, hipStream_t s, const int* ndev = nullptr,
                    uint32_t** kres = nullptr, uint32_t** vres = nullptr);

// Device-wide exclusive scan of int2 (x, y) counts: out[i] = sum_{j<i} in[j], out[n] = total.
// scan2_into runs it on caller-provided scratch of scan2_part_elems(n) int2 (allocates nothing).
size_t scan2_part_elems(long long n);
int scan2_into(const int2* in, int2* out, long long n, int2* part, hipStream_t s);
struct Scan2 {
  Buf part;
  int run(const int2* in, int2* out, long long n, hipStream_t s);
  int reserve(long long n);
  void release() { part.release(); }
};

// FP64 deterministic reduction helper: partials[b * nq + q] -> out[q] (fixed order).
int reduce_partials(const double* partials, int nblocks, int nq, double* out, hipStream_t s);

}  // namespace drb
