// Device-side BGK physics shared by the stage kernels, the step engine and the hex LBM.
//
// f_eq_k = rho W_k (1 + inv cu + inv^2 cu^2 / 2 - inv |u|^2 / 2), inv = 1/cs^2 = 4, evaluated in
// the operation order of direbm/warp/physics.py:39-42; collide relaxes f += (f_eq - f) / tau as
// direbm/warp/physics.py:52-62, with the rho == 0 guard (u = 0) of direbm/physics.py:24.
//
// Mass: BGK conserves sum_k f_k exactly, but the fp32 weights (W rounded to fp32, as the
// reference uploads them, direbm/warp/physics.py:25-28) sum to 1 + 1.5e-8, so each collide in fp32
// would add 1.5e-8 rho / tau to every moment -- a drift of 1.5e-6 over a 100-step run.  The rest
// equilibrium therefore takes the remainder, f_eq_0 = rho - sum_{k>0} f_eq_k (equal to the formula
// in exact arithmetic, since sum_k f_eq_k = rho), and every population relaxes as before: only
// unbiased rounding is left in the mass, and the dependence on tau keeps its direct form.
#pragma once

#include "common.h"

namespace drb {

__device__ __forceinline__ float feq_k(float rho, float ux, float uy, int k) {
  const float inv = 4.0f;
  float cu = fadd(fmul(kCx[k], ux), fmul(kCy[k], uy));
  float usq = fadd(fmul(ux, ux), fmul(uy, uy));
  float a = fadd(1.0f, fmul(inv, cu));
  float b = fmul(fmul(fmul(fmul(0.5f, inv), inv), cu), cu);
  float c = fmul(fmul(0.5f, inv), usq);
  return fmul(fmul(rho, kW[k]), fsub(fadd(a, b), c));
}

__device__ __forceinline__ void collide7(float f[7], float tau, unsigned& st) {
  float rho = 0.f, mx = 0.f, my = 0.f;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    rho = fadd(rho, f[k]);
    mx = fadd(mx, fmul(f[k], kCx[k]));
    my = fadd(my, fmul(f[k], kCy[k]));
  }
  float ux = 0.f, uy = 0.f;
  if (rho != 0.f) {
    ux = fdiv(mx, rho);
    uy = fdiv(my, rho);
  } else {
    st |= ST_EMPTY_RHO;
  }
  float fe[7], moving = 0.f;
#pragma unroll
  for (int k = 1; k < 7; ++k) {
    fe[k] = feq_k(rho, ux, uy, k);
    moving = fadd(moving, fe[k]);
  }
  fe[0] = fsub(rho, moving);
#pragma unroll
  for (int k = 0; k < 7; ++k) f[k] = fadd(f[k], fdiv(fsub(fe[k], f[k]), tau));
}

// Vector-Jacobian products (the adjoint).  Plain fp32; no reference counterpart.
//
// d f_eq_k / d rho = W_k (1 + inv cu + inv^2 cu^2 / 2 - inv |u|^2 / 2)
// d f_eq_k / d u   = rho W_k (inv c_k + inv^2 cu c_k - inv u)
__device__ __forceinline__ void feq_vjp(float rho, float ux, float uy, const float g[7], float& grho, float& gux,
                                        float& guy) {
  const float inv = 4.0f;
  float usq = ux * ux + uy * uy;
  grho = 0.f;
  gux = 0.f;
  guy = 0.f;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    float cu = kCx[k] * ux + kCy[k] * uy;
    float base = kW[k] * (1.f + inv * cu + 0.5f * inv * inv * cu * cu - 0.5f * inv * usq);
    float rw = rho * kW[k];
    float a = rw * (inv + inv * inv * cu);
    grho += g[k] * base;
    gux += g[k] * (a * kCx[k] - rw * inv * ux);
    guy += g[k] * (a * kCy[k] - rw * inv * uy);
  }
}

// Adjoint of collide7: given the pre-collision f and the cotangent g of the collided values,
// returns the cotangent of f and the contribution to d/d tau.  rho, u = sum f, sum f c / rho (u = 0
// and no velocity path when rho == 0, as the forward guard).
__device__ __forceinline__ void collide7_vjp(const float f[7], float tau, const float g[7], float gin[7],
                                             float& gtau) {
  float rho = 0.f, mx = 0.f, my = 0.f;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    rho += f[k];
    mx += f[k] * kCx[k];
    my += f[k] * kCy[k];
  }
  const bool live = rho != 0.f;
  const float ux = live ? mx / rho : 0.f, uy = live ? my / rho : 0.f;
  const float it = 1.f / tau;
  float gp[7];
  gtau = 0.f;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    gp[k] = g[k] * it;
    gtau -= gp[k] * (feq_k(rho, ux, uy, k) - f[k]) * it;
  }
  float grho, gux, guy;
  feq_vjp(rho, ux, uy, gp, grho, gux, guy);
  const float ir = live ? 1.f / rho : 0.f;
#pragma unroll
  for (int k = 0; k < 7; ++k)
    gin[k] = (1.f - it) * g[k] + grho + (gux * (kCx[k] - ux) + guy * (kCy[k] - uy)) * ir;
}

}  // namespace drb
