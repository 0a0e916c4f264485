// Device-side BGK physics shared by the stage kernels, the step engine and the hex LBM.
//
// f_eq_k = rho W_k (1 + inv cu + inv^2 cu^2 / 2 - inv |u|^2 / 2), inv = 1/cs^2 = 4, evaluated in
// the operation order of direbm/warp/physics.py:39-42; collide relaxes f += (f_eq - f) / tau as
// direbm/warp/physics.py:52-62, with the rho == 0 guard (u = 0) of direbm/physics.py:24.
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
#pragma unroll
  for (int k = 0; k < 7; ++k) f[k] = fadd(f[k], fdiv(fsub(feq_k(rho, ux, uy, k), f[k]), tau));
}

}  // namespace drb
