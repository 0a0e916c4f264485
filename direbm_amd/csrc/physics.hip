// BGK collide / equilibrium / dispersion stage kernels and their C-ABI entry points.
//
// Reference: direbm/warp/physics.py:31-97 (_equilibrium_row, _collide, _equilibrium and the
// host wrappers) and direbm/warp/propagation.py:21-36,115-127 (_disperse, disperse()).
// One thread per point; the distribution is addressed through (stride_i, stride_k) so the same
// kernel serves the reference's row-major (N, 7) arrays and the engine's SoA planes.
#include "physics.h"

namespace drb {

__global__ __launch_bounds__(256) void k_collide_strided(float* __restrict__ f, long long n, long long si,
                                                         long long sk, float tau) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v[7];
  float* row = f + i * si;
#pragma unroll
  for (int k = 0; k < 7; ++k) v[k] = row[k * sk];
  unsigned st = 0;
  collide7(v, tau, st);
#pragma unroll
  for (int k = 0; k < 7; ++k) row[k * sk] = v[k];
}

__global__ __launch_bounds__(256) void k_equilibrium_strided(const float* __restrict__ rho,
                                                             const float* __restrict__ u, long long n,
                                                             float* __restrict__ out, long long si,
                                                             long long sk) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float r = rho[i], ux = u[2 * i], uy = u[2 * i + 1];
#pragma unroll
  for (int k = 0; k < 7; ++k) out[i * si + k * sk] = feq_k(r, ux, uy, k);
}

__global__ __launch_bounds__(256) void k_disperse_aos(const float* __restrict__ pos_m, const float* __restrict__ f_m,
                                                      long long m, float dx, float* __restrict__ pos_c,
                                                      float* __restrict__ val_c, int* __restrict__ dir_c) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  float x = pos_m[2 * i], y = pos_m[2 * i + 1];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    long long c = 7 * i + k;
    pos_c[2 * c] = fadd(x, fmul(kCx[k], dx));
    pos_c[2 * c + 1] = fadd(y, fmul(kCy[k], dx));
    val_c[c] = f_m[c];
    dir_c[c] = k;
  }
}

__global__ __launch_bounds__(256) void k_collide_vjp(const float* __restrict__ f, long long n, long long si,
                                                     long long sk, float tau, const float* __restrict__ gout,
                                                     float* __restrict__ gin, double* __restrict__ gtau) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v[7], g[7], r[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    v[k] = f[i * si + k * sk];
    g[k] = gout[i * si + k * sk];
  }
  float gt;
  collide7_vjp(v, tau, g, r, gt);
#pragma unroll
  for (int k = 0; k < 7; ++k) gin[i * si + k * sk] = r[k];
  if (gtau) gtau[i] = (double)gt;
}

__global__ __launch_bounds__(256) void k_equilibrium_vjp(const float* __restrict__ rho, const float* __restrict__ u,
                                                         long long n, const float* __restrict__ g, long long si,
                                                         long long sk, float* __restrict__ grho,
                                                         float* __restrict__ gu) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gg[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) gg[k] = g[i * si + k * sk];
  float gr, gx, gy;
  feq_vjp(rho[i], u[2 * i], u[2 * i + 1], gg, gr, gx, gy);
  if (grho) grho[i] = gr;
  if (gu) {
    gu[2 * i] = gx;
    gu[2 * i + 1] = gy;
  }
}

}  // namespace drb

using namespace drb;

static inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

extern "C" int drb_collide(float* f, int64_t n, int64_t stride_i, int64_t stride_k, float tau, drb_stream_t s) {
  if (n < 0 || (n > 0 && !f)) return fail(DRB_ERR_INVALID_ARG, "drb_collide: bad array");
  if (!(tau > 0.f)) return fail(DRB_ERR_INVALID_ARG, "drb_collide: tau must be > 0");
  if (n == 0) return DRB_OK;
  hipLaunchKernelGGL(k_collide_strided, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)s, f, (long long)n,
                     (long long)stride_i, (long long)stride_k, tau);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

extern "C" int drb_equilibrium(const float* rho, const float* u, int64_t n, float* out, int64_t stride_i,
                               int64_t stride_k, drb_stream_t s) {
  if (n < 0 || (n > 0 && (!rho || !u || !out))) return fail(DRB_ERR_INVALID_ARG, "drb_equilibrium: bad array");
  if (n == 0) return DRB_OK;
  hipLaunchKernelGGL(k_equilibrium_strided, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)s, rho, u,
                     (long long)n, out, (long long)stride_i, (long long)stride_k);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

extern "C" int drb_disperse(const float* pos_m, const float* f_m, int64_t m, float dx, float* pos_c, float* val_c,
                            int32_t* dir_c, drb_stream_t s) {
  if (m < 0 || (m > 0 && (!pos_m || !f_m || !pos_c || !val_c || !dir_c)))
    return fail(DRB_ERR_INVALID_ARG, "drb_disperse: bad array");
  if (m == 0) return DRB_OK;
  hipLaunchKernelGGL(k_disperse_aos, dim3(blocks_for(m)), dim3(256), 0, (hipStream_t)s, pos_m, f_m, (long long)m,
                     dx, pos_c, val_c, dir_c);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

extern "C" int drb_collide_vjp(const float* f, int64_t n, int64_t stride_i, int64_t stride_k, float tau,
                               const float* gout, float* gin, double* gtau_i, drb_stream_t s) {
  if (n < 0 || (n > 0 && (!f || !gout || !gin))) return fail(DRB_ERR_INVALID_ARG, "drb_collide_vjp: bad array");
  if (!(tau > 0.f)) return fail(DRB_ERR_INVALID_ARG, "drb_collide_vjp: tau must be > 0");
  if (n == 0) return DRB_OK;
  hipLaunchKernelGGL(k_collide_vjp, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)s, f, (long long)n,
                     (long long)stride_i, (long long)stride_k, tau, gout, gin, gtau_i);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

extern "C" int drb_equilibrium_vjp(const float* rho, const float* u, int64_t n, const float* g, int64_t stride_i,
                                   int64_t stride_k, float* grho, float* gu, drb_stream_t s) {
  if (n < 0 || (n > 0 && (!rho || !u || !g))) return fail(DRB_ERR_INVALID_ARG, "drb_equilibrium_vjp: bad array");
  if (n == 0) return DRB_OK;
  hipLaunchKernelGGL(k_equilibrium_vjp, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)s, rho, u, (long long)n, g,
                     (long long)stride_i, (long long)stride_k, grho, gu);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}
