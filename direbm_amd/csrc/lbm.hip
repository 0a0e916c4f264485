// D2Q7 hexagonal lattice Boltzmann on SoA planes f[k][j][i].
//
// Reference: direbm/warp/lbm.py:24-105 (_lbm_collide, _lbm_stream, _lbm_set_center,
// GpuHexLBM.step).  The reference runs two kernels per step over an (nj, ni, 7) AoS array:
// collide in place, then a pull stream out[j,i,k] = f[j - dj_k, i - di_k, k] (periodic).  Here
// one kernel per step reads a node's 7 populations once (coalesced per plane), collides them in
// registers and pushes each to its downstream neighbour: out[j + dj_k][i + di_k][k] = f*[j][i][k].
// That is the same map (collide then stream) with 28 B read + 28 B written per node.  Rows of a
// multiple of 4 nodes take k_lbm_step4 (four nodes per thread, 16-byte accesses).
#include "physics.h"

namespace drb {

__device__ __constant__ static const int kDi[7] = {0, 1, 0, -1, -1, 0, 1};
__device__ __constant__ static const int kDj[7] = {0, 0, 1, 1, 0, -1, -1};

// Four floats at a 4-byte-aligned address (one 16-byte access).
struct __attribute__((aligned(4))) Fl4u {
  float x, y, z, w;
};

__global__ __launch_bounds__(256) void k_lbm_step(const float* __restrict__ f, float* __restrict__ out, int ni, int nj,
                                                  float tau) {
  const long long n = (long long)ni * nj;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int j = (int)(g / ni), i = (int)(g - (long long)j * ni);
  float v[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) v[k] = f[k * n + g];
  unsigned st = 0;
  collide7(v, tau, st);
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    int ii = i + kDi[k], jj = j + kDj[k];
    ii = ii < 0 ? ii + ni : (ii >= ni ? ii - ni : ii);
    jj = jj < 0 ? jj + nj : (jj >= nj ? jj - nj : jj);
    out[k * n + (long long)jj * ni + ii] = v[k];
  }
}

// The same step, four nodes per thread (rows of a multiple of 4 nodes): each plane is read as one
// 16-byte vector per thread and written back shifted by its axial offset.  A write with di = 0 is
// an aligned vector; with di = +-1 it is four consecutive floats one node off (a 4-byte-aligned
// vector store), except at the row's end (di = +1) or start (di = -1), where the node past the edge
// wraps to the other end of the row and is written on its own.
__global__ __launch_bounds__(256) void k_lbm_step4(const float* __restrict__ f, float* __restrict__ out, int ni, int nj,
                                                   float tau) {
  const unsigned n = (unsigned)ni * (unsigned)nj;
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (n >> 2)) return;
  const unsigned g0 = t << 2;
  const int j = (int)(g0 / (unsigned)ni), i0 = (int)(g0 - (unsigned)j * (unsigned)ni);
  float v[4][7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const float4 a = *reinterpret_cast<const float4*>(f + (size_t)k * n + g0);
    v[0][k] = a.x;
    v[1][k] = a.y;
    v[2][k] = a.z;
    v[3][k] = a.w;
  }
  unsigned st = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) collide7(v[q], tau, st);
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    int jj = j + kDj[k];
    jj = jj < 0 ? jj + nj : (jj >= nj ? jj - nj : jj);
    float* row = out + (size_t)k * n + (size_t)jj * (unsigned)ni;
    const int di = kDi[k];
    if (di == 0) {
      *reinterpret_cast<float4*>(row + i0) = make_float4(v[0][k], v[1][k], v[2][k], v[3][k]);
    } else if (di > 0) {
      if (i0 + 4 < ni) {
        *reinterpret_cast<Fl4u*>(row + i0 + 1) = Fl4u{v[0][k], v[1][k], v[2][k], v[3][k]};
      } else {
        row[i0 + 1] = v[0][k];
        row[i0 + 2] = v[1][k];
        row[i0 + 3] = v[2][k];
        row[0] = v[3][k];
      }
    } else {
      if (i0 > 0) {
        *reinterpret_cast<Fl4u*>(row + i0 - 1) = Fl4u{v[0][k], v[1][k], v[2][k], v[3][k]};
      } else {
        row[ni - 1] = v[0][k];
        row[0] = v[1][k];
        row[1] = v[2][k];
        row[2] = v[3][k];
      }
    }
  }
}

__global__ void k_lbm_set_node(float* __restrict__ f, long long n, long long node, float rho) {
  const int k = threadIdx.x;
  if (k < 7) f[k * n + node] = fmul(rho, kW[k]);
}

__global__ __launch_bounds__(256) void k_lbm_init(float* __restrict__ f, const float* __restrict__ rho, long long n) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const float r = rho[g];
#pragma unroll
  for (int k = 0; k < 7; ++k) f[k * n + g] = feq_k(r, 0.f, 0.f, k);
}

}  // namespace drb

using namespace drb;

extern "C" int drb_lbm_step(float* f, float* out, int32_t ni, int32_t nj, float tau, int32_t nsteps, drb_stream_t s) {
  if (!f || !out || ni < 1 || nj < 1 || nsteps < 0 || !(tau > 0.f) || (long long)ni * nj > (1LL << 31))
    return fail(DRB_ERR_INVALID_ARG, "drb_lbm_step: bad arguments");
  const long long n = (long long)ni * nj;
  // rows of a multiple of 4 nodes: four nodes per thread, 16-byte accesses
  const bool quad = ni % 4 == 0;
  const unsigned nb = (unsigned)(((quad ? n / 4 : n) + 255) / 256);
  float* a = f;
  float* b = out;
  for (int t = 0; t < nsteps; ++t) {
    if (quad)
      hipLaunchKernelGGL(k_lbm_step4, dim3(nb), dim3(256), 0, (hipStream_t)s, a, b, ni, nj, tau);
    else
      hipLaunchKernelGGL(k_lbm_step, dim3(nb), dim3(256), 0, (hipStream_t)s, a, b, ni, nj, tau);
    float* tmp = a;
    a = b;
    b = tmp;
  }
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

extern "C" int drb_lbm_set_node(float* f, int32_t ni, int32_t nj, int32_t ci, int32_t cj, float rho, drb_stream_t s) {
  if (!f || ni < 1 || nj < 1 || ci < 0 || ci >= ni || cj < 0 || cj >= nj)
    return fail(DRB_ERR_INVALID_ARG, "drb_lbm_set_node: bad arguments");
  const long long n = (long long)ni * nj;
  hipLaunchKernelGGL(k_lbm_set_node, dim3(1), dim3(64), 0, (hipStream_t)s, f, n, (long long)cj * ni + ci, rho);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

extern "C" int drb_lbm_init(float* f, const float* rho, int32_t ni, int32_t nj, drb_stream_t s) {
  if (!f || !rho || ni < 1 || nj < 1) return fail(DRB_ERR_INVALID_ARG, "drb_lbm_init: bad arguments");
  const long long n = (long long)ni * nj;
  hipLaunchKernelGGL(k_lbm_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)s, f, rho, n);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}
