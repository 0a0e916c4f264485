// Replay of a recorded step at another input state, with the recorded discrete choices.
//
// No reference counterpart.  The adjoint (adjoint.hip) differentiates the step with its discrete
// choices -- thinning runs, refine hard/soft decision, pairs within the radii (all decided on the
// recorded positions), empty directions, dropped components -- held fixed.  This evaluates that
// same function: every radius predicate and window is taken from the recorded step, every value
// (collide, component positions, centroids, exp-weighted refine, scatter weights and normalisers,
// gap-fill sums) from the new input.  Central differences of replayed rollouts therefore check
// the adjoint's gradient at full size without discrete jumps.  Positions are kept unwrapped
// (within a fraction of a cell of the recorded ones); periodic differences use the minimum image.
#include "engine.h"

namespace drb {

// Value arithmetic of the replay in T (float: the engine's own fp32 arithmetic; double: the same
// function evaluated in FP64, whose central differences are then exact to O(eps^2)).  Radius
// predicates, windows and every other discrete choice come from the recorded fp32 step either way.
__device__ __forceinline__ float tsqrt(float v) { return sqrtf(v); }
__device__ __forceinline__ double tsqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float texp(float v) { return expf(v); }
__device__ __forceinline__ double texp(double v) { return exp(v); }

template <class T>
__device__ __forceinline__ T dmi(T a, T b, T L, int per) {
  const T d = a - b, h = T(0.5) * L;
  if (!per) return d;
  return d > h ? d - L : (d < -h ? d + L : d);
}

// collide7 in T, for T = double: the same map, constants the engine's fp32 velocity set and
// weights, the rest equilibrium taking the remainder as in collide7.
__device__ __forceinline__ void collide7_t(double f[7], double tau) {
  double rho = 0.0, mx = 0.0, my = 0.0;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    rho += f[k];
    mx += f[k] * (double)kCx[k];
    my += f[k] * (double)kCy[k];
  }
  const double ux = rho != 0.0 ? mx / rho : 0.0, uy = rho != 0.0 ? my / rho : 0.0;
  const double usq = ux * ux + uy * uy;
  double fe[7], moving = 0.0;
#pragma unroll
  for (int k = 1; k < 7; ++k) {
    const double cu = (double)kCx[k] * ux + (double)kCy[k] * uy;
    fe[k] = rho * (double)kW[k] * (1.0 + 4.0 * cu + 8.0 * cu * cu - 2.0 * usq);
    moving += fe[k];
  }
  fe[0] = rho - moving;
#pragma unroll
  for (int k = 0; k < 7; ++k) f[k] += (fe[k] - f[k]) / tau;
}
__device__ __forceinline__ void collide7_t(float f[7], float tau) {
  unsigned st = 0;
  collide7(f, tau, st);
}

template <class T>
__global__ __launch_bounds__(TB) void k_rep_collide(int m, const T* __restrict__ f_in, const int* __restrict__ si,
                                                    const Inst* __restrict__ ins, T tau, T* __restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  T v[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) v[k] = f_in[7 * (size_t)i + k];
  collide7_t(v, tau > T(0) ? tau : (T)ins[si[i]].tau);
#pragma unroll
  for (int k = 0; k < 7; ++k) g[(size_t)k * m + i] = v[k];
}

// Component values in sorted order: position = recorded position + (new moment - recorded moment).
template <class T>
__global__ __launch_bounds__(TB) void k_rep_comp(int n, int m, const Comp* __restrict__ comp,
                                                 const uint32_t* __restrict__ vals, const float* __restrict__ sx,
                                                 const float* __restrict__ sy, const T* __restrict__ pos_in,
                                                 const T* __restrict__ g, const Inst* __restrict__ ins,
                                                 T* __restrict__ rq, T* __restrict__ rv) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const Comp c = comp[s];
  const Inst& p = ins[c.bits >> 3];
  const uint32_t o = vals[s], i = o / 7u, k = o - 7u * i;
  const T L = (T)p.L;
  rq[2 * (size_t)s] = (T)c.x + dmi<T>(pos_in[2 * (size_t)i], (T)sx[i], L, p.periodic);
  rq[2 * (size_t)s + 1] = (T)c.y + dmi<T>(pos_in[2 * (size_t)i + 1], (T)sy[i], L, p.periodic);
  rv[s] = g[(size_t)k * m + i];
}

template <class T>
__global__ __launch_bounds__(TB) void k_rep_centroid(int P, const int* __restrict__ cps, const int* __restrict__ cpn,
                                                     const T* __restrict__ rq, T* __restrict__ ra) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P) return;
  const int a = cps[r], n = cpn[r];
  T x = T(0), y = T(0);
  for (int t = a; t < a + n; ++t) {
    x += rq[2 * (size_t)t];
    y += rq[2 * (size_t)t + 1];
  }
  ra[2 * (size_t)r] = x / (T)n;
  ra[2 * (size_t)r + 1] = y / (T)n;
}

template <class T>
__global__ __launch_bounds__(TB) void k_rep_refine(int P, const float* __restrict__ cp0x, const float* __restrict__ cp0y,
                                                   const int* __restrict__ cpi, const Inst* __restrict__ ins,
                                                   const int* __restrict__ rp2, const Comp* __restrict__ comp,
                                                   const unsigned char* __restrict__ kap, const T* __restrict__ rq,
                                                   const T* __restrict__ rv, const T* __restrict__ ra,
                                                   T* __restrict__ rx) {
  const int lane = threadIdx.x & (kGroup - 1);
  const int ngroups = gridDim.x * (blockDim.x / kGroup);
  const int rlast = ((P + ngroups - 1) / ngroups) * ngroups;
  for (int r = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup; r < rlast; r += ngroups) {
    const bool active = r < P;
    const Inst& p = ins[active ? cpi[r] : 0];
    const float x0 = active ? cp0x[r] : 0.f, y0 = active ? cp0y[r] : 0.f, t2 = p.t_dx, L = p.L;
    const int per = p.periodic;
    const T ax = active ? ra[2 * (size_t)r] : T(0), ay = active ? ra[2 * (size_t)r + 1] : T(0);
    T sw = T(0), sx = T(0), sy = T(0);
    if (active)
      for_window(p, rp2, x0, y0, p.dx, 0, [&](int lo, int e) {
        for (int s = lo + lane; s < e; s += kGroup) {
          const Comp c = comp[s];
          const float ddx = min_image(fsub(c.x, x0), L, per), ddy = min_image(fsub(c.y, y0), L, per);
          if (fadd(fmul(ddx, ddx), fmul(ddy, ddy)) <= t2) {
            const T w = texp(rv[s]);
            sw += w;
            sx += w * dmi<T>(rq[2 * (size_t)s], ax, (T)L, per);
            sy += w * dmi<T>(rq[2 * (size_t)s + 1], ay, (T)L, per);
          }
        }
      });
    sw = group_sum_t(sw);
    sx = group_sum_t(sx);
    sy = group_sum_t(sy);
    if (!active || lane != 0) continue;
    const bool hard = kap[r] <= p.kappa_hard;
    rx[2 * (size_t)r] = hard ? ax : ax + sx / sw;
    rx[2 * (size_t)r + 1] = hard ? ay : ay + sy / sw;
  }
}

// Scatter normaliser and q = v / S (components dropped in the recorded step stay dropped).
template <class T>
__global__ __launch_bounds__(TB) void k_rep_norm(int n, const Comp* __restrict__ comp, const float* __restrict__ swc,
                                                 const Inst* __restrict__ ins, const int* __restrict__ rp2,
                                                 const float* __restrict__ cpx, const float* __restrict__ cpy,
                                                 const int* __restrict__ disp, const T* __restrict__ rq,
                                                 const T* __restrict__ rv, const T* __restrict__ rx,
                                                 T* __restrict__ rcq) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = s < n;
  Comp c;
  c.x = c.y = 0.f;
  c.bits = 0;
  T qx = T(0), qy = T(0);
  if (active) {
    c = comp[s];
    qx = rq[2 * (size_t)s];
    qy = rq[2 * (size_t)s + 1];
  }
  const int b = c.bits >> 3;
  const Inst& p = ins[b];
  const float t2 = p.t_rr, L = p.L;
  const T rr = (T)p.rr;
  const int per = p.periodic;
  T S = T(0);
  for_wave_window(ins, rp2, active, b, c.x, c.y, p.rr, disp, [&](int j) {
    if (dist2_32(cpx[j], cpy[j], c.x, c.y, L, per) <= t2) {
      const T dx = dmi<T>(rx[2 * (size_t)j], qx, (T)L, per), dy = dmi<T>(rx[2 * (size_t)j + 1], qy, (T)L, per);
      S += rr - tsqrt(dx * dx + dy * dy);
    }
  });
  if (!active) return;
  rcq[s] = swc[s] > 0.f ? rv[s] / S : T(0);
}

template <class T>
__global__ __launch_bounds__(TB) void k_rep_gather(int P, const float* __restrict__ cpx, const float* __restrict__ cpy,
                                                   const int* __restrict__ cpi, const Inst* __restrict__ ins,
                                                   const int* __restrict__ rp2, const Comp* __restrict__ comp,
                                                   const T* __restrict__ rq, const T* __restrict__ rcq,
                                                   const T* __restrict__ rx, T* __restrict__ rf) {
  const int lane = threadIdx.x & (kGroup - 1);
  const int ngroups = gridDim.x * (blockDim.x / kGroup);
  const int rlast = ((P + ngroups - 1) / ngroups) * ngroups;
  for (int r = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup; r < rlast; r += ngroups) {
    const bool active = r < P;
    const Inst& p = ins[active ? cpi[r] : 0];
    const float x = active ? cpx[r] : 0.f, y = active ? cpy[r] : 0.f, t2 = p.t_rr, L = p.L;
    const T rr = (T)p.rr;
    const int per = p.periodic;
    const T xrx = active ? rx[2 * (size_t)r] : T(0), xry = active ? rx[2 * (size_t)r + 1] : T(0);
    T acc[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) acc[k] = T(0);
    if (active)
      for_window(p, rp2, x, y, p.rr, 0, [&](int lo, int e) {
        for (int s = lo + lane; s < e; s += kGroup) {
          const Comp c = comp[s];
          if (dist2_32(x, y, c.x, c.y, L, per) <= t2 && c.w != 0.f) {  // c.w = val / sw
            const T dx = dmi<T>(xrx, rq[2 * (size_t)s], (T)L, per), dy = dmi<T>(xry, rq[2 * (size_t)s + 1], (T)L, per);
            const T t = (rr - tsqrt(dx * dx + dy * dy)) * rcq[s];
            const int kk = c.bits & 7;
#pragma unroll
            for (int k = 0; k < 7; ++k) acc[k] += k == kk ? t : T(0);
          }
        }
      });
#pragma unroll
    for (int k = 0; k < 7; ++k) acc[k] = group_sum_t(acc[k]);
    if (!active || lane != 0) continue;
#pragma unroll
    for (int k = 0; k < 7; ++k) rf[(size_t)k * P + r] = acc[k];
  }
}

template <class T>
__global__ __launch_bounds__(TB) void k_rep_emit(int P, const float* __restrict__ cpx, const float* __restrict__ cpy,
                                                 const int* __restrict__ cpi, const Inst* __restrict__ ins,
                                                 const int* __restrict__ rp2, const int* __restrict__ disp,
                                                 const unsigned char* __restrict__ emask, const T* __restrict__ rx,
                                                 const T* __restrict__ rf, T* __restrict__ pos_out,
                                                 T* __restrict__ f_out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P) return;
  const int b = cpi[r];
  const Inst& p = ins[b];
  const float x = cpx[r], y = cpy[r], t2 = p.t_rr, L = p.L;
  const T rr = (T)p.rr;
  const int per = p.periodic;
  const T xrx = rx[2 * (size_t)r], xry = rx[2 * (size_t)r + 1];
  T out[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) out[k] = rf[(size_t)k * P + r];
  const unsigned e = emask[r];
  if (e) {
    T G[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) G[k] = T(0);
    for_window(p, rp2, x, y, p.rr, disp[b], [&](int lo, int en) {
      for (int j = lo; j < en; ++j) {
        const unsigned m = e & emask[j];
        if (m && dist2_32(cpx[j], cpy[j], x, y, L, per) <= t2) {
          const T dx = dmi<T>(rx[2 * (size_t)j], xrx, (T)L, per), dy = dmi<T>(rx[2 * (size_t)j + 1], xry, (T)L, per);
          const T w = rr - (j == r ? T(0) : tsqrt(dx * dx + dy * dy));
#pragma unroll
          for (int k = 0; k < 7; ++k)
            if ((m >> k) & 1u) G[k] += w;
        }
      }
    });
#pragma unroll
    for (int k = 0; k < 7; ++k)
      if (((e >> k) & 1u) && G[k] > T(0)) out[k] = rr / G[k] * (T)p.frest[k];
  }
  pos_out[2 * (size_t)r] = xrx;
  pos_out[2 * (size_t)r + 1] = xry;
#pragma unroll
  for (int k = 0; k < 7; ++k) f_out[7 * (size_t)r + k] = out[k];
}

}  // namespace drb

using namespace drb;

template <class T>
int drb_sim::replay_t(int slot, T tau, const T* pos_in, const T* f_in, T* pos_out, T* f_out) {
  if (slot < 0 || slot >= tape_n) return fail(DRB_ERR_INVALID_ARG, "drb_sim_replay: bad slot");
  Work& w = tape[slot];
  if (!w.taped) return fail(DRB_ERR_STATE, "drb_sim_replay: slot not recorded");
  const long long P = w.P, N = w.N, Mi = w.M;
  if (Mi == 0) return DRB_OK;
  if (!pos_in || !f_in || (P > 0 && (!pos_out || !f_out))) return fail(DRB_ERR_INVALID_ARG, "drb_sim_replay: null array");
  AdjWork& a = adj;
  const size_t Pz = (size_t)std::max<long long>(P, 1), z = sizeof(T);
  DRB_CHECK(a.rg.ensure((size_t)Mi * 7 * z));
  DRB_CHECK(a.rq.ensure((size_t)N * 2 * z));
  DRB_CHECK(a.rv.ensure((size_t)N * z));
  DRB_CHECK(a.rcq.ensure((size_t)N * z));
  DRB_CHECK(a.ra.ensure(Pz * 2 * z));
  DRB_CHECK(a.rx.ensure(Pz * 2 * z));
  DRB_CHECK(a.rf.ensure(Pz * 7 * z));
  const Inst* I = w.d_inst.as<Inst>();
  const int* rp = w.rp.as<int>();
  hipLaunchKernelGGL(k_rep_collide<T>, dim3(nblk(Mi)), dim3(TB), 0, s, (int)Mi, f_in, w.si.as<int>(), I, tau,
                     a.rg.as<T>());
  hipLaunchKernelGGL(k_rep_comp<T>, dim3(nblk(N)), dim3(TB), 0, s, (int)N, (int)Mi, w.comp.as<Comp>(),
                     w.vals.as<uint32_t>(), w.sx.as<float>(), w.sy.as<float>(), pos_in, a.rg.as<T>(), I, a.rq.as<T>(),
                     a.rv.as<T>());
  if (P > 0) {
    hipLaunchKernelGGL(k_rep_centroid<T>, dim3(nblk(P)), dim3(TB), 0, s, (int)P, w.cps.as<int>(), w.cpn.as<int>(),
                       a.rq.as<T>(), a.ra.as<T>());
    hipLaunchKernelGGL(k_rep_refine<T>, dim3(gblk(P)), dim3(TB), 0, s, (int)P, w.cp0x.as<float>(), w.cp0y.as<float>(),
                       w.cpi.as<int>(), I, rp, w.comp.as<Comp>(), w.kap.as<unsigned char>(), a.rq.as<T>(), a.rv.as<T>(),
                       a.ra.as<T>(), a.rx.as<T>());
  }
  hipLaunchKernelGGL(k_rep_norm<T>, dim3(nblk(N)), dim3(TB), 0, s, (int)N, w.comp.as<Comp>(), w.swc.as<float>(), I,
                     rp + 1, w.cpx.as<float>(), w.cpy.as<float>(), w.dsp.as<int>(), a.rq.as<T>(), a.rv.as<T>(),
                     a.rx.as<T>(), a.rcq.as<T>());
  if (P > 0) {
    hipLaunchKernelGGL(k_rep_gather<T>, dim3(gblk(P)), dim3(TB), 0, s, (int)P, w.cpx.as<float>(), w.cpy.as<float>(),
                       w.cpi.as<int>(), I, rp, w.comp.as<Comp>(), a.rq.as<T>(), a.rcq.as<T>(), a.rx.as<T>(),
                       a.rf.as<T>());
    hipLaunchKernelGGL(k_rep_emit<T>, dim3(nblk(P)), dim3(TB), 0, s, (int)P, w.cpx.as<float>(), w.cpy.as<float>(),
                       w.cpi.as<int>(), I, rp + 1, w.dsp.as<int>(), w.emask.as<unsigned char>(), a.rx.as<T>(),
                       a.rf.as<T>(), pos_out, f_out);
  }
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

int drb_sim::replay(int slot, float tau, const float* pos_in, const float* f_in, float* pos_out, float* f_out) {
  return replay_t<float>(slot, tau, pos_in, f_in, pos_out, f_out);
}
int drb_sim::replay64(int slot, double tau, const double* pos_in, const double* f_in, double* pos_out,
                      double* f_out) {
  return replay_t<double>(slot, tau, pos_in, f_in, pos_out, f_out);
}

extern "C" int drb_sim_replay(drb_sim* sim, int32_t slot, float tau, const float* pos_in, const float* f_in,
                              float* pos_out, float* f_out) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim_replay: null handle");
  return sim->replay(slot, tau, pos_in, f_in, pos_out, f_out);
}

extern "C" int drb_sim_replay64(drb_sim* sim, int32_t slot, double tau, const double* pos_in, const double* f_in,
                                double* pos_out, double* f_out) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim_replay64: null handle");
  return sim->replay64(slot, tau, pos_in, f_in, pos_out, f_out);
}
