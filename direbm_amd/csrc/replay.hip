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

__device__ __forceinline__ float dmi(float a, float b, float L, int per) { return min_image(a - b, L, per); }

__global__ __launch_bounds__(TB) void k_rep_collide(int m, const float* __restrict__ f_in, const int* __restrict__ si,
                                                    const Inst* __restrict__ ins, float tau, float* __restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  float v[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) v[k] = f_in[7 * (size_t)i + k];
  unsigned st = 0;
  collide7(v, tau > 0.f ? tau : ins[si[i]].tau, st);
#pragma unroll
  for (int k = 0; k < 7; ++k) g[(size_t)k * m + i] = v[k];
}

// Component values in sorted order: position = recorded position + (new moment - recorded moment).
__global__ __launch_bounds__(TB) void k_rep_comp(int n, int m, const Comp* __restrict__ comp,
                                                 const uint32_t* __restrict__ vals, const float* __restrict__ sx,
                                                 const float* __restrict__ sy, const float* __restrict__ pos_in,
                                                 const float* __restrict__ g, const Inst* __restrict__ ins,
                                                 float2* __restrict__ rq, float* __restrict__ rv) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const Comp c = comp[s];
  const Inst& p = ins[c.bits >> 3];
  const uint32_t o = vals[s], i = o / 7u, k = o - 7u * i;
  rq[s] = make_float2(c.x + dmi(pos_in[2 * (size_t)i], sx[i], p.L, p.periodic),
                      c.y + dmi(pos_in[2 * (size_t)i + 1], sy[i], p.L, p.periodic));
  rv[s] = g[(size_t)k * m + i];
}

__global__ __launch_bounds__(TB) void k_rep_centroid(int P, const int* __restrict__ cps, const int* __restrict__ cpn,
                                                     const float2* __restrict__ rq, float2* __restrict__ ra) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P) return;
  const int a = cps[r], n = cpn[r];
  float x = 0.f, y = 0.f;
  for (int t = a; t < a + n; ++t) {
    x += rq[t].x;
    y += rq[t].y;
  }
  ra[r] = make_float2(x / (float)n, y / (float)n);
}

__global__ __launch_bounds__(TB) void k_rep_refine(int P, const float* __restrict__ cp0x, const float* __restrict__ cp0y,
                                                   const int* __restrict__ cpi, const Inst* __restrict__ ins,
                                                   const int* __restrict__ rp2, const Comp* __restrict__ comp,
                                                   const unsigned char* __restrict__ kap, const float2* __restrict__ rq,
                                                   const float* __restrict__ rv, const float2* __restrict__ ra,
                                                   float2* __restrict__ rx) {
  const int lane = threadIdx.x & (kGroup - 1);
  const int ngroups = gridDim.x * (blockDim.x / kGroup);
  const int rlast = ((P + ngroups - 1) / ngroups) * ngroups;
  for (int r = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup; r < rlast; r += ngroups) {
    const bool active = r < P;
    const Inst& p = ins[active ? cpi[r] : 0];
    const float x0 = active ? cp0x[r] : 0.f, y0 = active ? cp0y[r] : 0.f, t2 = p.t_dx, L = p.L;
    const int per = p.periodic;
    const float2 a = active ? ra[r] : make_float2(0.f, 0.f);
    float sw = 0.f, sx = 0.f, sy = 0.f;
    if (active)
      for_window(p, rp2, x0, y0, p.dx, 0, [&](int lo, int e) {
        for (int s = lo + lane; s < e; s += kGroup) {
          const Comp c = comp[s];
          const float ddx = min_image(fsub(c.x, x0), L, per), ddy = min_image(fsub(c.y, y0), L, per);
          if (fadd(fmul(ddx, ddx), fmul(ddy, ddy)) <= t2) {
            const float w = expf(rv[s]);
            const float2 q = rq[s];
            sw += w;
            sx += w * dmi(q.x, a.x, L, per);
            sy += w * dmi(q.y, a.y, L, per);
          }
        }
      });
    sw = group_sum(sw);
    sx = group_sum(sx);
    sy = group_sum(sy);
    if (!active || lane != 0) continue;
    rx[r] = kap[r] <= p.kappa_hard ? a : make_float2(a.x + sx / sw, a.y + sy / sw);
  }
}

// Scatter normaliser and q = v / S (components dropped in the recorded step stay dropped).
__global__ __launch_bounds__(TB) void k_rep_norm(int n, const Comp* __restrict__ comp, const float* __restrict__ swc,
                                                 const Inst* __restrict__ ins, const int* __restrict__ rp2,
                                                 const float* __restrict__ cpx, const float* __restrict__ cpy,
                                                 const int* __restrict__ disp, const float2* __restrict__ rq,
                                                 const float* __restrict__ rv, const float2* __restrict__ rx,
                                                 float* __restrict__ rcq) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = s < n;
  Comp c;
  c.x = c.y = 0.f;
  c.bits = 0;
  float2 q = make_float2(0.f, 0.f);
  if (active) {
    c = comp[s];
    q = rq[s];
  }
  const int b = c.bits >> 3;
  const Inst& p = ins[b];
  const float rr = p.rr, t2 = p.t_rr, L = p.L;
  const int per = p.periodic;
  float S = 0.f;
  for_wave_window(ins, rp2, active, b, c.x, c.y, rr, disp, [&](int j) {
    if (dist2_32(cpx[j], cpy[j], c.x, c.y, L, per) <= t2) {
      const float2 x = rx[j];
      const float dx = dmi(x.x, q.x, L, per), dy = dmi(x.y, q.y, L, per);
      S += rr - sqrtf(dx * dx + dy * dy);
    }
  });
  if (!active) return;
  rcq[s] = swc[s] > 0.f ? rv[s] / S : 0.f;
}

__global__ __launch_bounds__(TB) void k_rep_gather(int P, const float* __restrict__ cpx, const float* __restrict__ cpy,
                                                   const int* __restrict__ cpi, const Inst* __restrict__ ins,
                                                   const int* __restrict__ rp2, const Comp* __restrict__ comp,
                                                   const float2* __restrict__ rq,
                                                   const float* __restrict__ rcq, const float2* __restrict__ rx,
                                                   float* __restrict__ rf) {
  const int lane = threadIdx.x & (kGroup - 1);
  const int ngroups = gridDim.x * (blockDim.x / kGroup);
  const int rlast = ((P + ngroups - 1) / ngroups) * ngroups;
  for (int r = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup; r < rlast; r += ngroups) {
    const bool active = r < P;
    const Inst& p = ins[active ? cpi[r] : 0];
    const float x = active ? cpx[r] : 0.f, y = active ? cpy[r] : 0.f, rr = p.rr, t2 = p.t_rr, L = p.L;
    const int per = p.periodic;
    const float2 xr = active ? rx[r] : make_float2(0.f, 0.f);
    float acc[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) acc[k] = 0.f;
    if (active)
      for_window(p, rp2, x, y, rr, 0, [&](int lo, int e) {
        for (int s = lo + lane; s < e; s += kGroup) {
          const Comp c = comp[s];
          if (dist2_32(x, y, c.x, c.y, L, per) <= t2 && c.w != 0.f) {  // c.w = val / sw
            const float2 q = rq[s];
            const float dx = dmi(xr.x, q.x, L, per), dy = dmi(xr.y, q.y, L, per);
            const float t = (rr - sqrtf(dx * dx + dy * dy)) * rcq[s];
            const int kk = c.bits & 7;
#pragma unroll
            for (int k = 0; k < 7; ++k) acc[k] += k == kk ? t : 0.f;
          }
        }
      });
#pragma unroll
    for (int k = 0; k < 7; ++k) acc[k] = group_sum(acc[k]);
    if (!active || lane != 0) continue;
#pragma unroll
    for (int k = 0; k < 7; ++k) rf[(size_t)k * P + r] = acc[k];
  }
}

__global__ __launch_bounds__(TB) void k_rep_emit(int P, const float* __restrict__ cpx, const float* __restrict__ cpy,
                                                 const int* __restrict__ cpi, const Inst* __restrict__ ins,
                                                 const int* __restrict__ rp2, const int* __restrict__ disp,
                                                 const unsigned char* __restrict__ emask,
                                                 const float2* __restrict__ rx, const float* __restrict__ rf,
                                                 float* __restrict__ pos_out, float* __restrict__ f_out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P) return;
  const int b = cpi[r];
  const Inst& p = ins[b];
  const float x = cpx[r], y = cpy[r], rr = p.rr, t2 = p.t_rr, L = p.L;
  const int per = p.periodic;
  const float2 xr = rx[r];
  float out[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) out[k] = rf[(size_t)k * P + r];
  const unsigned e = emask[r];
  if (e) {
    float G[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) G[k] = 0.f;
    for_window(p, rp2, x, y, rr, disp[b], [&](int lo, int en) {
      for (int j = lo; j < en; ++j) {
        const unsigned m = e & emask[j];
        if (m && dist2_32(cpx[j], cpy[j], x, y, L, per) <= t2) {
          const float2 xj = rx[j];
          const float dx = dmi(xj.x, xr.x, L, per), dy = dmi(xj.y, xr.y, L, per);
          const float w = rr - (j == r ? 0.f : sqrtf(dx * dx + dy * dy));
#pragma unroll
          for (int k = 0; k < 7; ++k)
            if ((m >> k) & 1u) G[k] += w;
        }
      }
    });
#pragma unroll
    for (int k = 0; k < 7; ++k)
      if (((e >> k) & 1u) && G[k] > 0.f) out[k] = rr / G[k] * p.frest[k];
  }
  pos_out[2 * (size_t)r] = xr.x;
  pos_out[2 * (size_t)r + 1] = xr.y;
#pragma unroll
  for (int k = 0; k < 7; ++k) f_out[7 * (size_t)r + k] = out[k];
}

}  // namespace drb

using namespace drb;

int drb_sim::replay(int slot, float tau, const float* pos_in, const float* f_in, float* pos_out, float* f_out) {
  if (slot < 0 || slot >= tape_n) return fail(DRB_ERR_INVALID_ARG, "drb_sim_replay: bad slot");
  Work& w = tape[slot];
  if (!w.taped) return fail(DRB_ERR_STATE, "drb_sim_replay: slot not recorded");
  const long long P = w.P, N = w.N, Mi = w.M;
  if (Mi == 0) return DRB_OK;
  if (!pos_in || !f_in || (P > 0 && (!pos_out || !f_out))) return fail(DRB_ERR_INVALID_ARG, "drb_sim_replay: null array");
  AdjWork& a = adj;
  const size_t Pz = (size_t)std::max<long long>(P, 1);
  DRB_CHECK(a.rg.ensure((size_t)Mi * 28));
  DRB_CHECK(a.rq.ensure((size_t)N * 8));
  DRB_CHECK(a.rv.ensure((size_t)N * 4));
  DRB_CHECK(a.rcq.ensure((size_t)N * 4));
  DRB_CHECK(a.ra.ensure(Pz * 8));
  DRB_CHECK(a.rx.ensure(Pz * 8));
  DRB_CHECK(a.rf.ensure(Pz * 28));
  const Inst* I = w.d_inst.as<Inst>();
  const int* rp = w.rp.as<int>();
  hipLaunchKernelGGL(k_rep_collide, dim3(nblk(Mi)), dim3(TB), 0, s, (int)Mi, f_in, w.si.as<int>(), I, tau,
                     a.rg.as<float>());
  hipLaunchKernelGGL(k_rep_comp, dim3(nblk(N)), dim3(TB), 0, s, (int)N, (int)Mi, w.comp.as<Comp>(),
                     w.vals.as<uint32_t>(), w.sx.as<float>(), w.sy.as<float>(), pos_in, a.rg.as<float>(), I,
                     a.rq.as<float2>(), a.rv.as<float>());
  if (P > 0) {
    hipLaunchKernelGGL(k_rep_centroid, dim3(nblk(P)), dim3(TB), 0, s, (int)P, w.cps.as<int>(), w.cpn.as<int>(),
                       a.rq.as<float2>(), a.ra.as<float2>());
    hipLaunchKernelGGL(k_rep_refine, dim3(gblk(P)), dim3(TB), 0, s, (int)P, w.cp0x.as<float>(), w.cp0y.as<float>(),
                       w.cpi.as<int>(), I, rp, w.comp.as<Comp>(), w.kap.as<unsigned char>(), a.rq.as<float2>(),
                       a.rv.as<float>(), a.ra.as<float2>(), a.rx.as<float2>());
  }
  hipLaunchKernelGGL(k_rep_norm, dim3(nblk(N)), dim3(TB), 0, s, (int)N, w.comp.as<Comp>(), w.swc.as<float>(), I,
                     rp + 1, w.cpx.as<float>(), w.cpy.as<float>(), w.dsp.as<int>(), a.rq.as<float2>(),
                     a.rv.as<float>(), a.rx.as<float2>(), a.rcq.as<float>());
  if (P > 0) {
    hipLaunchKernelGGL(k_rep_gather, dim3(gblk(P)), dim3(TB), 0, s, (int)P, w.cpx.as<float>(), w.cpy.as<float>(),
                       w.cpi.as<int>(), I, rp, w.comp.as<Comp>(), a.rq.as<float2>(),
                       a.rcq.as<float>(), a.rx.as<float2>(), a.rf.as<float>());
    hipLaunchKernelGGL(k_rep_emit, dim3(nblk(P)), dim3(TB), 0, s, (int)P, w.cpx.as<float>(), w.cpy.as<float>(),
                       w.cpi.as<int>(), I, rp + 1, w.dsp.as<int>(), w.emask.as<unsigned char>(), a.rx.as<float2>(),
                       a.rf.as<float>(), pos_out, f_out);
  }
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

extern "C" int drb_sim_replay(drb_sim* sim, int32_t slot, float tau, const float* pos_in, const float* f_in,
                              float* pos_out, float* f_out) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim_replay: null handle");
  return sim->replay(slot, tau, pos_in, f_in, pos_out, f_out);
}
