// Adjoint (vector-Jacobian product) of one recorded DiReBM step.
//
// No reference counterpart: the reference step (direbm/warp/simulator.py:50-73) is not
// differentiated.  The discrete choices of the step -- the thinning runs, the refine hard/soft
// decision, which pairs lie within the radii, which directions are empty -- are held fixed, as
// the step recorded them; every continuous dependence is differentiated exactly:
//
//   collide      g = f + (f_eq(rho(f), u(f)) - f) / tau                       (physics.h)
//   disperse     q_c = p_i + C_k dx,  v_c = g_ik                               (c = 7 i + k)
//   centroid     a_r = sum_{c in run r} q_c / n_r
//   refine       x_r = a_r (kappa <= kappa_hard), else sum_c w_c q_c / W_r,  w_c = exp(v_c)
//   normaliser   S_c = sum_{r: |x_r - q_c| <= R} (R - d_rc),   cq_c = v_c / S_c
//   gather       F_rk = sum_{c: dir k, |x_r - q_c| <= R} (R - d_rc) cq_c
//   gap-fill     f'_rk = R frest_k / G_rk (k empty), G_rk = sum_{j empty in k} (R - d_jr); else F_rk
//   emit         p'_r = x_r
//
// The backward pass runs as six kernels, every one a gather over the same cell windows the
// forward walked (no atomics, so results are reproducible):
//   k_adj_emit   per control point: split the output cotangent into dF and dG
//   k_adj_gap    per control point with a gap: position cotangent of the gap-fill sums
//   k_adj_comp   per component (over control points): d cq, d S -> d v, d q (scatter terms)
//   k_adj_cp     per control point (over components): d x_r of the scatter, then refine/centroid
//   k_adj_refine per component (over control points, pre-refine positions): refine and centroid
//                terms, written to the component's original slot
//   k_adj_moment per moment: sum of its 7 components, collide adjoint, d tau
#include "engine.h"

namespace drb {

__device__ __forceinline__ float delta_mi(float a, float b, float L, int per) { return min_image(fsub(a, b), L, per); }

// Output cotangent -> dF (kept directions; 8 floats per control point, the last one 0, so a wave
// reads one point's seven values with one scalar load) and dG (gap-filled directions, SoA pitch P).
__global__ __launch_bounds__(TB) void k_adj_emit(int P, const int* __restrict__ cpi, const Inst* __restrict__ ins,
                                                 const float* __restrict__ gpos, const float* __restrict__ gf,
                                                 const unsigned char* __restrict__ emask, const float* __restrict__ cpg,
                                                 long long gpitch, float* __restrict__ fbar, float* __restrict__ gbar,
                                                 float* __restrict__ xb) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P) return;
  const Inst& p = ins[cpi[r]];
  const unsigned e = emask[r];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const float g = gf[7 * (size_t)r + k];
    const float G = cpg[k * gpitch + r];
    const bool filled = ((e >> k) & 1u) && G > 0.f;
    // f' = (R / G) frest  ->  df'/dG = -f' / G
    const float out = filled ? fmul(fdiv(p.rr, G), p.frest[k]) : 0.f;
    fbar[8 * (size_t)r + k] = filled ? 0.f : g;
    gbar[(size_t)k * P + r] = filled ? -g * out / G : 0.f;
  }
  fbar[8 * (size_t)r + 7] = 0.f;
  xb[2 * (size_t)r] = gpos ? gpos[2 * (size_t)r] : 0.f;
  xb[2 * (size_t)r + 1] = gpos ? gpos[2 * (size_t)r + 1] : 0.f;
}

// Position cotangent of the gap-fill sums.  For a pair (r, j) within R that are both empty in
// direction k, G_rk and G_jk both contain (R - d_rj), so
//   dx_r -= sum_j (x_r - x_j) / d_rj * sum_k E_rk E_jk (dG_rk + dG_jk).
template <int PM>
__global__ __launch_bounds__(TB) void k_adj_gap(int P, const float* __restrict__ cpx, const float* __restrict__ cpy,
                                                const int* __restrict__ cpi, const Inst* __restrict__ ins,
                                                const int* __restrict__ rp2, const int* __restrict__ disp,
                                                const unsigned char* __restrict__ emask,
                                                const float* __restrict__ gbar, float* __restrict__ xb) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P) return;
  const unsigned e = emask[r];
  if (!e) return;
  const int b = cpi[r];
  const Inst& p = ins[b];
  const float x = cpx[r], y = cpy[r], t2 = p.t_rr, L = p.L;
  const int per = PM == 2 ? p.periodic : PM;
  float own[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) own[k] = gbar[(size_t)k * P + r];
  float ax = 0.f, ay = 0.f;
  for_window(p, rp2, x, y, p.rr, disp[b], [&](int a, int en) {
    for (int j = a; j < en; ++j) {
      const unsigned m = e & emask[j];
      if (!m || j == r) continue;
      const float s2 = dist2_32(cpx[j], cpy[j], x, y, L, per);
      if (s2 > t2 || s2 <= 0.f) continue;
      float coef = 0.f;
#pragma unroll
      for (int k = 0; k < 7; ++k)
        if ((m >> k) & 1u) coef += gbar[(size_t)k * P + j] + own[k];
      const float id = fast_rsq(s2);
      ax += coef * delta_mi(x, cpx[j], L, per) * id;
      ay += coef * delta_mi(y, cpy[j], L, per) * id;
    }
  });
  xb[2 * (size_t)r] -= ax;
  xb[2 * (size_t)r + 1] -= ay;
}

// Per component: dcq = sum_r dF_{r,k} (R - d), then dv = dcq / S, dS = -dcq cq / S and
// dq = -(cq T + dS U) with T = sum_r dF_{r,k} (q - x_r) / d, U = sum_r (q - x_r) / d.
template <int PM>
__global__ __launch_bounds__(TB) void k_adj_comp(int n, const Comp* __restrict__ comp, const float* __restrict__ swc,
                                                 const Inst* __restrict__ ins,
                                                 const int* __restrict__ rp2, const float* __restrict__ cpx,
                                                 const float* __restrict__ cpy, const int* __restrict__ disp, int P,
                                                 const float* __restrict__ fbar, float* __restrict__ sbar,
                                                 float* __restrict__ vbs, float2* __restrict__ qbs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = s < n;
  Comp c;
  c.x = c.y = 0.f;
  c.bits = 0;
  if (active) c = comp[s];
  const int b = c.bits >> 3, k = c.bits & 7;
  const Inst& p = ins[b];
  const float rr = p.rr, t2 = p.t_rr, L = p.L;
  const int per = PM == 2 ? p.periodic : PM;
  const float4* __restrict__ f4 = reinterpret_cast<const float4*>(fbar);
  float cqb = 0.f, tx = 0.f, ty = 0.f, ux = 0.f, uy = 0.f;
  for_wave_window(ins, rp2, active, b, c.x, c.y, rr, disp, [&](int j) {
    const float s2 = dist2_32(cpx[j], cpy[j], c.x, c.y, L, per);
    if (s2 <= t2) {
      const float d = fast_sqrt(s2);
      // the point's seven dF values (uniform j: scalar loads), this lane's direction selected
      const float4 lo = f4[2 * (size_t)j], hi = f4[2 * (size_t)j + 1];
      float F = lo.x;
      F = k == 1 ? lo.y : F;
      F = k == 2 ? lo.z : F;
      F = k == 3 ? lo.w : F;
      F = k == 4 ? hi.x : F;
      F = k == 5 ? hi.y : F;
      F = k == 6 ? hi.z : F;
      cqb += F * (rr - d);
      if (d > 0.f) {
        const float id = fast_rcp(d);
        const float ex = delta_mi(c.x, cpx[j], L, per) * id, ey = delta_mi(c.y, cpy[j], L, per) * id;
        tx += F * ex;
        ty += F * ey;
        ux += ex;
        uy += ey;
      }
    }
  });
  if (!active) return;
  const float S = swc[s];
  float vb = 0.f, sb = 0.f, qx = 0.f, qy = 0.f;
  if (S > 0.f) {
    const float q = c.w;  // val / sw (written by the forward's normaliser)
    vb = cqb / S;
    sb = -cqb * q / S;
    qx = -(q * tx + sb * ux);
    qy = -(q * ty + sb * uy);
  }
  sbar[s] = sb;
  vbs[s] = vb;
  qbs[s] = make_float2(qx, qy);
}

// Per control point (a group of kGroup lanes): dx_r -= sum_c (dF_{r,k(c)} cq_c + dS_c) (x_r - q_c) / d,
// then the refine adjoint: a hard point passes dx_r to its centroid (dA_r / n_r per member); a
// refined point passes z_r = dx_r / W_r to the components of its refine disk (z_r = 0 when hard).
template <int PM>
__global__ __launch_bounds__(TB) void k_adj_cp(int P, const float* __restrict__ cpx, const float* __restrict__ cpy,
                                               const int* __restrict__ cpi, const Inst* __restrict__ ins,
                                               const int* __restrict__ rp2, const Comp* __restrict__ comp,
                                               const float* __restrict__ sbar,
                                               const float* __restrict__ fbar, const float* __restrict__ xb,
                                               const unsigned char* __restrict__ kap, const float* __restrict__ cpw,
                                               const int* __restrict__ cpn, const float* __restrict__ cp0x,
                                               const float* __restrict__ cp0y, float2* __restrict__ abn,
                                               float4* __restrict__ z) {
  const int lane = threadIdx.x & (kGroup - 1);
  const int ngroups = gridDim.x * (blockDim.x / kGroup);
  const int rlast = ((P + ngroups - 1) / ngroups) * ngroups;
  for (int r = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup; r < rlast; r += ngroups) {
    const bool active = r < P;
    const Inst& p = ins[active ? cpi[r] : 0];
    const float x = active ? cpx[r] : 0.f, y = active ? cpy[r] : 0.f, rr = p.rr, t2 = p.t_rr, L = p.L;
    const int per = PM == 2 ? p.periodic : PM;
    float F[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) F[k] = active ? fbar[8 * (size_t)r + k] : 0.f;
    float ax = 0.f, ay = 0.f;
    if (active)
      for_window(p, rp2, x, y, rr, 0, [&](int a, int e) {
        // kInflight candidates per lane in flight (records loaded before use, added in order)
        for (int s0 = a + lane; s0 < e; s0 += kInflight * kGroup) {
          Comp cc[kInflight];
          float sb[kInflight];
#pragma unroll
          for (int u = 0; u < kInflight; ++u) {
            const int s = min(s0 + u * kGroup, e - 1);
            cc[u] = comp[s];  // .w = q = val / sw (written by the forward's normaliser)
            sb[u] = sbar[s];
          }
#pragma unroll
          for (int u = 0; u < kInflight; ++u) {
            const Comp& c = cc[u];
            const float s2 = dist2_32(x, y, c.x, c.y, L, per);
            const float q = c.w;
            // branch-free: a rejected candidate adds +0 (its divisor is replaced by 1)
            const bool ok = s0 + u * kGroup < e && s2 <= t2 && q != 0.f && s2 > 0.f;
            const int kk = c.bits & 7;
            float fk = 0.f;
#pragma unroll
            for (int k = 0; k < 7; ++k) fk = k == kk ? F[k] : fk;
            const float coef = ok ? (fk * q + sb[u]) * fast_rsq(ok ? s2 : 1.f) : 0.f;
            ax += coef * delta_mi(x, c.x, L, per);
            ay += coef * delta_mi(y, c.y, L, per);
          }
        }
      });
    ax = group_sum(ax);
    ay = group_sum(ay);
    if (!active || lane != 0) continue;
    const float gx = xb[2 * (size_t)r] - ax, gy = xb[2 * (size_t)r + 1] - ay;
    // z[r] = (pre-refine position, dx_r / W_r): one 16-byte record per point for k_adj_refine
    if (kap[r] <= p.kappa_hard) {
      const float in = 1.f / (float)cpn[r];
      abn[r] = make_float2(gx * in, gy * in);
      z[r] = make_float4(cp0x[r], cp0y[r], 0.f, 0.f);
    } else {
      const float iw = 1.f / cpw[r];
      abn[r] = make_float2(0.f, 0.f);
      z[r] = make_float4(cp0x[r], cp0y[r], gx * iw, gy * iw);
    }
  }
}

// Per component: refine terms dq += w_c z_r, dv += w_c (q_c - x_r) . z_r over refined points r whose
// pre-refine position a_r lies within dx (the forward refine predicate), the centroid term
// dA_run / n_run, plus the scatter terms of k_adj_comp; written to the component's original slot.
template <int PM>
__global__ __launch_bounds__(TB) void k_adj_refine(int n, const Comp* __restrict__ comp, const float* __restrict__ cv,
                                                   const uint32_t* __restrict__ vals,
                                                   const int* __restrict__ crun, const Inst* __restrict__ ins,
                                                   const int* __restrict__ rp2, const float* __restrict__ cpx,
                                                   const float* __restrict__ cpy, const float4* __restrict__ z,
                                                   const float2* __restrict__ abn, const float* __restrict__ vbs,
                                                   const float2* __restrict__ qbs, float* __restrict__ gv,
                                                   float2* __restrict__ gq) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = s < n;
  Comp c;
  c.x = c.y = 0.f;
  c.w = 0.f;
  c.bits = 0;
  if (active) c = comp[s];
  const int b = c.bits >> 3;
  const Inst& p = ins[b];
  const float t2 = p.t_dx, L = p.L;
  const int per = PM == 2 ? p.periodic : PM;
  float zx = 0.f, zy = 0.f, vz = 0.f;
  for_wave_window(ins, rp2, active, b, c.x, c.y, p.dx, nullptr, [&](int j) {
    // forward refine: ddx = min_image(q - a_r) and the dx threshold; hard points have z = 0
    const float4 zj = z[j];
    const float ddx = delta_mi(c.x, zj.x, L, per), ddy = delta_mi(c.y, zj.y, L, per);
    if (fadd(fmul(ddx, ddx), fmul(ddy, ddy)) <= t2 && (zj.z != 0.f || zj.w != 0.f)) {
      zx += zj.z;
      zy += zj.w;
      vz += delta_mi(c.x, cpx[j], L, per) * zj.z + delta_mi(c.y, cpy[j], L, per) * zj.w;
    }
  });
  if (!active) return;
  const float2 a = abn[crun[s]];
  const float2 qs = qbs[s];
  const uint32_t o = vals[s];
  // the refine weight exp(val), as k_mark computed it (the record's slot holds val / sw by now)
  const float w = expf(cv[s]);
  gq[o] = make_float2(qs.x + a.x + w * zx, qs.y + a.y + w * zy);
  gv[o] = vbs[s] + w * vz;
}

// Per moment: position cotangent = sum of its components', then the collide adjoint with the
// recorded pre-collision f (SoA, pitch m).  tt[i] receives the moment's d tau.
__global__ __launch_bounds__(TB) void k_adj_moment(int m, const float* __restrict__ sf, const int* __restrict__ si,
                                                   const Inst* __restrict__ ins, const float* __restrict__ gv,
                                                   const float2* __restrict__ gq, float* __restrict__ gpos,
                                                   float* __restrict__ gf, double* __restrict__ tt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const Inst& p = ins[si[i]];
  float f[7], g[7], r[7];
  float px = 0.f, py = 0.f;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    f[k] = sf[(size_t)k * m + i];
    g[k] = gv[7 * (size_t)i + k];
    const float2 q = gq[7 * (size_t)i + k];
    px += q.x;
    py += q.y;
  }
  float gt;
  collide7_vjp(f, p.tau, g, r, gt);
  if (gpos) {
    gpos[2 * (size_t)i] = px;
    gpos[2 * (size_t)i + 1] = py;
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) gf[7 * (size_t)i + k] = r[k];
  tt[i] = (double)gt;
}

// Per-instance FP64 sum of tt over the instance's moments: kTauChunks blocks per instance each
// reduce a slice with a fixed tree, then one thread per instance adds the slices in order.
constexpr int kTauChunks = 64;

__global__ __launch_bounds__(TB) void k_tau_partial(const double* __restrict__ tt, const long long* __restrict__ moff,
                                                    double* __restrict__ part) {
  __shared__ double sh[TB];
  const int b = blockIdx.y, chunk = blockIdx.x, t = threadIdx.x;
  const long long lo = moff[b], hi = moff[b + 1];
  const long long per = (hi - lo + kTauChunks - 1) / kTauChunks;
  const long long s0 = min(hi, lo + per * chunk), s1 = min(hi, s0 + per);
  double a = 0.0;
  for (long long i = s0 + t; i < s1; i += TB) a += tt[i];
  sh[t] = a;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (t < w) sh[t] += sh[t + w];
    __syncthreads();
  }
  if (t == 0) part[(size_t)b * kTauChunks + chunk] = sh[0];
}

__global__ void k_tau_final(const double* __restrict__ part, int ninst, double* __restrict__ acc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ninst) return;
  double s = 0.0;
  for (int c = 0; c < kTauChunks; ++c) s += part[(size_t)b * kTauChunks + c];
  acc[b] += s;
}

__global__ void k_zero_f64(double* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

}  // namespace drb

using namespace drb;

// Launch the adjoint kernels of one recorded step; d tau of each instance is added to the device
// accumulator tacc[ninst].
int drb_sim::vjp_launch(int slot, const float* gpos_out, const float* gf_out, float* gpos_in, float* gf_in,
                        double* tacc) {
  Work& w = tape[slot];
  const long long P = w.P, N = w.N, Mi = w.M;
  if (Mi == 0) return DRB_OK;
  AdjWork& a = adj;
  const size_t Pz = (size_t)std::max<long long>(P, 1);
  DRB_CHECK(a.fbar.ensure(Pz * 8 * 4));
  DRB_CHECK(a.gbar.ensure(Pz * 7 * 4));
  DRB_CHECK(a.xb.ensure(Pz * 8));
  DRB_CHECK(a.ab.ensure(Pz * 8));
  DRB_CHECK(a.z.ensure(Pz * 16));
  DRB_CHECK(a.sbar.ensure((size_t)N * 4));
  DRB_CHECK(a.vbs.ensure((size_t)N * 4));
  DRB_CHECK(a.qbs.ensure((size_t)N * 8));
  DRB_CHECK(a.gv.ensure((size_t)N * 4));
  DRB_CHECK(a.gq.ensure((size_t)N * 8));
  DRB_CHECK(a.tt.ensure((size_t)Mi * 8));
  DRB_CHECK(a.tpart.ensure(sizeof(double) * kTauChunks * ninst));
  const Inst* I = w.d_inst.as<Inst>();
  const int* rp = w.rp.as<int>();
  const int* dsp = w.dsp.as<int>();
  if (P > 0) {
    hipLaunchKernelGGL(k_adj_emit, dim3(nblk(P)), dim3(TB), 0, s, (int)P, w.cpi.as<int>(), I, gpos_out, gf_out,
                       w.emask.as<unsigned char>(), w.cpg.as<float>(), w.cpcap, a.fbar.as<float>(), a.gbar.as<float>(),
                       a.xb.as<float>());
    DRB_LAUNCH_PM(k_adj_gap, pmode, dim3(nblk(P)), dim3(TB), 0, s, (int)P, w.cpx.as<float>(), w.cpy.as<float>(),
                       w.cpi.as<int>(), I, rp + 1, dsp, w.emask.as<unsigned char>(), a.gbar.as<float>(),
                       a.xb.as<float>());
  }
  DRB_LAUNCH_PM(k_adj_comp, pmode, dim3(nblk(N)), dim3(TB), 0, s, (int)N, w.comp.as<Comp>(), w.swc.as<float>(),
                     I, rp + 1, w.cpx.as<float>(), w.cpy.as<float>(), dsp, (int)P, a.fbar.as<float>(),
                     a.sbar.as<float>(), a.vbs.as<float>(), a.qbs.as<float2>());
  if (P > 0)
    DRB_LAUNCH_PM(k_adj_cp, pmode, dim3(gblk(P)), dim3(TB), 0, s, (int)P, w.cpx.as<float>(), w.cpy.as<float>(),
                       w.cpi.as<int>(), I, rp, w.comp.as<Comp>(), a.sbar.as<float>(),
                       a.fbar.as<float>(), a.xb.as<float>(), w.kap.as<unsigned char>(), w.cpw.as<float>(),
                       w.cpn.as<int>(), w.cp0x.as<float>(), w.cp0y.as<float>(), a.ab.as<float2>(), a.z.as<float4>());
  DRB_LAUNCH_PM(k_adj_refine, pmode, dim3(nblk(N)), dim3(TB), 0, s, (int)N, w.comp.as<Comp>(), w.cv.as<float>(),
                     w.vals.as<uint32_t>(),
                     w.crun.as<int>(), I, rp + 1, w.cpx.as<float>(), w.cpy.as<float>(), a.z.as<float4>(),
                     a.ab.as<float2>(), a.vbs.as<float>(), a.qbs.as<float2>(), a.gv.as<float>(), a.gq.as<float2>());
  hipLaunchKernelGGL(k_adj_moment, dim3(nblk(Mi)), dim3(TB), 0, s, (int)Mi, w.sf.as<float>(), w.si.as<int>(), I,
                     a.gv.as<float>(), a.gq.as<float2>(), gpos_in, gf_in, a.tt.as<double>());
  hipLaunchKernelGGL(k_tau_partial, dim3(kTauChunks, ninst), dim3(TB), 0, s, a.tt.as<double>(), w.dmoff.as<long long>(),
                     a.tpart.as<double>());
  hipLaunchKernelGGL(k_tau_final, dim3((ninst + 63) / 64), dim3(64), 0, s, a.tpart.as<double>(), ninst, tacc);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

int drb_sim::tau_start() {
  DRB_CHECK(adj.tsum.ensure(sizeof(double) * ninst));
  hipLaunchKernelGGL(k_zero_f64, dim3((ninst + 63) / 64), dim3(64), 0, s, adj.tsum.as<double>(), ninst);
  DRB_HIP(hipGetLastError());
  return DRB_OK;
}

int drb_sim::tau_finish(double* gtau) {
  std::vector<double> h(ninst);
  DRB_HIP(hipMemcpyAsync(h.data(), adj.tsum.p, sizeof(double) * ninst, hipMemcpyDeviceToHost, s));
  DRB_HIP(hipStreamSynchronize(s));
  if (gtau)
    for (int b = 0; b < ninst; ++b) gtau[b] += h[b];
  return DRB_OK;
}

int drb_sim::step_vjp(int slot, const float* gpos_out, const float* gf_out, float* gpos_in, float* gf_in,
                      double* gtau) {
  if (slot < 0 || slot >= tape_n) return fail(DRB_ERR_INVALID_ARG, "drb_sim_step_vjp: bad slot");
  Work& w = tape[slot];
  if (!w.taped) return fail(DRB_ERR_STATE, "drb_sim_step_vjp: slot not recorded");
  if (w.P > 0 && !gf_out) return fail(DRB_ERR_INVALID_ARG, "drb_sim_step_vjp: null output cotangent");
  if (w.M > 0 && !gf_in) return fail(DRB_ERR_INVALID_ARG, "drb_sim_step_vjp: null input cotangent");
  DRB_CHECK(tau_start());
  DRB_CHECK(vjp_launch(slot, gpos_out, gf_out, gpos_in, gf_in, adj.tsum.as<double>()));
  return tau_finish(gtau);
}

int drb_sim::tape_vjp(const float* gpos_out, const float* gf_out, float* gpos_in, float* gf_in, double* gtau) {
  const int n = tape_pos;
  if (n < 1) return fail(DRB_ERR_STATE, "drb_sim_tape_vjp: no recorded step");
  const Work& wl = tape[n - 1];
  if (wl.P > 0 && !gf_out) return fail(DRB_ERR_INVALID_ARG, "drb_sim_tape_vjp: null output cotangent");
  if (tape[0].M > 0 && !gf_in) return fail(DRB_ERR_INVALID_ARG, "drb_sim_tape_vjp: null input cotangent");
  long long mmax = 1;
  for (int t = 0; t < n; ++t) mmax = std::max(mmax, tape[t].M);
  for (int i = 0; i < 2; ++i) {
    DRB_CHECK(adj.cpos[i].ensure((size_t)mmax * 8));
    DRB_CHECK(adj.cf[i].ensure((size_t)mmax * 28));
  }
  DRB_CHECK(tau_start());
  const float* go_p = gpos_out;
  const float* go_f = gf_out;
  for (int t = n - 1; t >= 0; --t) {
    float* gi_p = t == 0 ? gpos_in : adj.cpos[t & 1].as<float>();
    float* gi_f = t == 0 ? gf_in : adj.cf[t & 1].as<float>();
    if (t == 0 && !gi_p) gi_p = adj.cpos[0].as<float>();  // position cotangent not wanted: scratch
    DRB_CHECK(vjp_launch(t, go_p, go_f, gi_p, gi_f, adj.tsum.as<double>()));
    go_p = gi_p;
    go_f = gi_f;
  }
  return tau_finish(gtau);
}

extern "C" int drb_sim_step_vjp(drb_sim* sim, int32_t slot, const float* gpos_out, const float* gf_out,
                                float* gpos_in, float* gf_in, double* gtau) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim_step_vjp: null handle");
  return sim->step_vjp(slot, gpos_out, gf_out, gpos_in, gf_in, gtau);
}

extern "C" int drb_sim_tape_vjp(drb_sim* sim, const float* gpos_out, const float* gf_out, float* gpos_in,
                                float* gf_in, double* gtau) {
  if (!sim) return fail(DRB_ERR_INVALID_ARG, "drb_sim_tape_vjp: null handle");
  return sim->tape_vjp(gpos_out, gf_out, gpos_in, gf_in, gtau);
}
