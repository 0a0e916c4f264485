// Host-side state of the step engine, shared by the forward step (step.hip) and its adjoint
// (adjoint.hip).
#pragma once

#include "physics.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

namespace drb {

constexpr int TB = 256;
// stages of a step timed by drb_sim_stage_ms: collide, keys, sort, thin, refine, bin, normaliser,
// gather, gap-fill + emit
constexpr int kStages = 9;
// lanes cooperating on one query point in the neighbourhood kernels (see k_refine)
#ifndef DRB_GROUP
#define DRB_GROUP 4
#endif
constexpr int kGroup = DRB_GROUP;
#ifndef DRB_XCD
#define DRB_XCD 0
#endif
// Block order of the 2D refine and gather (xcd_block in step.hip): chunks of this many consecutive
// blocks, dealt to the 8 XCDs in turn (0: the hardware's round-robin of single blocks).
#ifndef DRB_XCD_CHUNK
#define DRB_XCD_CHUNK 128
#endif
static inline unsigned nblk(long long n) { return (unsigned)std::max<long long>(1, (n + TB - 1) / TB); }
// blocks for one kGroup-lane (or g-lane) group per item
static inline unsigned gblk(long long n, int g = kGroup) { return nblk(n * g); }
// Grid of the gap-fill/emit kernel, which strides over the points: at most 8192 blocks.  (2048
// kept its bounding-box atomics few when the emit widened the boxes; the boxes now come from the
// bins, and 8192 blocks hide more of the gap walks' latency: gap-fill 37.7 -> 32 us at 1M points.)
#ifndef DRB_EMIT_BLOCKS
#define DRB_EMIT_BLOCKS 8192
#endif
static inline unsigned emit_blocks(long long n) { return std::min<unsigned>(nblk(n), (unsigned)DRB_EMIT_BLOCKS); }
// grid rounded up to a multiple of the 8 XCDs (see xcd_block)
static inline unsigned xblk(unsigned nb) {
  // A multiple of the 8 XCDs always (the 3D kernels' XCD-major order needs it).  The 2D chunked
  // order (xcd_block): grids of at least four rounds of 8 chunks are padded to whole rounds (at
  // most 25% idle blocks); smaller grids keep the hardware order.
  const unsigned m = 8u * (DRB_XCD_CHUNK > 0 ? DRB_XCD_CHUNK : 1);
  if (nb < 4u * m) return (nb + 7u) & ~7u;
  return (nb + m - 1u) / m * m;
}

// Launch a kernel template instantiated on the periodic mode of the instances: 0 all open, 1 all
// periodic, 2 mixed (decided per point at run time).  With 0 or 1 the minimum-image selects of
// the neighbourhood loops are resolved at compile time.
#define DRB_LAUNCH_PM(kern, pm, grid, block, shm, stream, ...)                  \
  do {                                                                          \
    if ((pm) == 0) hipLaunchKernelGGL((kern<0>), grid, block, shm, stream, __VA_ARGS__); \
    else if ((pm) == 1) hipLaunchKernelGGL((kern<1>), grid, block, shm, stream, __VA_ARGS__); \
    else hipLaunchKernelGGL((kern<2>), grid, block, shm, stream, __VA_ARGS__); \
  } while (0)

// Unroll of the normaliser's wave-uniform candidate loop (scalar loads in flight).
#ifndef DRB_NORM_UNROLL
#define DRB_NORM_UNROLL 4
#endif

// Candidate records per lane loaded ahead of use in the point-group neighbourhood loops.
#ifndef DRB_INFLIGHT
#define DRB_INFLIGHT 4
#endif
constexpr int kInflight = DRB_INFLIGHT;

// Sums over the kGroup lanes of a group with a butterfly (__shfl_xor): both partners add the same
// two values, so every lane ends with identical bits.
template <int G = kGroup>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v = fadd(v, __shfl_xor(v, o, G));
  return v;
}
// group_sum for the replay's value type (float: group_sum itself; double: the same butterfly).
__device__ __forceinline__ float group_sum_t(float v) { return group_sum<kGroup>(v); }
__device__ __forceinline__ double group_sum_t(double v) {
#pragma unroll
  for (int o = kGroup / 2; o > 0; o >>= 1) v = v + __shfl_xor(v, o, kGroup);
  return v;
}
template <int G = kGroup>
__device__ __forceinline__ unsigned group_or(unsigned v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v |= __shfl_xor(v, o, G);
  return v;
}

// Launch a neighbourhood kernel template instantiated on the periodic mode and on the number of
// lanes per query point (2, 4 or 8: the tile parameter the autotuner chooses).
#define DRB_LAUNCH_PMG(kern, pm, g, grid, block, shm, stream, ...)                                     \
  do {                                                                                                  \
    if ((g) == 2) DRB_LAUNCH_PM2(kern, pm, 2, grid, block, shm, stream, __VA_ARGS__);                  \
    else if ((g) == 8) DRB_LAUNCH_PM2(kern, pm, 8, grid, block, shm, stream, __VA_ARGS__);             \
    else DRB_LAUNCH_PM2(kern, pm, 4, grid, block, shm, stream, __VA_ARGS__);                            \
  } while (0)
#define DRB_LAUNCH_PM2(kern, pm, G, grid, block, shm, stream, ...)                                      \
  do {                                                                                                  \
    if ((pm) == 0) hipLaunchKernelGGL((kern<0, G>), grid, block, shm, stream, __VA_ARGS__);              \
    else if ((pm) == 1) hipLaunchKernelGGL((kern<1, G>), grid, block, shm, stream, __VA_ARGS__);         \
    else hipLaunchKernelGGL((kern<2, G>), grid, block, shm, stream, __VA_ARGS__);                        \
  } while (0)
// The same for the point-group kernels (k_refine, k_gather) with their tile: g lanes per point
// (2, 4, 8) and k candidates in flight per lane (2, 4).
#define DRB_LAUNCH_PMGK(kern, pm, g, k, grid, block, shm, stream, ...)                                 \
  do {                                                                                                  \
    if ((k) == 2) {                                                                                     \
      if ((g) == 2) DRB_LAUNCH_PM3(kern, pm, 2, 2, grid, block, shm, stream, __VA_ARGS__);             \
      else if ((g) == 8) DRB_LAUNCH_PM3(kern, pm, 8, 2, grid, block, shm, stream, __VA_ARGS__);        \
      else DRB_LAUNCH_PM3(kern, pm, 4, 2, grid, block, shm, stream, __VA_ARGS__);                      \
    } else {                                                                                            \
      if ((g) == 2) DRB_LAUNCH_PM3(kern, pm, 2, 4, grid, block, shm, stream, __VA_ARGS__);             \
      else if ((g) == 8) DRB_LAUNCH_PM3(kern, pm, 8, 4, grid, block, shm, stream, __VA_ARGS__);        \
      else DRB_LAUNCH_PM3(kern, pm, 4, 4, grid, block, shm, stream, __VA_ARGS__);                      \
    }                                                                                                   \
  } while (0)
#define DRB_LAUNCH_PM3(kern, pm, G, K, grid, block, shm, stream, ...)                                   \
  do {                                                                                                  \
    if ((pm) == 0) hipLaunchKernelGGL((kern<0, G, K>), grid, block, shm, stream, __VA_ARGS__);           \
    else if ((pm) == 1) hipLaunchKernelGGL((kern<1, G, K>), grid, block, shm, stream, __VA_ARGS__);      \
    else hipLaunchKernelGGL((kern<2, G, K>), grid, block, shm, stream, __VA_ARGS__);                     \
  } while (0)

// Walk the control points that may lie within r of each lane's point (x, y) of instance b, one
// lane per point.  When the wave's active lanes share one instance and span a compact box, the
// wave walks the window of that box with wave-uniform bounds (scalar loads, scalar loop control)
// and every lane sees every candidate; otherwise each active lane walks its own window.  fn(j) is
// called per candidate and must apply the radius predicate itself (inactive lanes may call it
// with their dummy point).  rp2 is the control-point component of the cell prefix (rp + 1).
template <class F>
__device__ __forceinline__ void for_wave_window(const Inst* __restrict__ ins, const int* __restrict__ rp2,
                                                bool active, int b, float x, float y, float r,
                                                const int* __restrict__ disp, F&& fn) {
  int bmin = active ? b : INT_MAX, bmax = active ? b : INT_MIN;
  float x0 = active ? x : INFINITY, x1 = active ? x : -INFINITY;
  float y0 = active ? y : INFINITY, y1 = active ? y : -INFINITY;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    bmin = min(bmin, __shfl_xor(bmin, o, 64));
    bmax = max(bmax, __shfl_xor(bmax, o, 64));
    x0 = fminf(x0, __shfl_xor(x0, o, 64));
    x1 = fmaxf(x1, __shfl_xor(x1, o, 64));
    y0 = fminf(y0, __shfl_xor(y0, o, 64));
    y1 = fmaxf(y1, __shfl_xor(y1, o, 64));
  }
  if (bmax == INT_MIN) return;
  const Inst& p0 = ins[bmin];
  const float lim = 3.f * (r + p0.qcs);
  if (bmin == bmax && x1 - x0 <= lim && y1 - y0 <= lim) {
    const int ub = __builtin_amdgcn_readfirstlane(bmin);
    const Inst& pu = ins[ub];
    const float ux0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x0)));
    const float ux1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x1)));
    const float uy0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(y0)));
    const float uy1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(y1)));
    const int uext = disp ? __builtin_amdgcn_readfirstlane(disp[ub]) : 0;
    for_box_window(pu, rp2, ux0, ux1, uy0, uy1, r, uext, [&](int a, int e) {
      const int ua = __builtin_amdgcn_readfirstlane(a), ue = __builtin_amdgcn_readfirstlane(e);
#pragma unroll 4
      for (int j = ua; j < ue; ++j) fn(j);
    });
  } else if (active) {
    for_window(ins[b], rp2, x, y, r, disp ? disp[b] : 0, [&](int a, int e) {
      for (int j = a; j < e; ++j) fn(j);
    });
  }
}

// Graph mode, open instances: the arguments of the device layout of a step (graph_layout in
// step.hip), passed by value; ins == nullptr: no device layout (periodic grids only).
struct GraphLayout {
  Inst* ins;
  int ninst;
  BBox* box;
  int* disp;
  const int* counts;
  long long ncap;
  int* sv;
  unsigned* status;
};

// Buffers of one step.  A plain run reuses one set; a taped run (the adjoint's recompute) keeps
// one set per step so the backward pass can read every intermediate of that step.
struct Work {
  // forward intermediates (sorted components, cell lists, control points)
  Buf keys, vals, comp, cv, swc, cells, rp, cp0x, cp0y, cpq, cpi, cpx, cpy, cpf, emask, kap;
  // taped only: input state (positions, pre-collision f as SoA with pitch M, instance), the run
  // of each sorted component, run lengths and starts, refine weight sums, gap-fill sums
  Buf sx, sy, sf, si, crun, cpn, cps, cpw, cpg, dsp, dmoff;
  Buf d_inst;              // the layout (segment table) this step ran with
  std::vector<Inst> hin;   // host copy of it
  long long N = 0;         // components (7 M)
  long long M = 0;         // moments entering the step
  long long P = 0;         // control points = moments leaving the step
  long long ncells = 0;
  long long cpcap = 0;     // capacity (pitch) of the control-point buffers
  int sub_bits = 0;
  bool taped = false;
  std::vector<long long> counts_in, counts_out;  // per-instance moments in / out
  std::vector<long long> hmoff;                   // host copy of dmoff (kept alive for the async copy)
  void release() {
    for (Buf* b : {&keys, &vals, &comp, &cv, &swc, &cells, &rp, &cp0x, &cp0y, &cpq, &cpi, &cpx, &cpy, &cpf,
                   &emask, &kap, &sx, &sy, &sf, &si, &crun, &cpn, &cps, &cpw, &cpg, &dsp, &dmoff, &d_inst})
      b->release();
    cpcap = 0;
  }
};

// Scratch of the backward pass (sized by the largest step seen).
struct AdjWork {
  Buf fbar, gbar, xb, ab, z, sbar, vbs, qbs, gv, gq, tt, tpart, tsum;
  Buf cpos[2], cf[2];  // cotangents between the steps of a segment
  Buf rg, rq, rv, ra, rx, rcq, rf;  // replay: collided f, component positions / values, centroids, ...
  void release() {
    for (Buf* b : {&fbar, &gbar, &xb, &ab, &z, &sbar, &vbs, &qbs, &gv, &gq, &tt, &tpart, &tsum, &cpos[0], &cpos[1],
                   &cf[0], &cf[1], &rg, &rq, &rv, &ra, &rx, &rcq, &rf})
      b->release();
  }
};

}  // namespace drb

// Workspace of the stage entry points.  Each call builds a dense grid over its own points.
struct drb_stage {
  drb::Buf d_inst, box, status, keys, vals, ckeys, cvals, comp, cv, swc, cells, rp, cpsx, cpsy, cpx, cpy, cpf, emask;
  drb::RadixSort sorter;
  drb::Scan2 scan;
  void* hpin = nullptr;
  ~drb_stage() {
    for (drb::Buf* b : {&d_inst, &box, &status, &keys, &vals, &ckeys, &cvals, &comp, &cv, &swc, &cells, &rp, &cpsx,
                        &cpsy, &cpx, &cpy, &cpf, &emask})
      b->release();
    sorter.release();
    scan.release();
    if (hpin) (void)hipHostFree(hpin);
  }
};

struct drb_sim {
  int ninst = 0;
  int pmode = 2;  // periodic mode of the instances (DRB_LAUNCH_PM)
  hipStream_t s = nullptr;
  std::vector<drb::Inst> hin;  // parameters + current layout
  drb::Buf d_inst;             // parameters for set_state's checks
  // state (double-buffered SoA)
  drb::Buf mx[2], my[2], mf[2], mi[2];
  long long mcap[2] = {0, 0};
  int cur = 0;
  long long M = 0;
  std::vector<long long> counts;
  bool have_state = false;
  // per-step buffers: w0 for plain steps, tape[] for recorded steps
  drb::Work w0;
  std::vector<drb::Work> tape;
  int tape_n = 0, tape_pos = 0;
  drb::Work* last = nullptr;  // buffers of the last step (diagnostics)
  drb::AdjWork adj;
  drb::Buf box, disp, status, dcounts, moff, diag, diagpart;
  drb::Buf hio;  // the device block holding box, dcounts and status (views into it)
  // refined control points binned by their own cell (the normaliser's and the gap-fill's
  // candidates): bin of each point, per-bin counts and prefix, positions in bin order, bin slot of
  // each point, its arrival slot in the bin's count, empty-direction masks in bin order
  drb::Buf bk, bcells, brp, bx, by, bslot, bsl, bem;
  drb::Buf bpb, bpq;  // per-block partials of the next step's boxes (k_binsort, k_box_merge)
  drb::Buf pxy;       // single simulations: the moment positions as pairs (k_head -> k_mark)
  drb::Buf pxyi;      // several instances, graph steps: {x, y, instance, -} per moment (k_head -> k_mark)
  // one-level grids: the tile-major normaliser (k_norm_tiles): each instance's first tile, each
  // tile's components and first position, each wave's first tile; graph mode's tile count
  drb::Buf ntoff, tcnt, tpos, tseg, wstart;
  long long gtiles = 0;
  int tile_buffers(long long ntiles, long long ncomp);
  int norm_tiles(const drb::Work& w, const drb::Inst* I, long long ntiles, long long ncomp, hipStream_t q);
  // circular obstacles (any instance): bounce parts per moment and their offsets, the parts'
  // records, the kept control points' flags and offsets
  bool obst = false;
  drb::Buf bcnt, boff, ex, kc, koff;
  // soft_outer spawn (kappa_soft > kappa_hard, single simulations): flags and positions per refined
  // point, their prefix, the point total P + S on the device; and the point of each bin slot
  bool spawn = false;
  drb::Buf spflag, spoff, spx, spy, sptot, bidx;
  drb::RadixSort sorter;
  drb::Scan2 scan;
  // pinned host mirror of (box, counts, status)
  void* hpin = nullptr;
  size_t hpin_bytes = 0;
  unsigned timing = 0;  // stages timed by HIP events (bit i: stage i), 0: off
  float stage_ms[drb::kStages] = {};
  // two event sets, alternating by step: a step's later stages are read after the next step has
  // been queued
  hipEvent_t ev[2][drb::kStages + 1] = {};
  unsigned tpar = 0;        // event set of the step being queued
  long long updates = 0;    // point-updates queued so far: the moments entering each step, summed
  float stage_sum[drb::kStages] = {};
  int stage_n = 0;          // steps summed in stage_sum by the current drb_sim_step call
  bool back_unread = false; // the previous step's stages from the normaliser on are still to be read
  // Read back of a step: boxes, counts and status copied to the pinned mirror as soon as the next
  // step's layout is known (after the refined points are binned; after the emit with obstacles),
  // with ev_mid recorded behind the copy.  step_finish waits on it, so the host lays out and queues
  // the next step while the GPU runs the rest of this one.
  hipEvent_t ev_mid = nullptr;
  drb::Work* pend_w = nullptr;
  int pend_nxt = 0;
  bool pend_rec = false;

  // graph mode: fixed capacities, counts on the device, one captured step per buffer parity,
  // replayed with one host synchronisation per drb_sim_step call (open instances: the layout of
  // each step made on the device within a cell capacity, graph_ready)
  bool graph_on = false;
  int g_refine = drb::kGroup, g_gather = drb::kGroup;  // lanes per point (autotuned)
  int d_refine = drb::kInflight, d_gather = drb::kInflight;  // candidates in flight per lane (autotuned)
  hipStream_t gs = nullptr;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<const void*> gsig;
  long long gpcap = 0;
  long long gncells = 0;    // cells of the captured step (open instances: the capacity)
  long long gncap = 0;      // open instances: the cell capacity, grown as the clouds grow
  long long gpcap_open = 0; // open instances: the control-point capacity, grown as needed
  int ghorizon = 16;        // steps of the largest growth the open capacity covers
  long long gcaptures = 0, gresumes = 0;  // graph captures, replays resumed after an empty step
  drb::Buf gsv;             // graph_layout's record of an empty step
  drb::Buf ginst1;          // open instances: the layout table of the steps from state buffer 1
  drb::Buf gcnt;  // {moments entering the step, components}, then the replayed point-updates (int64)
  unsigned long long* graph_upd() const { return reinterpret_cast<unsigned long long*>(gcnt.as<char>() + 8); }

  ~drb_sim();
  int graph_ready(int horizon);
  drb::GraphLayout graph_layout_args(int parity);
  drb::Inst* graph_table(int parity);
  int graph_body(int from);
  int step_graph(int nsteps);
  int autotune(int* out);
  void graph_drop();
  int* hm() const { return (int*)(hstatus() + 1); }
  int* hE() const { return (int*)(hstatus() + 2); }  // bounce parts of the step
  size_t pin_need() const { return (size_t)ninst * (sizeof(drb::BBox) + sizeof(int)) + 16; }
  drb::BBox* hbox() const { return (drb::BBox*)hpin; }
  int* hcounts() const { return (int*)((char*)hpin + (size_t)ninst * sizeof(drb::BBox)); }
  unsigned* hstatus() const { return (unsigned*)((char*)hpin + (size_t)ninst * (sizeof(drb::BBox) + sizeof(int))); }
  int ensure_state(int which, long long cap);
  int ensure_bins(long long ncells, long long pcap);
  // I: the segment table of the step (default w.d_inst)
  int bin_control_points(const drb::Work& w, long long ncells, long long pcap, const int* Ptot, hipStream_t q,
                         bool graph, bool keyed, const drb::Inst* I = nullptr);
  int read_back();
  // defer: a single simulation's table is left for the step's first kernel to upload (k_head)
  int layout(drb::Work& w, bool reset_box, bool defer = false);
  // event i opens stage i and closes stage i - 1: recorded only where a timed stage needs it (an
  // event between two kernels leaves the GPU idle for a few microseconds)
  void mark(int i) {
    if (timing & ((1u << i) | (1u << i >> 1))) (void)hipEventRecord(ev[tpar][i], s);
  }
  // add the stages [i0, i1) of event set par, recorded and complete, to stage_sum
  int stage_read(unsigned par, int i0, int i1);
  int step_once();    // queue one step (returns before it completes)
  int step_finish();  // wait for the queued step's read back and take its counts
  int step_sync();    // one step, completed
  int init(const drb_params* prm, int n, hipStream_t stream);
  int set_state(const float* pos, const float* f, long long m, const int64_t* cnt, bool device);
  int rebox();
  int get_state(float* pos, float* f, bool device);
  int diagnostics(double* out);
  // adjoint of recorded steps (adjoint.hip)
  int vjp_launch(int slot, const float* gpos_out, const float* gf_out, float* gpos_in, float* gf_in, double* tacc);
  int tau_start();
  int tau_finish(double* gtau);
  int step_vjp(int slot, const float* gpos_out, const float* gf_out, float* gpos_in, float* gf_in, double* gtau);
  int tape_vjp(const float* gpos_out, const float* gf_out, float* gpos_in, float* gf_in, double* gtau);
  int replay(int slot, float tau, const float* pos_in, const float* f_in, float* pos_out, float* f_out);
  int replay64(int slot, double tau, const double* pos_in, const double* f_in, double* pos_out, double* f_out);
  template <class T>
  int replay_t(int slot, T tau, const T* pos_in, const T* f_in, T* pos_out, T* f_out);
};
