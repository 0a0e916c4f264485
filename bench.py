#!/usr/bin/env python3
"""Throughput of the DiReBM step on one MI355X, in million point-updates per second (MPUPS).

Workload: a 2D D2Q7 Gaussian density pulse (amplitude 1e-3, sigma 8h) on rho = 1 at rest,
tau = 0.6, about 1.0M points in an unbounded domain, cell thinning at two levels (h inside radius
0.3L, 2h outside), dx = 4h.  A step is one pass of collide -> disperse -> thin -> refine ->
resample over the whole point cloud; MPUPS = sum over timed steps of the moments entering the
step / wall time.  Inputs are resident in HBM before the timed region.  Also reported:

* roofline of the dominant kernel (HIP-event time on the engine's stream, algorithmic bytes);
* radix-sort keys/s on the step's component keys and on uniform-random 20-bit keys, per-stage
  times;
* the hex D2Q7 LBM on a 1024 x 1024 lattice with the same density field and tau, and on a
  4096 x 4096 lattice, whose two lattices exceed the 256 MB Infinity Cache;
* a parity gate (teacher-forced step vs the FP64 oracle on a small case of the same shape, and
  the mass budget of the full-size run);
* the CPU baseline: the restatement of the reference's CPU solver (direbm/reference) timed on a
  bounded sample of the same workload, with the vectorised FP64 oracle beside it.

Prints one JSON line.  `--gpus N` with N > 1 exits: the engine runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MPUPS (million point-updates/s) at fixed L2 err; batched gain; adjoint:forward"
HBM_PEAK = 8.0e12  # MI355X HBM3E bytes/s

# Algorithmic bytes per unit of each kernel: each array the kernel needs read once and written once
# (the component records, values and normalisers of a point's window counted once per point).
BYTES = {
    "collide": (56.0, "moment"),   # R f 28, W f 28
    "keys": (40.0, "moment"),      # R pos 8 + inst 4, W 7 keys 28
    "thin": (358.0, "moment"),     # mark 208 + scan/centroid 150
    "refine": (137.0, "cp"),       # R own 16 + 7 comp records 112, W pos 8 + kappa 1
    "norm": (147.0, "moment"),     # per moment: 7 x (R comp 16 + W sw 4) + CP pos share 7
    "gather": (153.0, "cp"),       # R own 12 + 7 comp records {x, y, val/sw, dir} 112, W f 28 + mask 1
    "gapfill": (90.0, "cp"),       # R f 28 + mask 1 + own 12 + nbr share 9, W pos 8 + f 28 + inst 4
}


CLOCK_HZ = 2.4e9  # MI355X peak engine clock
N_CU = 256
# VALU issue peak: one wave64 VALU instruction per CU per cycle (4 SIMD16 units, 4 cycles each)
VALU_PEAK = N_CU * CLOCK_HZ
# vector L1 (TCP): one 64-byte access per CU per cycle
L1_ACCESS_BYTES = 64
L1_PEAK = N_CU * CLOCK_HZ * L1_ACCESS_BYTES


def committed_counters(kernel, lanes=None, depth=None):
    """Per-dispatch counters of the step's `kernel` on this workload from the committed rocprofv3
    passes (profiles/traffic.json): HBM bytes (FETCH_SIZE / WRITE_SIZE scaled by the calibration
    kernel) and SQ_INSTS_VALU, for the instantiation of the open-domain step with `lanes` lanes per
    point (refine, gather), or {} when the file lacks it."""
    try:
        data = json.loads((ROOT / "profiles" / "traffic.json").read_text())
    except (OSError, ValueError):
        return {}
    inst = data.get("instantiations", {})
    names = [f"k_{kernel}<0, {lanes}, {depth}>"] if lanes else []
    names += [f"k_{kernel}<0>", f"k_{kernel}"]
    if kernel == "gapfill":
        names.insert(0, "k_gapfill_emit<0>")
    for n in names:
        if n in inst:
            return {"instantiation": n, **inst[n]}
    return {}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def parity_gate(drb, workloads):
    """Teacher-forced steps of a small case of the benchmark's shape vs the FP64 oracle."""
    import numpy as np

    from oracle import v2

    case = workloads.gaussian_pulse_two_level(n=96, amp=0.05, sigma_cells=4.0)
    kw = case.kw
    prm = v2.Params(tau=kw["tau"], dx=kw["dx"], alpha=kw["alpha"], rho_rest=kw["rho_rest"],
                    two_level=kw["two_level"])
    sim = drb.GpuSimulator(case.pos, case.f, **kw)
    worst = 0.0
    same = True
    for _ in range(2):
        p0, f0 = sim.moments()
        ref = v2.step(p0, f0, prm)
        sim.step()
        p1, f1 = sim.moments()
        same &= len(p1) == len(ref.pos)
        if not same:
            break
        e = float(np.linalg.norm(f1 - ref.f) / np.linalg.norm(ref.f))
        worst = max(worst, e)
    sim.close()
    return {"case": f"two-level Gaussian pulse, {len(case.pos)} points, 2 teacher-forced steps",
            "rel_l2_f": worst, "same_control_points": bool(same), "tol": 1e-5,
            "passed": bool(same and worst <= 1e-5)}


def cpu_baseline(workloads, seconds=8.0):
    """The reference solver's CPU path, timed on a bounded sample of the benchmark workload.

    Primary figure: oracle/v1.py, the restatement of direbm/reference (per-object Python loops over
    a dict-of-cells grid, FP64, one core), on the same workload shape at 64 x 64 seeding cells for
    two steps (about 10 s on the GPU hosts' cores).  Also reported: oracle/v2.py (the vectorised FP64 restatement of the GPU algorithm,
    numpy + scipy cKDTree, single thread) at 192 x 192 seeding cells.
    """
    import numpy as np

    from oracle import v1, v2

    model, ncpu = cpu_info()
    case = workloads.gaussian_pulse_two_level(n=64)
    kw = case.kw
    n1 = len(case.pos)
    t0 = time.perf_counter()
    sim = v1.Simulator(case.pos.astype(np.float64), case.f.astype(np.float64), kw["tau"], dx=kw["dx"],
                       alpha=kw["alpha"], rho_rest=kw["rho_rest"])
    m1 = 0
    for _ in range(2):
        m1 += len(sim.moments_)  # moments entering the step
        sim.step()
    wall1 = time.perf_counter() - t0
    v1_mpups = m1 / wall1 / 1e6

    case = workloads.gaussian_pulse_two_level(n=192)
    kw = case.kw
    prm = v2.Params(tau=kw["tau"], dx=kw["dx"], alpha=kw["alpha"], rho_rest=kw["rho_rest"],
                    two_level=kw["two_level"])
    pos, f = case.pos, case.f
    total, steps = 0, 0
    t0 = time.perf_counter()
    while steps < 1 or (time.perf_counter() - t0 < seconds and steps < 5):
        total += len(pos)
        r = v2.step(pos, f, prm)
        pos, f = r.pos.astype(np.float32), r.f.astype(np.float32)
        steps += 1
    wall2 = time.perf_counter() - t0
    return {"value": v1_mpups, "unit": "MPUPS", "cores": 1, "kind": "port",
            "sample": f"oracle/v1.py (restatement of direbm/reference) on the benchmark workload at "
                      f"64x64 seeding cells ({n1} points), 2 steps, {wall1:.1f} s",
            "v2_port": {"value": total / wall2 / 1e6, "unit": "MPUPS", "cores": 1,
                        "sample": f"oracle/v2.py at 192x192 seeding cells ({len(case.pos)} points), "
                                  f"{steps} steps, {wall2:.1f} s"},
            "cpu_model": model, "host_cpus": ncpu}


def bench_lbm(drb, workloads, torch, ni=1024, nj=1024, steps=500, warmup=20):
    dens = workloads.lbm_gaussian_density(ni, nj, amp=1e-3, sigma=8.0)
    lbm = drb.GpuHexLBM(ni, nj, tau=0.6)
    lbm.set_density(dens)
    lbm.step(warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    lbm.step(steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    nodes = ni * nj
    achieved = 56.0 * nodes / (ms * 1e-3)
    return {"lattice": f"{ni}x{nj}", "steps": steps, "ms_per_step": ms, "glups": nodes / (ms * 1e-3) / 1e9,
            "bytes_per_node_update": 56, "achieved_GBps": achieved / 1e9, "roofline_frac": achieved / HBM_PEAK}


def bench_sort(torch, n, bits=20, reps=10):
    """Radix-sort rate on n uniform-random `bits`-bit keys with their indices (the worst case for
    the sort: no order to exploit), through drb_radix_sort_pairs on torch's current stream."""
    import ctypes as C

    from direbm_amd import _lib
    from direbm_amd._device import stream_ptr

    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    src = torch.randint(0, 2**bits, (n,), dtype=torch.int32, device=dev, generator=g)
    keys = torch.empty(2 * n, dtype=torch.int32, device=dev)
    vals = torch.empty(2 * n, dtype=torch.int32, device=dev)
    iota = torch.arange(n, dtype=torch.int32, device=dev)
    h = C.c_void_p()
    _lib.call("drb_stage_create", C.byref(h))
    ms = []
    try:
        for rep in range(reps + 1):
            keys[:n].copy_(src)
            vals[:n].copy_(iota)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            _lib.call("drb_radix_sort_pairs", h, keys.data_ptr(), vals.data_ptr(), n, bits, stream_ptr(dev))
            e1.record()
            torch.cuda.synchronize()
            if rep:  # the first call sizes the workspace
                ms.append(e0.elapsed_time(e1))
        ok = bool(torch.all(keys[1:n] >= keys[: n - 1]).item())
    finally:
        _lib.call("drb_stage_destroy", h)
    t = sorted(ms)[len(ms) // 2]
    return {"keys": n, "bits": bits, "passes": (bits + 7) // 8, "ms": t, "gkeys_per_s": n / (t * 1e-3) / 1e9,
            "sorted": ok}


def bench_batch(drb, workloads, torch, ninst=256, steps=20, warmup=1):
    """Batched gain: the batch's MPUPS over the MPUPS of the same instances run with per-instance
    launches, two ways: stepped in lockstep (one GpuSimulator step after another, each waiting for
    its step: `gain`) and run one after another for all the steps, each instance's steps queued in
    one call (`gain_vs_sequential_runs`, the stronger baseline)."""
    cases = workloads.batch_cases(ninst=ninst)
    batch = drb.BatchSimulator([(c.pos, c.f, c.kw) for c in cases])
    lanes = batch.autotune()  # one step, counted as warm-up
    batch.set_graph(True)  # every instance is periodic: captured HIP graphs
    batch.step(warmup)
    torch.cuda.synchronize()
    n_b = int(batch.counts().sum())
    u0 = batch.updates
    t0 = time.perf_counter()
    batch.step(steps)
    torch.cuda.synchronize()
    wall_b = time.perf_counter() - t0
    total_b = batch.updates - u0
    batch.close()
    sims = [drb.GpuSimulator(c.pos, c.f, **c.kw) for c in cases]
    for sm in sims:
        sm.step(warmup)
    torch.cuda.synchronize()
    u0 = sum(sm.updates for sm in sims)
    t0 = time.perf_counter()
    for _ in range(steps):
        for sm in sims:
            sm.step()
    torch.cuda.synchronize()
    wall_s = time.perf_counter() - t0
    total_s = sum(sm.updates for sm in sims) - u0
    u0 = total_s + u0
    t0 = time.perf_counter()
    for sm in sims:
        sm.step(steps)
    torch.cuda.synchronize()
    wall_q = time.perf_counter() - t0
    total_q = sum(sm.updates for sm in sims) - u0
    for sm in sims:
        sm.close()
    b = total_b / wall_b / 1e6
    single = total_s / wall_s / 1e6
    seq = total_q / wall_q / 1e6
    return {"instances": ninst, "points": n_b, "steps": steps,
            "workload": "periodic Taylor-Green / double shear layer 50/50, counts log-uniform in [4k, 256k], "
                        "tau ~ U(0.55, 1.5), u0 ~ U(0.01, 0.08)",
            "batched_form": {"hip_graph": True, "autotuned_tiles": lanes, "timed": f"one {steps}-step call"},
            "batched_mpups": b, "per_instance_mpups": single, "gain": b / single,
            "sequential_runs_mpups": seq, "gain_vs_sequential_runs": b / seq,
            "batched_ms_per_step": wall_b / steps * 1e3, "per_instance_ms_per_step": wall_s / steps * 1e3,
            "sequential_runs_ms_per_step": wall_q / steps * 1e3}


def bench_3d(drb, workloads, torch, n=200, steps=5, warmup=1, sphere=0.1):
    """The 3D D3Q13 step at n^3 seeding cells: periodic double shear layer (u0 0.05, delta 0.05 L,
    seeded 1e-4 v perturbation, tau 0.7) around a spherical obstacle of radius sphere * L."""
    case = workloads.shear_layer_3d(n=n, sphere=sphere)
    sim = drb.GpuSimulator3D(case.pos, case.f, **case.kw)
    m0 = sim.count()
    mass_init = sim.diagnostics()["mass"]
    sim.step(warmup)
    torch.cuda.synchronize()
    mass0 = sim.diagnostics()["mass"]
    total = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        total += sim.count()
        sim.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    mass1 = sim.diagnostics()["mass"]
    m1 = sim.count()
    sim.close()
    return {"workload": f"3D D3Q13 periodic double shear layer, {n}^3 seeding cells (h = L/{n}), "
                        f"sphere r = {sphere:g} L at the centre, tau 0.7",
            "points_initial": m0, "points_final": m1, "steps": steps, "warmup": warmup,
            "ms_per_step": wall / steps * 1e3, "mpups": total / wall / 1e6,
            "mass_relative_change_over_timed_steps": (mass1 - mass0) / mass0,
            "mass_relative_change_from_initial_state": (mass1 - mass_init) / mass_init,
            "gapfill_eps": case.kw["eps"]}


def bench_adjoint(workloads, torch, n=512, steps=200, every=20, fd_dirs=64):
    """Differentiable rollout: forward time against recompute + backward time, and central
    differences (eps 1e-3) of the loss along `fd_dirs` random directions in the initial velocities
    (half white noise, half smooth Fourier fields, scaled to the field's largest speed) and in tau,
    evaluated by the GPU replay of the recorded rollout in FP64 (discrete choices held fixed, as the
    adjoint holds them).
    Loss: binned density and velocity against a target run from a different seeded velocity field."""
    import numpy as np

    from direbm_amd import adjoint as adj
    from direbm_amd.fields import bin_edges, bin_moments, fields_t

    dev = "cuda:0"
    pos, rho0, u0, kw = workloads.adjoint_case(n=n, seed=0)
    _, _, u_tgt, _ = workloads.adjoint_case(n=n, seed=1)
    tau0 = kw.pop("tau")
    edges = bin_edges(0.0, float(np.float32(kw["box"])), kw["dx"])
    pos_t = torch.tensor(pos, device=dev)
    rho_t = torch.tensor(rho0, device=dev)
    with torch.no_grad():
        pT, fT = adj.rollout(pos_t, adj.equilibrium_t(rho_t, torch.tensor(u_tgt, device=dev)), tau0, steps, **kw)
        rho_g, u_g = fields_t(*bin_moments(pT, fT, edges), kw["dx"])

    def run(u, tau, every_, timings=None):
        f0 = adj.equilibrium_t(rho_t, u)
        p, f = adj.rollout(pos_t, f0, tau, steps, checkpoint_every=every_, timings=timings, **kw)
        return adj.fields_loss(p, f, edges, kw["dx"], rho_g, u_g)

    # warm-up (allocations, tape buffers)
    u = torch.tensor(u0, device=dev, requires_grad=True)
    tau = torch.tensor(tau0, dtype=torch.float64, device=dev, requires_grad=True)
    run(u, tau, every).backward()
    torch.cuda.synchronize()
    u = torch.tensor(u0, device=dev, requires_grad=True)
    tau = torch.tensor(tau0, dtype=torch.float64, device=dev, requires_grad=True)
    timings = {}
    t0 = time.perf_counter()
    loss = run(u, tau, every, timings)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    gu, gtau = u.grad.double(), tau.grad.item()
    out = {"workload": f"periodic [0, 2pi)^2, {n}x{n} points, 8 random Fourier velocity modes (amp 0.03), "
                       f"tau {tau0}, {steps} steps, checkpoints every {every}",
           "points": n * n, "steps": steps, "checkpoint_every": every,
           "forward_s": t_fwd, "backward_s": t_bwd,
           "recompute_s": timings.get("recompute_s"), "adjoint_kernels_s": timings.get("backward_s"),
           "adjoint_over_forward": t_bwd / t_fwd, "forward_mpups": n * n * steps / t_fwd / 1e6,
           "loss": loss.item(), "dloss_dtau": gtau, "grad_u0_norm": float(gu.norm())}
    if fd_dirs:
        t0 = time.perf_counter()
        with torch.no_grad():
            u0_t = torch.tensor(u0, device=dev)
            fr = adj.FrozenRollout(pos_t, adj.equilibrium_t(rho_t, u0_t), tau0, steps, **kw)
            # the differenced function: the recorded rollout replayed in FP64 (choices fixed), its
            # final fields binned by the recorded final positions in FP64
            bins = adj.fixed_bins(fr.final[0], edges)
            nb = len(edges) - 1
            pos64, rho64 = pos_t.double(), rho_t.double()

            def L(uu, tt):
                _, f_end = fr(pos64, adj.equilibrium64(rho64, uu), tt)
                return adj.fields_loss_fixed(bins, f_end, nb, kw["dx"], rho_g, u_g).item()

            rng = np.random.default_rng(2)
            p64 = pos.astype(np.float64)
            white, smooth = [], []
            eps = 1e-3
            for j in range(fd_dirs):
                if j % 2 == 0:
                    d, acc = rng.standard_normal(p64.shape), white
                else:
                    d, acc = workloads.fourier_velocity(n, 8, 1.0, seed=1000 + j)(p64), smooth
                # eps relative to the field: max |d| = max |u0| (0.03); with max |d| = 1 the
                # difference's own O(eps^2) term reaches 1e-3 in the most curved directions
                d = torch.tensor(d / np.abs(d).max() * np.abs(u0).max(), device=dev)
                fd = (L(u0_t.double() + eps * d, tau0) - L(u0_t.double() - eps * d, tau0)) / (2 * eps)
                an = float((gu * d).sum())
                acc.append(abs(fd - an) / abs(an))
            fd = (L(u0_t.double(), tau0 + eps) - L(u0_t.double(), tau0 - eps)) / (2 * eps)
            fr.close()
        allr = np.array(white + smooth)
        out["fd_check"] = {
            "method": "central differences of the recorded rollout replayed in FP64 (choices fixed), "
                      "fields binned by the recorded final positions",
            "directions": fd_dirs, "eps": eps, "eps_tau": eps,
            "direction_scale": "max |d| = max |u0|: u0 +- eps d is a relative perturbation of the field",
            "rel_err_median": float(np.median(allr)), "rel_err_max": float(allr.max()),
            "rel_err_max_white": float(max(white)) if white else None,
            "rel_err_max_smooth": float(max(smooth)) if smooth else None,
            "frac_within_1e-3": float((allr <= 1e-3).mean()),
            "rel_err_tau": abs(fd - gtau) / abs(gtau), "seconds": time.perf_counter() - t0}
    return out


def host_round_trip(sim, torch, units, args, m_end, ms_step):
    """The state's PCIe round trip after the timed steps (moments out, set_state back), then the
    sort on uniform-random keys as many as the step's component keys."""
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pos_h, f_h = sim.moments()
    t2 = time.perf_counter()
    sim.set_state(pos_h, f_h)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    d2h_ms, h2d_ms = (t2 - t1) * 1e3, (t3 - t2) * 1e3
    pcie = {"state_bytes": int(pos_h.nbytes + f_h.nbytes), "d2h_ms": d2h_ms, "h2d_ms": h2d_ms,
            "mpups_with_state_round_trip_per_step": m_end / ((ms_step + d2h_ms + h2d_ms) * 1e-3) / 1e6}
    log("pcie:", pcie)
    del pos_h, f_h
    sim.close()
    sort_random = bench_sort(torch, int(round(7 * units["moment"] / args.steps)))
    log("sort, random 20-bit keys:", sort_random)
    return pcie, sort_random


def bench_graph_open(drb, case, lanes, torch, steps):
    """The headline workload in graph mode: the same steps replayed from captured HIP graphs (each
    step's grid laid out on the device), one host synchronisation per call, the tuned tiles."""
    sim = drb.GpuSimulator(case.pos, case.f, **case.kw)
    if lanes:
        sim.set_tiles(**lanes)
    sim.set_graph(True)
    # warm-up: 3 steps, the first capture.  Not a call as long as the timed one: the open pulse's
    # exp(f) weights overflow near step 38 (the reference algorithm's own blow-up), so the timed
    # call (steps 4 to 3 + steps) includes the capture sized for its length and any resume.
    sim.step(3)
    torch.cuda.synchronize()
    u0 = sim.updates
    t0 = time.perf_counter()
    sim.step(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {"ms_per_step": wall / steps * 1e3, "mpups": (sim.updates - u0) / wall / 1e6, "steps": steps,
           "graph": sim.graph_info(),
           "timed": f"one {steps}-step call after a 3-step warm-up call; it includes its own capture and resumes"}
    sim.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=25)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1472, help="seeding cells per side (1472 -> ~1.0M points)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-lbm", action="store_true", help="skip the hex LBM line")
    ap.add_argument("--no-parity", action="store_true", help="skip the parity gate")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-gain line")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint:forward line")
    ap.add_argument("--fd-dirs", type=int, default=64, help="central-difference directions of the adjoint check")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D line")
    ap.add_argument("--n3d", type=int, default=200, help="3D seeding cells per side (200 -> ~8.0M points)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 256 MiB device-to-device copy (reference kernel for HBM counters)")
    ap.add_argument("--only", choices=["2d", "sort", "lbm", "batch", "3d", "adjoint"],
                    help="run one leg alone and print its object (per-leg profiles); '2d' prints the "
                         "headline line with every other leg skipped")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the engine targets one GPU; multi-GPU runs (--gpus > 1) are not supported.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import direbm_amd as drb
    from direbm_amd import workloads

    if not torch.cuda.is_available():
        print("bench.py: no HIP GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.calib_copy:
        # 256 MiB read + 256 MiB written by one elementwise kernel (mul): the reference dispatch
        # for scaling the FETCH_SIZE / WRITE_SIZE counters to HBM bytes
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        del a, b
    if args.only and args.only != "2d":
        leg = {
            "sort": lambda: {"uniform_random_20bit": bench_sort(torch, 7 * 1001296)},
            "lbm": lambda: {"lbm": bench_lbm(drb, workloads, torch),
                            "lbm_4096": bench_lbm(drb, workloads, torch, ni=4096, nj=4096, steps=200, warmup=10)},
            "batch": lambda: {"batched": bench_batch(drb, workloads, torch)},
            "3d": lambda: {"d3q13_3d": bench_3d(drb, workloads, torch, n=args.n3d)},
            "adjoint": lambda: {"adjoint": bench_adjoint(workloads, torch, fd_dirs=args.fd_dirs)},
        }[args.only]()
        print(json.dumps({"leg": args.only, **leg}, default=float), flush=True)
        return
    if args.only == "2d":
        args.no_parity = args.no_lbm = args.no_batch = args.no_3d = args.no_adjoint = args.no_cpu = True
    parity = None if args.no_parity else parity_gate(drb, workloads)
    if parity is not None:
        log("parity gate:", parity)

    case = workloads.gaussian_pulse_two_level(n=args.n)
    log(f"workload: {case.name}, {len(case.pos)} points")
    sim = drb.GpuSimulator(case.pos, case.f, **case.kw)
    m0 = sim.count()
    lanes = sim.autotune() if args.warmup > 0 else None  # one step, counted as warm-up
    # The stage split: HIP events on every stage of the other warm-up steps.  The timed steps
    # record events around the dominant kernel only (each event leaves the GPU idle a few us).
    sim.set_timing(True)
    split = {k: 0.0 for k in drb.GpuSimulator.STAGES}
    nsplit = split_moments = 0
    for _ in range(max(0, args.warmup - 1)):
        split_moments += sim.count()
        sim.step()
        for k, v in sim.stage_ms().items():
            split[k] += v
        nsplit += 1
    per = {k: v / max(nsplit, 1) for k, v in split.items()}
    cand = {k: per[k] for k in ("collide", "keys", "refine", "norm", "gather", "gapfill")}
    dom = max(cand, key=cand.get) if nsplit else "gather"  # dominant single-kernel stage
    sim.set_timing(True, stages=[dom])
    torch.cuda.synchronize()
    mass_prev = sim.diagnostics()["mass"]

    # The timed steps in one call: the engine queues each step behind the previous one (it reads
    # the next layout back mid-step), as a caller stepping a simulation does.
    u0, m_first = sim.updates, sim.count()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total = sim.updates - u0  # moments entering the timed steps
    m_end = sim.count()
    units = {"moment": total, "cp": total - m_first + m_end}  # control points = moments leaving
    ms_launch = sim.stage_ms()[dom]  # mean over the timed steps
    diag = sim.diagnostics()
    mpups = total / wall / 1e6
    ms_step = wall / args.steps * 1e3

    b_unit, unit = BYTES[dom]
    bytes_per_launch = b_unit * units[unit] / args.steps
    achieved = bytes_per_launch / (ms_launch * 1e-3)
    g = k = None
    if lanes and dom in ("refine", "gather"):
        g, k = lanes[f"{dom}_lanes"], lanes[f"{dom}_depth"]
    ctr = committed_counters(dom, g, k)
    valu = None
    if ctr.get("SQ_INSTS_VALU"):
        valu = {"insts_per_launch": ctr["SQ_INSTS_VALU"], "peak_insts_per_s": VALU_PEAK,
                "frac": ctr["SQ_INSTS_VALU"] / (VALU_PEAK * ms_launch * 1e-3),
                "source": f"SQ_INSTS_VALU of {ctr['instantiation']} (profiles/traffic.json) over this run's "
                          "launch time; peak: one wave64 VALU instruction per CU per cycle at 2.4 GHz"}
    l1 = None
    if ctr.get("TCP_TOTAL_CACHE_ACCESSES_sum"):
        l1_bytes = ctr["TCP_TOTAL_CACHE_ACCESSES_sum"] * L1_ACCESS_BYTES
        l1 = {"accesses_per_launch": ctr["TCP_TOTAL_CACHE_ACCESSES_sum"], "achieved": l1_bytes / (ms_launch * 1e-3) / 1e9,
              "peak": L1_PEAK / 1e9, "unit": "GB/s", "frac": l1_bytes / (L1_PEAK * ms_launch * 1e-3),
              "source": f"TCP_TOTAL_CACHE_ACCESSES_sum of {ctr['instantiation']} (profiles/traffic.json) x 64 B over "
                        "this run's launch time; peak: one 64-byte vector-L1 access per CU per cycle at 2.4 GHz"}
    sort_gkeys = (7 * split_moments / nsplit / (per["sort"] * 1e-3) / 1e9) if nsplit else None

    # Host boundary, outside the timed region: the state leaves over PCIe (moments) and comes back
    # (set_state), as for a caller that holds it in host memory between steps.
    if args.only == "2d":
        pcie = None
        sim.close()
        sort_random = None
    else:
        pcie, sort_random = host_round_trip(sim, torch, units, args, m_end, ms_step)
    graph_open = None if args.only == "2d" else bench_graph_open(drb, case, lanes, torch, args.steps)
    if graph_open is not None:
        log("graph mode, open pulse:", graph_open)
    lbm = None if args.no_lbm else bench_lbm(drb, workloads, torch)
    # 4096^2: the two lattices (2 x 470 MB) exceed the 256 MB Infinity Cache, so the figure is HBM's
    lbm_large = None if args.no_lbm else bench_lbm(drb, workloads, torch, ni=4096, nj=4096, steps=200, warmup=10)
    if lbm_large is not None:
        log("lbm 4096^2:", lbm_large)
    batched = None if args.no_batch else bench_batch(drb, workloads, torch)
    d3 = None if args.no_3d else bench_3d(drb, workloads, torch, n=args.n3d)
    if d3 is not None:
        log("3d:", d3)
    adjoint = None if args.no_adjoint else bench_adjoint(workloads, torch, fd_dirs=args.fd_dirs)
    if adjoint is not None:
        log("adjoint:", adjoint)
    cpu = None if args.no_cpu else cpu_baseline(workloads)

    out = {
        "metric": METRIC,
        "value": mpups,
        "unit": "MPUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "2D D2Q7 DiReBM, Gaussian density pulse (amp 1e-3, sigma 8h) on rho=1 at rest, "
                        "unbounded domain, two-level thinning (h inside 0.3L, 2h outside)",
            "points_initial": m0,
            "points_final": m_end,
            "seeding_cells": args.n,
            "h": 0.25,
            "dx": 1.0,
            "tau": 0.6,
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": ctr.get("hbm_bytes"),
            "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)",
            "valu": valu,
            "l1": l1,
            "bytes_per_launch": bytes_per_launch,
            "ms_per_launch": ms_launch,
            "timing": "HIP events around this kernel on the engine's stream, every timed step",
        },
        "stages_ms_per_step": per,
        "stages_source": f"HIP events on every stage over {nsplit} untimed warm-up steps",
        "autotuned_tiles": lanes,
        "sort_gkeys_per_s": sort_gkeys,
        "sort": {"component_keys": {"keys_per_step": 7 * split_moments / max(nsplit, 1), "ms": per["sort"],
                                    "gkeys_per_s": sort_gkeys},
                 "uniform_random_20bit": sort_random},
        "mass": {"relative_change_over_timed_steps": (diag["mass"] - mass_prev) / mass_prev,
                 "last_step_dropped": diag["dropped"], "last_step_injected": diag["injected"]},
        "parity": parity,
        "lbm": lbm,
        "lbm_4096": lbm_large,
        "batched": batched,
        "d3q13_3d": d3,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
        "gpu_over_cpu": (mpups / cpu["value"]) if cpu else None,
        "pcie_state_round_trip": pcie,
        "graph_open": graph_open,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
