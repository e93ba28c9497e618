#!/usr/bin/env python3
"""Contact grad+Hessian+CCD throughput of libipc_warp_amd on one MI355X.

Workload (the wrecking-ball ABD configuration): 64 icospheres (642 vertices,
1280 triangles, r = 0.5) on a 4x4x4 grid spaced 1.001, seeded N(0,1) body
velocities, h = 1e-2, dhat = 1e-3 (thickness 5e-4), kappa = 1e8, same-body
pairs excluded. Inputs are resident in HBM before timing starts.

One step is one Newton iteration's contact work, in the reference's order
(affine_body_dynamics.py:427-505, dynamic_contacts.py:299-308,
contact.py:539-585):
  1. detect_collision: LBVH culling + fp32 distance filter, EE and PT;
  2. contact_hessian_ee + _pt (one launch): barrier gradient + PSD Hessian
     into triplets;
  3. contact_energy_ee + _pt (one launch): barrier energy (FP64 sum);
  4. collision_free_step(dx = -h v): swept culling + FP64 CCD TOI.
value = (assembled contacts + CCD candidates) per second over the timed
steps.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

METRIC = "contact grad+Hessian+CCD pairs/s; Newton iters/step warm-start vs plain"
HBM_PEAK_GBS = 8000.0
H = 1.0e-2
THICKNESS = 5.0e-4
KAPPA = 1.0e8


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def build_workload(seed: int):
    from ipc_warp_amd import scenes

    soup, vel = scenes.icosphere_grid(seed=seed)
    dx = (-H * vel[soup.body]).astype(np.float32)
    return soup, dx


def cpu_baseline(soup, dx, threads: int, budget_s: float = 12.0):
    """The oracle restatement timed on the host over a bounded number of steps."""
    import oracle as O

    th = np.float32(THICKNESS)
    l0 = th * np.float32(2.0)
    x1 = (soup.x - dx).astype(np.float32)
    steps = 0
    pairs = 0
    t0 = time.perf_counter()
    while True:
        ee, _ = O.dcd_ee(soup.x, soup.edges, soup.body, th, th, True)
        pt, _ = O.dcd_pt(soup.x, soup.triangles, soup.body, th, th, True)
        O.contact_hessian("ee", soup.x, ee, l0, KAPPA, threads=threads)
        O.contact_hessian("pt", soup.x, pt, l0, KAPPA, threads=threads)
        O.contact_energy("ee", soup.x, ee, l0, KAPPA, threads=threads)
        O.contact_energy("pt", soup.x, pt, l0, KAPPA, threads=threads)
        qe, _ = O.ccd_candidates_ee(soup.x, x1, soup.edges, soup.body, 1e-7, True)
        qp, _ = O.ccd_candidates_pt(soup.x, x1, soup.triangles, soup.body, 1e-7, True)
        O.collision_time("ee", soup.x, x1, qe, relative=True, threads=threads)
        O.collision_time("pt", soup.x, x1, qp, relative=True, threads=threads)
        steps += 1
        pairs += len(ee) + len(pt) + len(qe) + len(qp)
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {
        "value": pairs / dt,
        "unit": "pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} full steps of the same workload (hash-grid DCD, Hessian, energy, CCD candidates + TOI) "
                  f"on {threads} thread(s) of {cpu_model()}",
    }


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "c2_contact_step_hbm_traffic.json")


def measured_traffic(kernel: str, stream: bool = False):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None when it has no entry.
    stream selects the 2^22-contact Hessian stream's launch over the C2 step's."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            entries = json.load(f)["kernels"].get(kernel, [])
    except (OSError, ValueError, KeyError):
        return None
    pick = [e for e in entries if (e["grid"] == 1 << 22) == stream]
    if not pick:
        return None
    return pick[0]["fetch_bytes"] + pick[0]["write_bytes"]


FP64_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "toi_stream_fp64.json")
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 vector


def measured_flops64(kind: str):
    """FP64 flops per launch of the TOI kernel on the 2^20-candidate CCD
    stream from the committed rocprofv3 PMC summary (VALU FP64 instruction
    counts x 64 lanes, FMA counted twice), or None when it has no entry."""
    try:
        with open(FP64_PROFILE) as f:
            return float(json.load(f)["kernels"][kind]["flops_per_launch"])
    except (OSError, ValueError, KeyError, TypeError):
        return None


WARMSTART_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ipc_warp_amd", "data",
                               "warmstart_cube_drop.npz")


def newton_comparison(seeds, steps, dev):
    """Newton iterations per time step on held-out seeded cube drops (two 8^3-cell
    Neo-Hookean cubes over a fixed ground, implicit Euler h = 1e-2, kappa = 1e9,
    dhat = 1e-3, 1e-6 relative energy tolerance) from the previous state
    (plain, the reference's start), from the constant-acceleration
    extrapolation, and from the learned warm start (committed network)."""
    from ipc_warp_amd import warmstart

    if not os.path.exists(WARMSTART_MODEL):
        return {"error": f"no trained model at {os.path.relpath(WARMSTART_MODEL)}"}
    ws = warmstart.WarmStart.load(WARMSTART_MODEL, dev)
    its = {"plain": [], "extrapolation": [], "warm_start": []}
    conv, cross, learned, rejected = True, 0, [], 0
    for seed in seeds:
        r = warmstart.compare_newton(seed, steps, ws, dev)
        for m in its:
            its[m] += r[m]["iterations"]
            conv = conv and all(r[m]["converged"])
            cross = max(cross, max(r[m]["intersections"]))
            rejected += r[m]["crossing_rejections"]
        learned += [c == 0 for c in r["warm_start"]["choice"]]
    mean = {m: float(np.mean(v)) for m, v in its.items()}
    return {
        "workload": "cube drop: two 8^3-cell Neo-Hookean tet cubes (E=1e5, nu=0.4, rho=1000) over a fixed ground, "
                    "h=1e-2, kappa=1e9, dhat=1e-3, Newton to 1e-6 relative energy decrease",
        "seeds": list(seeds),
        "steps_per_seed": steps,
        "iters_per_step": mean,
        "warm_start_vs_plain": mean["warm_start"] / mean["plain"],
        "all_converged": conv,
        "max_edge_triangle_crossings": cross,
        "line_search_moves_rejected_for_crossing": rejected,
        "learned_candidate_share": float(np.mean(learned)) if learned else 0.0,
    }


def fixed_path_step(soup, dx_np, steps, warmup, world_counts, dev):
    """ms per step of detect_collision + Hessian + energy + collision_free_step
    through FixedBodyContactSolver on the same state and motion; checks that
    the DCD sets have the world path's sizes and the TOI its value."""
    import torch

    from ipc_warp_amd.contact import FixedBodyContactSolver

    nb = int(soup.body.max()) + 1
    q = np.zeros((nb, 4, 3), np.float32)
    q[:, 1:] = np.eye(3, dtype=np.float32)[None]
    dq = np.zeros((nb, 4, 3), np.float32)
    first = np.searchsorted(soup.body, np.arange(nb))
    dq[:, 0] = dx_np[first]
    x = torch.from_numpy(soup.x).to(dev)
    S = FixedBodyContactSolver(x, x.clone(), torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                               torch.from_numpy(soup.body), thickness=THICKNESS, buffer_=THICKNESS, kappa=KAPPA, h=H,
                               disable_self_collision_=True)
    S.initialize_fixed_body_bvhs(torch.from_numpy(q.reshape(-1, 3)).to(dev))
    dx = torch.from_numpy(dx_np).to(dev)
    dq_t = torch.from_numpy(dq.reshape(-1, 3)).to(dev)
    b = torch.zeros_like(x)
    out = {}

    def step():
        b.zero_()
        S.detect_collision()
        S.compute_collision_hessian(b)
        e = S.collision_energy_raw()
        toi, fb = S.raw_toi_fixed(dx, dq_t)
        out.update(ee=S.n_contacts, pt=S.n_contacts_pt, toi=toi, fallback=fb, energy=e)
        return S.n_contacts + S.n_contacts_pt + sum(S.last_ccd_counts())

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pairs = 0
    for _ in range(steps):
        pairs += step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    S.set_kernel_timing(True)
    S.kernel_times()
    n_prof = max(1, min(steps, 50))
    for _ in range(n_prof):
        step()
    kt = {k: v[0] / n_prof for k, v in S.kernel_times().items() if v[1]}
    S.set_kernel_timing(False)
    return {
        "ms_per_step": el / steps * 1e3,
        "pairs_per_s": pairs / el,
        "ccd_candidates": list(S.last_ccd_counts()),
        "same_dcd_sizes_as_world": (out["ee"], out["pt"]) == (world_counts["ee"], world_counts["pt"]),
        "same_toi_as_world": out["toi"] == world_counts["toi"],
        "fallback": out["fallback"],
        "kernel_ms_per_step": kt,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0 = all host cores, max 16)")
    ap.add_argument("--newton-seeds", type=int, nargs="*", default=[100, 101])
    ap.add_argument("--newton-steps", type=int, default=20)
    ap.add_argument("--no-newton", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    from ipc_warp_amd.contact import ContactSolverBase

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    soup, dx_np = build_workload(args.seed)
    x = torch.from_numpy(soup.x).to(dev)
    xcs = x.clone()
    dx = torch.from_numpy(dx_np).to(dev)
    S = ContactSolverBase(x, xcs, torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                          torch.from_numpy(soup.body), thickness=THICKNESS, buffer_=THICKNESS, kappa=KAPPA, h=H,
                          disable_self_collision_=True)
    b = torch.zeros_like(x)
    e_dev = torch.zeros(1, dtype=torch.float64, device=dev)
    e_host = torch.zeros(1, dtype=torch.float64, pin_memory=True)

    counts = {}
    # Single-kernel entry points are bracketed with events on torch's current
    # stream (the stream the library launches on); builds, queries and TOI
    # launches inside the context are bracketed by the library itself.
    kev = {k: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for k in ("hessian", "energy")}
    kacc = {k: 0.0 for k in kev}

    def step(timed: bool):
        b.zero_()
        S.detect_collision()
        n_ee, n_pt = S.n_contacts, S.n_contacts_pt
        # the EE and PT halves of the Hessian and of the energy each share one
        # launch; the FP64 energy is read after the CCD
        if timed:
            kev["hessian"][0].record()
        S.compute_collision_hessian(b)
        if timed:
            kev["hessian"][1].record()
            kev["energy"][0].record()
        S.collision_energy_into(e_dev)
        if timed:
            kev["energy"][1].record()
        # the energy travels to pinned host memory ahead of the CCD, whose
        # own synchronisation then covers it
        e_host.copy_(e_dev, non_blocking=True)
        toi = S.raw_toi(dx)
        e = float(e_host[0])
        if timed:
            torch.cuda.synchronize()
            for k in kev:
                kacc[k] += kev[k][0].elapsed_time(kev[k][1])
        n_ccd = S.last_ccd_counts()
        counts.update(ee=n_ee, pt=n_pt, ccd_ee=n_ccd[0], ccd_pt=n_ccd[1], toi=toi, energy=e)
        return n_ee + n_pt + n_ccd[0] + n_ccd[1]

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pairs = 0
    for _ in range(args.steps):
        pairs += step(False)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    # instrumented pass: per-kernel durations
    n_prof = max(1, min(args.steps, 50))
    S.set_kernel_timing(True)
    S.kernel_times()
    for _ in range(n_prof):
        step(True)
    lib_times = S.kernel_times()
    S.set_kernel_timing(False)
    # detect_collision and collision_free_step launch each stage once for the
    # EE and PT halves together (recorded under the EE kind)
    fused = {"build_edge": "build", "query_dcd_ee": "query_dcd", "query_ccd_ee": "query_ccd", "toi_ee": "toi"}
    lib_times = {fused.get(k, k): v for k, v in lib_times.items()
                 if not (k in ("build_tri", "query_dcd_pt", "query_ccd_pt", "toi_pt") and v[1] == 0)}
    kernel_ms = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in lib_times.items()}
    per_step_ms = {k: v[0] / n_prof for k, v in lib_times.items()}
    for k in kacc:
        kernel_ms[k] = kacc[k] / n_prof
        per_step_ms[k] = kacc[k] / n_prof

    # Algorithmic bytes per launch: barrier assembly 816 B (PT) / 864 B (EE) per
    # contact; TOI 112 B per candidate; a query reads the tree once (leaf:
    # 24 B box + 4 B body + 4 B primitive id + 4 B per vertex id; node: 24 B
    # box + 8 B body range), one 32 B (EE) or 12/24 B (PT DCD/CCD) query box
    # per query and writes 16 B (+8 B ids in DCD) per emitted pair. A DCD or
    # CCD query launch serves both halves, so its bytes are the EE and PT sums.
    E, F, V = int(soup.edges.shape[0]), int(soup.triangles.shape[0]), int(soup.x.shape[0])
    tree = lambda n, nv: (32 + 4 * nv) * n + 32 * ((n + 6) // 7)
    alg = {
        "hessian": 864 * counts["ee"] + 816 * counts["pt"],
        "energy": 112 * counts["ee"] + 64 * counts["pt"],
        "query_dcd": tree(E, 2) + 32 * E + 24 * counts["ee"] + tree(F, 3) + 12 * V + 24 * counts["pt"],
        "query_ccd": tree(E, 2) + 32 * E + 16 * counts["ccd_ee"] + tree(F, 3) + 24 * V + 16 * counts["ccd_pt"],
        "toi": 112 * (counts["ccd_ee"] + counts["ccd_pt"]),
    }
    dominant = max((k for k in alg if kernel_ms.get(k, 0) > 0), key=lambda k: per_step_ms[k])
    ach = alg[dominant] / (kernel_ms[dominant] * 1e-3) / 1e9

    # the barrier Hessian kernel on a large synthetic contact stream (2^22 PT contacts)
    from ipc_warp_amd import ops, scenes
    n_big = 1 << 22
    xs, cs = scenes.pair_stream("pt", n_big, seed=args.seed)
    xs_t = torch.from_numpy(xs).to(dev)
    cs_t = torch.from_numpy(cs).to(dev)
    rows = torch.empty(16 * n_big, dtype=torch.int32, device=dev)
    cols = torch.empty_like(rows)
    vals = torch.empty((16 * n_big, 3, 3), dtype=torch.float32, device=dev)
    bb = torch.zeros_like(xs_t)
    l0 = float(np.float32(THICKNESS) * np.float32(2))
    for _ in range(3):
        ops.contact_hessian_pt(xs_t, cs_t, n_big, l0, KAPPA, rows, cols, vals, bb, 0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        ops.contact_hessian_pt(xs_t, cs_t, n_big, l0, KAPPA, rows, cols, vals, bb, 0)
    e1.record()
    torch.cuda.synchronize()
    big_ms = e0.elapsed_time(e1) / 10
    big_gbs = 816 * n_big / (big_ms * 1e-3) / 1e9
    del rows, cols, vals, xs_t, cs_t, bb

    # The TOI kernel on synthetic CCD candidate streams (2^20 EE and 2^20 PT
    # candidates, coordinates ~10, ~30 % impacts), the minimum folded on the
    # device: HBM fraction from the 112 B per candidate, FP64-VALU fraction
    # from the committed PMC counts.
    n_ccd = 1 << 20
    toi_stream = {}
    for kind in ("ee", "pt"):
        arrays = scenes.ccd_stream(kind, n_ccd, seed=args.seed)[:3]
        x0t, x1t, qt = (torch.from_numpy(a).to(dev) for a in arrays)
        tmin = ops.new_toi_min(dev)
        for _ in range(3):
            ops.collision_time(kind, x0t, x1t, qt, n_ccd, None, tmin)
        e0.record()
        for _ in range(10):
            ops.collision_time(kind, x0t, x1t, qt, n_ccd, None, tmin)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        gbs = 112 * n_ccd / (ms * 1e-3) / 1e9
        flops = measured_flops64(kind)
        tflops = None if flops is None else flops / (ms * 1e-3) / 1e12
        toi_stream[kind] = {
            "kernel": f"collision_time_{kind}",
            "candidates": n_ccd,
            "ms": ms,
            "candidates_per_s": n_ccd / (ms * 1e-3),
            "hbm": {"achieved": gbs, "unit": "GB/s", "peak": HBM_PEAK_GBS, "frac": gbs / HBM_PEAK_GBS,
                    "algorithmic_bytes_per_launch": 112 * n_ccd},
            "fp64": None if flops is None else {"achieved": tflops, "unit": "TFLOP/s", "peak": FP64_PEAK_TFLOPS,
                                                "frac": tflops / FP64_PEAK_TFLOPS, "flops_per_launch": flops,
                                                "source": os.path.relpath(FP64_PROFILE, os.path.dirname(
                                                    os.path.abspath(__file__)))},
        }
        del x0t, x1t, qt

    # The same step through the fixed per-body BVHs (the reference's
    # FixedBVHWreckingBalls option): identity frames, each body translated by
    # its dx over the step.
    fixed = fixed_path_step(soup, dx_np, args.steps, args.warmup, counts, dev)

    # STREAM-style copy of 256 MiB: the measured HBM ceiling, and the
    # calibration kernel for the FETCH_SIZE / WRITE_SIZE counters.
    copy_src = torch.empty(64 << 20, dtype=torch.float32, device=dev)
    copy_dst = torch.empty_like(copy_src)
    for _ in range(3):
        copy_dst.copy_(copy_src)
    e0.record()
    for _ in range(10):
        copy_dst.copy_(copy_src)
    e1.record()
    torch.cuda.synchronize()
    copy_ms = e0.elapsed_time(e1) / 10
    copy_gbs = 2 * copy_src.numel() * 4 / (copy_ms * 1e-3) / 1e9
    del copy_src, copy_dst

    result = {
        "metric": METRIC,
        "value": pairs / elapsed,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "wrecking-ball ABD: 64 icospheres (642 v / 1280 t, r=0.5), 4x4x4 grid spaced 1.001, "
                        "N(0,1) velocities, h=1e-2, dhat=1e-3, kappa=1e8, same-body pairs excluded",
            "seed": args.seed,
            "vertices": V,
            "edges": E,
            "triangles": F,
        },
        "pairs_per_step": {"ee_contacts": counts["ee"], "pt_contacts": counts["pt"],
                           "ccd_ee_candidates": counts["ccd_ee"], "ccd_pt_candidates": counts["ccd_pt"]},
        "kernel_ms_per_launch": kernel_ms,
        "kernel_ms_per_step": per_step_ms,
        "toi": counts["toi"],
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": ach,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS,
            "traffic": measured_traffic(dominant),
            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, os.path.dirname(os.path.abspath(__file__))),
            "algorithmic_bytes_per_launch": alg[dominant],
            "avg_launch_ms": kernel_ms[dominant],
        },
        "hessian_stream": {
            "kernel": "contact_hessian_pt",
            "contacts": n_big,
            "ms": big_ms,
            "contacts_per_s": n_big / (big_ms * 1e-3),
            "achieved": big_gbs,
            "unit": "GB/s",
            "frac": big_gbs / HBM_PEAK_GBS,
            "traffic": measured_traffic("hessian_pt", stream=True),
            "algorithmic_bytes_per_launch": 816 * n_big,
        },
        "toi_stream": toi_stream,
        "hbm_copy_gbs": copy_gbs,
        "fixed_body_bvh_step": fixed,
    }
    if not args.no_newton:
        result["newton"] = newton_comparison(args.newton_seeds, args.newton_steps, dev)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        result["cpu_baseline"] = cpu_baseline(soup, dx_np, threads)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
