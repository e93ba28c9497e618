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
  2. contact_hessian_ee / _pt: barrier gradient + PSD Hessian into triplets;
  3. contact_energy_ee / _pt: barrier energy (FP64 sum);
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


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0 = all host cores, max 16)")
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

    # per-kernel-stage timers on torch's current stream (the library's stream)
    stage_names = ("dcd", "hessian", "energy", "ccd")
    ev = {k: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for k in stage_names}
    acc = {k: 0.0 for k in stage_names}
    counts = {}

    def step(timed: bool):
        b.zero_()
        if timed:
            ev["dcd"][0].record()
        S.detect_collision()
        if timed:
            ev["dcd"][1].record()
            ev["hessian"][0].record()
        S.compute_collision_hessian(b)
        if timed:
            ev["hessian"][1].record()
            ev["energy"][0].record()
        e = S.collision_energy_raw()
        if timed:
            ev["energy"][1].record()
            ev["ccd"][0].record()
        toi = S.raw_toi(dx)
        if timed:
            ev["ccd"][1].record()
            torch.cuda.synchronize()
            for k in stage_names:
                acc[k] += ev[k][0].elapsed_time(ev[k][1])
        n_ccd = S.last_ccd_counts()
        counts.update(ee=S.n_contacts, pt=S.n_contacts_pt, ccd_ee=n_ccd[0], ccd_pt=n_ccd[1], toi=toi, energy=e)
        return S.n_contacts + S.n_contacts_pt + n_ccd[0] + n_ccd[1]

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pairs = 0
    for _ in range(args.steps):
        pairs += step(False)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # stage breakdown in a separate, event-instrumented pass
    n_prof = max(1, min(args.steps, 50))
    for _ in range(n_prof):
        step(True)
    stage_ms = {k: acc[k] / n_prof for k in stage_names}

    # Roofline of the barrier assembly kernel family at this workload's
    # contact count: algorithmic bytes per contact (PT 816 B, EE 864 B).
    alg_bytes = counts["pt"] * 816 + counts["ee"] * 864
    hess_ms = stage_ms["hessian"]
    achieved = alg_bytes / (hess_ms * 1e-3) / 1e9 if hess_ms > 0 else 0.0

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
            "vertices": int(soup.x.shape[0]),
            "edges": int(soup.edges.shape[0]),
            "triangles": int(soup.triangles.shape[0]),
        },
        "pairs_per_step": {"ee_contacts": counts["ee"], "pt_contacts": counts["pt"],
                           "ccd_ee_candidates": counts["ccd_ee"], "ccd_pt_candidates": counts["ccd_pt"]},
        "stage_ms": stage_ms,
        "toi": counts["toi"],
        "roofline": {
            "kernel": "contact_hessian_ee + contact_hessian_pt",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        result["cpu_baseline"] = cpu_baseline(soup, dx_np, threads)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
