#!/usr/bin/env python3
"""Contact grad+Hessian+CCD throughput of libipc_warp_amd on one MI355X.

Headline workload (the largest configuration that fits one GPU): the peak
of a rod twist. 16 hexagonal-prism tet rods (radius 0.1, length 10, 400
segments, 28,800 tets each) in a 4x4 bundle spaced 0.25, both end caps
turned at 2 rad/s for 200 steps of 0.025 s (rods_twist.py:85); the scene is
scaled to unit extent as rods_twist.py:108-113 does, contact thickness and
buffer 1e-3 (rods_twist.py:38,91-92), kappa 1e9, self-contact on (FEM).
121,904 vertices, 231,552 edges, 154,368 triangles. The peak state is
generated (scenes.rod_bundle), not simulated. Inputs are resident in HBM
before timing starts.

One step is one Newton iteration's contact work in the reference's order
(RodComplexBC.compute_A, dynamic_contacts.py:275-297; RodBCBase.line_search,
time_integrator_base.py:514-546; compute_collision_energy,
dynamic_contacts.py:299-308; collision_free_step, contact.py:539-585):
  1. detect_collision at x_k: BVH culling + fp32 distance filter, EE and PT;
  2. contact_hessian_ee + _pt (one launch): barrier gradient + PSD Hessian
     into triplets, force into b;
  3. contact_energy_ee + _pt (one launch) at x_k (E0's contact term);
  4. collision_free_step(dx): swept culling + FP64 CCD TOI, alpha = 0.9 toi;
  5. the line search's first trial: x_k+1 = x_k - alpha dx, then
     detect_collision + energy at x_k+1 (E1's contact term).
The state moves: every 4 steps (a time step) x restarts from the next of 8
precomputed twist states (successive turns of the ends, each with its own
step direction dx), and within a time step x advances by alpha dx.
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
ROOT = os.path.dirname(os.path.abspath(__file__))

ROD_THICKNESS = 1.0e-3
ROD_KAPPA = 1.0e9
STATES = 8  # precomputed time-step states
ITERS_PER_STATE = 4  # Newton iterations per time step

GRID_H = 1.0e-2
GRID_THICKNESS = 5.0e-4
GRID_KAPPA = 1.0e8


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def rod_states(seed: int):
    """8 successive twist states (200, 201, ... steps) of the rod bundle, each
    with the step direction dx of its time step."""
    from ipc_warp_amd import scenes

    out = []
    for k in range(STATES):
        soup, dx, _ = scenes.rod_bundle(seed=seed, steps=200 + k)
        out.append((soup, dx))
    return out


# Algorithmic bytes per launch (the units of the measurement table): barrier
# assembly 816 B (PT) / 864 B (EE) per contact; energy 64 / 112 B per
# contact; TOI 112 B per candidate; a query reads the tree once (leaf: 24 B
# box + 4 B body + 4 B primitive id + 4 B per vertex id; node: 24 B box + 8 B
# body range), one 32 B (EE) or 12/24 B (PT DCD/CCD) query box per query and
# writes 16 B (+8 B ids in DCD) per emitted pair. A DCD or CCD query launch
# serves both halves, so its bytes are the EE and PT sums.
def algorithmic_bytes(E, F, V, c):
    tree = lambda n, nv: (32 + 4 * nv) * n + 32 * ((n + 6) // 7)
    return {
        "hessian": 864 * c["ee"] + 816 * c["pt"],
        "energy": 112 * c["ee"] + 64 * c["pt"],
        "query_dcd": tree(E, 2) + 32 * E + 24 * c["ee"] + tree(F, 3) + 12 * V + 24 * c["pt"],
        "query_ccd": tree(E, 2) + 32 * E + 16 * c["ccd_ee"] + tree(F, 3) + 24 * V + 16 * c["ccd_pt"],
        "toi": 112 * (c["ccd_ee"] + c["ccd_pt"]),
    }


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "rod_twist_hbm_traffic.json")
STREAM_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "kernel_streams_hbm_traffic.json")


def measured_traffic(kernel: str, path: str = TRAFFIC_PROFILE, grid: int | None = None):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None when it has no entry.
    grid selects the launch of that grid size (the synthetic streams)."""
    try:
        with open(path) as f:
            entries = json.load(f)["kernels"].get(kernel, [])
    except (OSError, ValueError, KeyError):
        return None
    pick = [e for e in entries if grid is None or e["grid"] == grid]
    if not pick:
        return None
    return pick[0]["fetch_bytes"] + pick[0]["write_bytes"]


ISSUE_PROFILE = os.path.join(ROOT, "profiles", "rod_twist_issue_counters.json")


def measured_issue(kernel: str):
    """VALU issue of `kernel` on the rod-twist step from the committed
    rocprofv3 PMC summary (VALUBusy: share of the CUs' time with a VALU
    instruction issuing; VALU instructions per launch), or None."""
    try:
        with open(ISSUE_PROFILE) as f:
            e = json.load(f)["kernels"][kernel]
    except (OSError, ValueError, KeyError):
        return None
    return {"valu_busy": e["valu_busy"], "valu_insts_per_launch": e["valu_insts"],
            "valu_utilization": e["valu_utilization"], "source": os.path.relpath(ISSUE_PROFILE, ROOT)}


FP64_PROFILE = os.path.join(ROOT, "profiles", "toi_stream_fp64.json")
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


def cpu_baseline_and_precision(soup, dx, gpu, threads: int, budget_s: float = 12.0):
    """The oracle restatement timed on the host over whole steps of the
    workload (at least one), and, as the checker, the GPU's first-step
    outputs against the oracle in the library's policy (mixed), the
    all-FP64 mode and the reference's FP32 semantics (ref32, and the
    absolute-coordinate CCD cubic)."""
    import oracle as O

    th = np.float32(ROD_THICKNESS)
    l0 = th * np.float32(2.0)
    x1 = (soup.x - dx).astype(np.float32)
    steps = 0
    pairs = 0
    t0 = time.perf_counter()
    while True:
        ee, _ = O.dcd_ee(soup.x, soup.edges, soup.body, th, th, False)
        pt, _ = O.dcd_pt(soup.x, soup.triangles, soup.body, th, th, False)
        O.contact_hessian("ee", soup.x, ee, l0, ROD_KAPPA, xcs=soup.xcs, threads=threads)
        O.contact_hessian("pt", soup.x, pt, l0, ROD_KAPPA, threads=threads)
        O.contact_energy("ee", soup.x, ee, l0, ROD_KAPPA, xcs=soup.xcs, threads=threads)
        O.contact_energy("pt", soup.x, pt, l0, ROD_KAPPA, threads=threads)
        qe, _ = O.ccd_candidates_ee(soup.x, x1, soup.edges, soup.body, 1e-7, False)
        qp, _ = O.ccd_candidates_pt(soup.x, x1, soup.triangles, soup.body, 1e-7, False)
        O.collision_time("ee", soup.x, x1, qe, relative=True, threads=threads)
        O.collision_time("pt", soup.x, x1, qp, relative=True, threads=threads)
        steps += 1
        pairs += len(ee) + len(pt) + len(qe) + len(qp)
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    baseline = {
        "value": pairs / dt,
        "unit": "pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} full step(s) of the same workload at its first state (hash-grid DCD, Hessian, energy, "
                  f"CCD candidates + TOI), {dt:.1f} s on {threads} thread(s) of {cpu_model()}; host nproc "
                  f"{os.cpu_count()}, one GPU's share is 16 threads",
        "nproc": os.cpu_count(),
    }

    # precision of the GPU's first-step outputs (the checker's comparison)
    g_ee, g_pt = gpu["ee"], gpu["pt"]
    n = len(g_ee) + len(g_pt)
    report = {"contacts": n}
    same = np.concatenate([
        O.contact_points("ee", soup.x, g_ee, xcs=soup.xcs)[2] ==
        O.contact_points("ee", soup.x, g_ee, xcs=soup.xcs, double=True)[2],
        O.contact_points("pt", soup.x, g_pt)[2] == O.contact_points("pt", soup.x, g_pt, double=True)[2]])
    for mode in ("mixed", "fp64", "ref32"):
        r_ee = O.contact_hessian("ee", soup.x, g_ee, l0, ROD_KAPPA, xcs=soup.xcs, n_total=n, mode=mode,
                                 threads=threads)
        r_pt = O.contact_hessian("pt", soup.x, g_pt, l0, ROD_KAPPA, offset=len(g_ee), n_total=n, mode=mode,
                                 threads=threads)
        r_vals = (r_ee[2] + r_pt[2]).reshape(n, -1)
        r_b = r_ee[3] + r_pt[3]
        r_e = O.contact_energy("ee", soup.x, g_ee, l0, ROD_KAPPA, xcs=soup.xcs, mode=mode, threads=threads) + \
            O.contact_energy("pt", soup.x, g_pt, l0, ROD_KAPPA, mode=mode, threads=threads)
        err = np.abs(gpu["vals"].reshape(n, -1) - r_vals).max(1) / np.maximum(np.abs(r_vals).max(1), 1e-300)
        report[mode] = {
            "block_max": float(err.max()),
            "block_max_same_feature": float(err[same].max()),
            "block_p999": float(np.quantile(err, 0.999)),
            "force": float(np.abs(gpu["b"] - r_b).max() / np.abs(r_b).max()),
            "energy": float(abs(gpu["energy"] - r_e) / abs(r_e)),
        }
    report["fp64"]["pairs_with_other_feature"] = int((~same).sum())
    tois = {}
    for kind, q in (("ee", gpu["q_ee"]), ("pt", gpu["q_pt"])):
        rel, _ = O.collision_time(kind, soup.x, x1, q, relative=True, threads=threads)
        ab, _ = O.collision_time(kind, soup.x, x1, q, relative=False, threads=threads)
        tois[kind] = (rel, ab)
    rel = np.concatenate([tois["ee"][0], tois["pt"][0]])
    ab = np.concatenate([tois["ee"][1], tois["pt"][1]])
    g_t = np.concatenate([gpu["toi_ee"], gpu["toi_pt"]])
    d = np.abs(g_t - ab)
    report["toi"] = {
        "candidates": int(len(rel)),
        "max_vs_oracle": float(np.abs(g_t - rel).max()),
        "vs_absolute_cubic": {"disagree_1e-5": int((d > 1e-5).sum()), "max_on_agreeing": float(d[d <= 1e-5].max()),
                              "step_bound": float(ab.min())},
        "step_bound": float(g_t.min()),
    }
    report["modes"] = ("mixed = the library's policy (fp32 classification, FP64 feature distance, barrier and log "
                       "term); fp64 = all-FP64 oracle (classification in double); ref32 = the reference's FP32 "
                       "semantics (fp32 distance in the barrier); errors: blocks per pair relative to the pair's "
                       "largest entry, force relative to the largest force, energy relative")
    return baseline, report


def newton_comparison(seeds, steps, dev):
    """Newton iterations per time step on held-out seeded cube drops (two 8^3-cell
    Neo-Hookean cubes over a fixed ground, implicit Euler h = 1e-2, kappa = 1e9,
    dhat = 1e-3, 1e-6 relative energy tolerance) from the previous state
    (plain, the reference's start), from the constant-acceleration
    extrapolation, and from the learned warm start (committed network)."""
    from ipc_warp_amd import warmstart

    model = os.path.join(ROOT, "ipc_warp_amd", "data", "warmstart_cube_drop.npz")
    if not os.path.exists(model):
        return {"error": f"no trained model at {os.path.relpath(model, ROOT)}"}
    ws = warmstart.WarmStart.load(model, dev)
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


def frozen_step(soup, dx_np, q, dq, exclude, thickness, kappa, steps, warmup, dev, fixed_bodies=None):
    """ms per step of detect_collision + Hessian + energy + collision_free_step on
    one state (world BVH), and the same through the per-body BVHs when q is
    given; checks that both paths find the same contact counts and TOI."""
    import torch

    from ipc_warp_amd.contact import ContactSolverBase, FixedBodyContactSolver

    x = torch.from_numpy(soup.x).to(dev)
    dx = torch.from_numpy(dx_np).to(dev)
    out = {}
    for path in ("world", "fixed") if q is not None else ("world",):
        cls = FixedBodyContactSolver if path == "fixed" else ContactSolverBase
        S = cls(x, torch.from_numpy(soup.xcs).to(dev), torch.from_numpy(soup.edges),
                torch.from_numpy(soup.triangles), torch.from_numpy(soup.body), thickness=thickness,
                buffer_=thickness, kappa=kappa, disable_self_collision_=exclude)
        if path == "fixed":
            S.initialize_fixed_body_bvhs(torch.from_numpy(q).to(dev), fixed_bodies=fixed_bodies)
            dq_t = torch.from_numpy(dq).to(dev)
        b = torch.zeros_like(x)
        res = {}

        def step():
            b.zero_()
            S.detect_collision()
            S.compute_collision_hessian(b)
            e = S.collision_energy_raw()
            if path == "fixed":
                toi, fb = S.raw_toi_fixed(dx, dq_t)
            else:
                toi, fb = S.raw_toi(dx), False
            res.update(ee=S.n_contacts, pt=S.n_contacts_pt, toi=toi, fallback=fb, energy=e)
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
        out[path] = {"ms_per_step": el / steps * 1e3, "pairs_per_s": pairs / el, "ee_contacts": res["ee"],
                     "pt_contacts": res["pt"], "ccd_candidates": list(S.last_ccd_counts()), "toi": res["toi"],
                     "fallback": res["fallback"]}
        del S
    if "fixed" in out:
        out["same_contacts_and_toi"] = (out["fixed"]["ee_contacts"], out["fixed"]["pt_contacts"],
                                        out["fixed"]["toi"]) == (out["world"]["ee_contacts"],
                                                                 out["world"]["pt_contacts"], out["world"]["toi"])
    return out


def bsr_assembly(S, x0, soup, dev, reps: int = 20):
    """compute_A's sparse assembly on the workload's first state
    (dynamic_contacts.py:290-297): the contact triplets to BSR, the same
    pattern refilled, and A = h^2 H_c + K over the union with an elastic
    pattern K (the tets' 16 vertex blocks each, random values)."""
    import torch

    from ipc_warp_amd import sparse

    S.x.copy_(x0)
    S.detect_collision()
    b = torch.zeros_like(S.x)
    trip = S.compute_collision_hessian(b)
    nv = S.x.shape[0]
    m = int(trip.rows.shape[0])
    H = sparse.bsr_from_triplets(nv, nv, trip.rows, trip.cols, trip.vals)
    tets = torch.from_numpy(soup.tets.astype(np.int32)).to(dev)
    kr = tets[:, :, None].expand(-1, 4, 4).reshape(-1).contiguous()
    kc = tets[:, None, :].expand(-1, 4, 4).reshape(-1).contiguous()
    g = torch.Generator(device=dev).manual_seed(0)
    K = sparse.bsr_from_triplets(nv, nv, kr, kc, torch.randn((kr.shape[0], 3, 3), device=dev, generator=g))
    A = sparse.BsrMatrix(nv, nv, None, None, None)

    def timed(fn):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    build_ms = timed(lambda: sparse.bsr_set_from_triplets(H, trip.rows, trip.cols, trip.vals))
    refill_ms = timed(lambda: H.refill(trip.vals))

    def axpy():
        A.offsets, A.columns, A.values = K.offsets, K.columns, K.values
        sparse.bsr_axpy(H, A, 0.025 ** 2, 1.0)

    axpy_ms = timed(axpy)
    nnz_h, nnz_k = H.nnz_sync(), K.nnz_sync()
    nnz_a = A.nnz_sync()
    kept = int((trip.vals.reshape(m, 9) != 0).any(1).sum())
    # algorithmic bytes: triplets read once (rows, cols, values: 44 B) and
    # blocks written once (column + values: 40 B) plus the offsets; a refill
    # reads the kept triplets' values and writes the block values; axpy reads
    # both operands' blocks and writes the union's
    alg = {"build": 44 * m + 40 * nnz_h + 4 * (nv + 1), "refill": 36 * kept + 36 * nnz_h,
           "axpy": 40 * (nnz_h + nnz_k + nnz_a) + 12 * (nv + 1)}
    ms = {"build": build_ms, "refill": refill_ms, "axpy": axpy_ms}
    out = {"triplets": m, "kept_triplets": kept, "contact_blocks": nnz_h, "elastic_blocks": nnz_k,
           "system_blocks": nnz_a}
    for k in ms:
        gbs = alg[k] / (ms[k] * 1e-3) / 1e9
        out[k] = {"ms": ms[k], "algorithmic_bytes": alg[k], "achieved": gbs, "unit": "GB/s",
                  "frac": gbs / HBM_PEAK_GBS}
    return out


def newton_system_cg(S, x0, soup, dev, h: float = 0.025):
    """The linear solve of one rod-twist Newton iteration (time_integrator_base.py:
    500-507): A = M + h^2 (K + H_c) assembled with ipc_bsr_from_triplets (K the
    stable Neo-Hookean tet Hessian of ipc_tet_hessian, fem/fem.py:138-203,
    E = 1e5, nu = 0.4, rho = 1000; H_c the contact Hessian) and solved by
    ipc_bsr_cg to the
    reference's tolerance 1e-4, unpreconditioned (as the reference calls cg)
    and with the Jacobi preconditioner, at most 5000 iterations (the barrier
    terms make this system far too ill-conditioned for CG to reach 1e-4; the
    reference's rod scenes use its direct solver)."""
    import torch

    from ipc_warp_amd import ops, sparse
    from ipc_warp_amd.integrator import lame

    S.x.copy_(x0)
    S.detect_collision()
    trip = S.compute_collision_hessian(torch.zeros_like(S.x))
    nv = S.x.shape[0]
    T = torch.from_numpy(soup.tets.astype(np.int32)).to(dev)
    n_tet = int(T.shape[0])
    mu, lam = lame(1e5, 0.4)
    Bm, W = ops.tet_rest(torch.from_numpy(soup.xcs).to(dev), T)
    k_rows = torch.empty(16 * n_tet, dtype=torch.int32, device=dev)
    k_cols = torch.empty_like(k_rows)
    k_vals = torch.empty((16 * n_tet, 3, 3), dtype=torch.float32, device=dev)
    ops.tet_hessian(S.x, T, Bm, W, mu, lam, k_rows, k_cols, k_vals, torch.zeros_like(S.x))
    m = torch.zeros(nv, dtype=torch.float64, device=dev)
    m.index_add_(0, T.reshape(-1).long(), (1000.0 * W.double() / 4.0).repeat_interleave(4))
    m[m == 0] = 1.0
    idx = torch.arange(nv, device=dev, dtype=torch.int32)
    rows = torch.cat([k_rows, trip.rows, idx]).contiguous()
    cols = torch.cat([k_cols, trip.cols, idx]).contiguous()
    vals = torch.cat([h * h * k_vals, h * h * trip.vals, m.float()[:, None, None] * torch.eye(3, device=dev)]).contiguous()
    A = sparse.bsr_from_triplets(nv, nv, rows, cols, vals)
    nnz = A.nnz_sync()
    g = torch.Generator(device=dev).manual_seed(0)
    b = torch.randn((nv, 3), device=dev, generator=g)
    out = {"block_rows": nv, "blocks": nnz, "tol": 1e-4}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    bn = float(torch.linalg.norm(b.double()))
    for name, jac in (("plain", False), ("jacobi", True)):
        x = torch.zeros_like(b)
        sparse.cg(A, b, x, tol=1e-4, maxiter=20, jacobi=jac)
        x.zero_()
        torch.cuda.synchronize()
        e0.record()
        it, res, goal = sparse.cg(A, b, x, tol=1e-4, maxiter=5000, jacobi=jac)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        # per iteration: the product reads each block (36 B values + 4 B column)
        # and gathers x; the vector passes move about 14 floats per unknown
        alg = 40 * nnz + 4 * (nv + 1) + 14 * 4 * 3 * nv
        out[name] = {"iterations": it, "converged": res <= goal, "relative_residual": res / bn, "ms": ms,
                     "ms_per_iteration": ms / max(it, 1),
                     "achieved_per_iteration": alg / (ms / max(it, 1) * 1e-3) / 1e9, "unit": "GB/s"}
    return out


def tet_elasticity(soup, x0, dev, reps: int = 20):
    """The rod twist's elastic tets (460,800; fem/fem.py:138-203 and
    time_integrator_base.py:124-137) at the workload's first state: the
    stable Neo-Hookean Hessian into triplets with the forces, and the
    energy. Algorithmic bytes per tet: tet 16 B, 4 vertices 48 B, Bm 36 B,
    W 4 B read; 16 triplets 704 B and 4 forces 48 B written."""
    import torch

    from ipc_warp_amd import ops

    T = torch.from_numpy(soup.tets.astype(np.int32)).to(dev)
    n = int(T.shape[0])
    Bm, W = ops.tet_rest(torch.from_numpy(soup.xcs).to(dev), T)
    rows = torch.empty(16 * n, dtype=torch.int32, device=dev)
    cols = torch.empty_like(rows)
    vals = torch.empty((16 * n, 3, 3), dtype=torch.float32, device=dev)
    b = torch.zeros_like(x0)
    e = torch.zeros(1, dtype=torch.float64, device=dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {"tets": n}
    for name, fn, alg in (("hessian", lambda: ops.tet_hessian(x0, T, Bm, W, 2.0e6, 2.0e7, rows, cols, vals, b), 856),
                          ("energy", lambda: ops.tet_energy(x0, T, Bm, W, 2.0e6, 2.0e7, None, e), 104)):
        for _ in range(3):
            fn()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gbs = alg * n / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "tets_per_s": n / (ms * 1e-3), "achieved": gbs, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes": alg * n}
    return out


def al_active_set(S, X0, DX, dev, reps: int = 20):
    """update_active_set (augmented_lagrangian.py:705-717) on the rod twist,
    the geometry RodsTwistAL runs the barrier-free solver on (:750-767):
    the active sets hold the pairs met between the first state and its full
    Newton step; the pairs met by the second state's full step are then
    recorded once (new_intersections), and the timed hash work is the rest
    of the update: per-vertex earliest TOI, create_hash (EE, PT) and merge
    into the active sets, with the host count readbacks ActiveSet.count
    makes."""
    import torch

    from ipc_warp_amd.active_set import ALActiveSets, accumulate_vertex_toi, xconstraint_records

    sets = ALActiveSets(X0[0].shape[0], dev, initial_capacity=1 << 12)
    sets.update_active_set(S, X0[0], (X0[0] - DX[0]).contiguous())
    old = (sets.active_set_ee.hash, sets.active_set_pt.hash)
    n_old = (sets.active_set_ee.count(), sets.active_set_pt.count())
    x1 = (X0[1] - DX[1]).contiguous()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        S.new_intersections(X0[1], x1)
    torch.cuda.synchronize()
    ccd_ms = (time.perf_counter() - t0) * 1e3 / reps
    recs = [xconstraint_records(C.list, C.e0e1, C.alpha, n, S.l0) for C, n in
            ((S.contacts_new, S.n_contacts), (S.contacts_pt, S.n_contacts_pt))]

    def hash_work():
        sets.active_set_ee.hash, sets.active_set_pt.hash = old
        sets.vertex_toi.fill_(1.0)
        for r in recs:
            accumulate_vertex_toi(r, r.shape[0], sets.vertex_toi)
        sets.detected_set_ee.create_hash(recs[0], recs[0].shape[0], sets.vertex_toi)
        sets.detected_set_pt.create_hash(recs[1], recs[1].shape[0], sets.vertex_toi)
        sets.active_set_ee.merge(sets.detected_set_ee)
        sets.active_set_pt.merge(sets.detected_set_pt)

    for _ in range(3):
        hash_work()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        hash_work()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    detected = recs[0].shape[0] + recs[1].shape[0]
    return {"recorded_pairs": [int(recs[0].shape[0]), int(recs[1].shape[0])],
            "accepted": [sets.detected_set_ee.count(), sets.detected_set_pt.count()],
            "old_active": list(n_old), "merged_active": [sets.active_set_ee.count(), sets.active_set_pt.count()],
            "new_intersections_ms": ccd_ms, "hash_ms_per_update": ms,
            "recorded_pairs_per_s": detected / (ms * 1e-3)}


def abd_reduction(dev, seed: int, reps: int = 10):
    """The affine-body contact reduction (affine_body_dynamics.py:144-190)
    of the screw-and-nut contact Hessian: vertex triplets onto the screw's
    12 degrees of freedom (nut fixed)."""
    import torch

    from ipc_warp_amd.abd import screw_and_nut_integrator

    I = screw_and_nut_integrator(seed, dev)
    S = I.contact
    S.detect_collision()
    force = torch.zeros_like(I.x)
    trip = S.compute_collision_hessian(force)
    m = int(trip.rows.shape[0])
    args = (I.nb, I.body32, I.rest32, I.fixed, trip.rows, trip.cols, trip.vals, force, I.h * I.h)
    for _ in range(2):
        I.reducer.reduce(*args)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(reps):
        I.reducer.reduce(*args)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    # algorithmic bytes: the triplets once (44 B) and the body/rest gathers of
    # their two vertices (2 x 16 B); the force pass reads 28 B per vertex
    alg = 76 * m + 28 * I.V
    return {"triplets": m, "bodies": I.nb, "ms_per_call": ms, "includes": "pattern + one host readback + reduction",
            "achieved": alg / (ms * 1e-3) / 1e9, "unit": "GB/s", "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "algorithmic_bytes": alg}


def newton_screw(dev, seed: int, steps: int):
    """Newton iterations per step of the screw turning in its nut (affine
    bodies, nut fixed), from the previous state and from the inertial
    extrapolation; both the 1e-6 relative-potential criterion and the
    reference's max|dq| <= 0.05 h (time_integrator_base.py:392-396)."""
    from ipc_warp_amd.abd import compare_newton

    r = compare_newton(seed, steps, dev)
    out = {"workload": "screw and nut as affine bodies: h=1e-2, motor -pi rad/s about y, nut fixed, affine "
                       "stiffness 1e8, kappa 1e9, thickness 5e-4; Newton to 1e-6 relative potential decrease",
           "seed": seed, "steps": steps}
    for mode, v in r.items():
        out[mode] = {"iters_per_step": float(np.mean(v["iterations"])),
                     "reference_criterion_iters_per_step": float(np.mean(v["reference_criterion_iterations"])),
                     "all_converged": all(v["converged"]), "max_edge_triangle_crossings": max(v["intersections"]),
                     "iterations": v["iterations"]}
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (0 = the host's cores, at most 16: one GPU's share of the box)")
    ap.add_argument("--newton-seeds", type=int, nargs="*", default=[100, 101])
    ap.add_argument("--newton-steps", type=int, default=20)
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the streams and the other workloads")
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
    states = rod_states(args.seed)
    soup = states[0][0]
    X0 = [torch.from_numpy(s.x).to(dev) for s, _ in states]
    DX = [torch.from_numpy(d).to(dev) for _, d in states]
    x = X0[0].clone()
    S = ContactSolverBase(x, torch.from_numpy(soup.xcs).to(dev), torch.from_numpy(soup.edges),
                          torch.from_numpy(soup.triangles), torch.from_numpy(soup.body), thickness=ROD_THICKNESS,
                          buffer_=ROD_THICKNESS, kappa=ROD_KAPPA, h=0.025, disable_self_collision_=False)
    b = torch.zeros_like(x)
    e_dev = torch.zeros(2, dtype=torch.float64, device=dev)
    e_host = torch.zeros(2, dtype=torch.float64, pin_memory=True)

    counts = {}
    kev = {k: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for k in ("hessian", "energy")}
    kacc = {k: 0.0 for k in kev}
    kn = {k: 0 for k in kev}

    def step(i: int, timed: bool):
        st = (i // ITERS_PER_STATE) % STATES
        if i % ITERS_PER_STATE == 0:
            x.copy_(X0[st])
        dx = DX[st]
        b.zero_()
        S.detect_collision()
        n_ee, n_pt = S.n_contacts, S.n_contacts_pt
        if timed:
            kev["hessian"][0].record()
        S.compute_collision_hessian(b)
        if timed:
            kev["hessian"][1].record()
            kev["energy"][0].record()
        S.collision_energy_into(e_dev[0:1])
        if timed:
            kev["energy"][1].record()
        toi = S.raw_toi(dx)
        alpha = 0.9 * toi if toi < 1.0 else 1.0
        x.sub_(dx, alpha=alpha)
        S.detect_collision()
        S.collision_energy_into(e_dev[1:2])
        e_host.copy_(e_dev)
        if timed:
            torch.cuda.synchronize()
            for k in kev:
                kacc[k] += kev[k][0].elapsed_time(kev[k][1])
                kn[k] += 1
        n_ccd = S.last_ccd_counts()
        counts.update(ee=n_ee, pt=n_pt, ccd_ee=n_ccd[0], ccd_pt=n_ccd[1], toi=toi)
        return n_ee + n_pt + n_ccd[0] + n_ccd[1]

    # first-step outputs for the precision report (checked on the host later)
    x.copy_(X0[0])
    S.detect_collision()
    b.zero_()
    trip = S.compute_collision_hessian(b)
    e_first = S.collision_energy_raw()
    first = {"ee": S.contacts_new.list[:S.n_contacts].cpu().numpy(),
             "pt": S.contacts_pt.list[:S.n_contacts_pt].cpu().numpy(),
             "vals": trip.vals.cpu().numpy().astype(np.float64), "b": b.cpu().numpy().astype(np.float64),
             "energy": e_first}
    x1_first = (X0[0] - DX[0]).contiguous()
    from ipc_warp_amd import ops
    for kind in ("ee", "pt"):
        q, _ = S.ccd_candidates(kind, X0[0], x1_first)
        t = torch.empty(len(q), dtype=torch.float64, device=dev)
        ops.collision_time(kind, X0[0], x1_first, q.contiguous(), len(q), t, None)
        first[f"q_{kind}"] = q.cpu().numpy()
        first[f"toi_{kind}"] = t.cpu().numpy()
    del trip

    for i in range(args.warmup):
        step(i, False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pairs = 0
    for i in range(args.steps):
        pairs += step(i, False)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    # instrumented pass: per-kernel durations (per step: 2 builds of each
    # tree, 2 DCD queries, 1 CCD query, 1 TOI launch, 1 Hessian, 2 energies)
    n_prof = max(ITERS_PER_STATE, min(args.steps, 48))
    S.set_kernel_timing(True)
    S.kernel_times()
    sums = {"ee": 0, "pt": 0, "ccd_ee": 0, "ccd_pt": 0}
    for i in range(n_prof):
        step(i, True)
        for k in sums:
            sums[k] += counts[k]
    lib_times = S.kernel_times()
    S.set_kernel_timing(False)
    fused = {"build_edge": "build", "query_dcd_ee": "query_dcd", "query_ccd_ee": "query_ccd", "toi_ee": "toi"}
    lib_times = {fused.get(k, k): v for k, v in lib_times.items()
                 if not (k in ("build_tri", "query_dcd_pt", "query_ccd_pt", "toi_pt") and v[1] == 0)}
    kernel_ms = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in lib_times.items() if v[1]}
    per_step_ms = {k: v[0] / n_prof for k, v in lib_times.items() if v[1]}
    kernel_ms["hessian"] = kacc["hessian"] / kn["hessian"]
    per_step_ms["hessian"] = kacc["hessian"] / n_prof
    # the energy launch at x_k is timed; the step runs it twice
    kernel_ms["energy"] = kacc["energy"] / kn["energy"]
    per_step_ms["energy"] = 2 * kacc["energy"] / n_prof
    mean_counts = {k: v / n_prof for k, v in sums.items()}
    E, F, V = int(soup.edges.shape[0]), int(soup.triangles.shape[0]), int(soup.x.shape[0])
    alg = algorithmic_bytes(E, F, V, mean_counts)
    dominant = max((k for k in alg if kernel_ms.get(k, 0) > 0), key=lambda k: per_step_ms[k])
    ach = alg[dominant] / (kernel_ms[dominant] * 1e-3) / 1e9
    hess_gbs = alg["hessian"] / (kernel_ms["hessian"] * 1e-3) / 1e9

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
            "workload": "rod twist peak: 16 hexagonal-prism tet rods (r=0.1, L=10, 400 segments, 28,800 tets each) "
                        "in a 4x4 bundle spaced 0.25, ends turned at 2 rad/s for 200 steps of 0.025 s, scene scaled "
                        "to unit extent, thickness=buffer=1e-3, kappa=1e9, self-contact on; one Newton iteration's "
                        "contact work per step (DCD, Hessian, energy, CCD, line-search DCD + energy), moving state",
            "seed": args.seed,
            "vertices": V,
            "edges": E,
            "triangles": F,
            "tets": int(soup.tets.shape[0]),
        },
        "pairs_per_step": {"ee_contacts": mean_counts["ee"], "pt_contacts": mean_counts["pt"],
                           "ccd_ee_candidates": mean_counts["ccd_ee"], "ccd_pt_candidates": mean_counts["ccd_pt"]},
        "kernel_ms_per_launch": kernel_ms,
        "kernel_ms_per_step": per_step_ms,
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": ach,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS,
            "traffic": measured_traffic(dominant),
            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT),
            "algorithmic_bytes_per_launch": alg[dominant],
            "avg_launch_ms": kernel_ms[dominant],
            # the query launches are bound by VALU issue, not by HBM
            "issue": measured_issue(dominant),
        },
        "hessian_on_workload": {"achieved": hess_gbs, "unit": "GB/s", "frac": hess_gbs / HBM_PEAK_GBS,
                                "traffic": measured_traffic("hessian"), "avg_launch_ms": kernel_ms["hessian"],
                                "algorithmic_bytes_per_launch": alg["hessian"]},
    }
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    if not args.no_extras:
        result["bsr_assembly"] = bsr_assembly(S, X0[0], soup, dev)
        result["newton_system_cg"] = newton_system_cg(S, X0[0], soup, dev)
        result["tet_elasticity"] = tet_elasticity(soup, X0[0], dev)
        result["al_active_set"] = al_active_set(S, X0, DX, dev)
    del S

    if not args.no_extras:
        from ipc_warp_amd import scenes
        # the barrier Hessian kernel on a large synthetic contact stream (2^22 PT contacts)
        n_big = 1 << 22
        xs, cs = scenes.pair_stream("pt", n_big, seed=args.seed)
        xs_t = torch.from_numpy(xs).to(dev)
        cs_t = torch.from_numpy(cs).to(dev)
        rows = torch.empty(16 * n_big, dtype=torch.int32, device=dev)
        cols = torch.empty_like(rows)
        vals = torch.empty((16 * n_big, 3, 3), dtype=torch.float32, device=dev)
        bb = torch.zeros_like(xs_t)
        l0 = float(np.float32(GRID_THICKNESS) * np.float32(2))
        for _ in range(3):
            ops.contact_hessian_pt(xs_t, cs_t, n_big, l0, GRID_KAPPA, rows, cols, vals, bb, 0)
        e0.record()
        for _ in range(10):
            ops.contact_hessian_pt(xs_t, cs_t, n_big, l0, GRID_KAPPA, rows, cols, vals, bb, 0)
        e1.record()
        torch.cuda.synchronize()
        big_ms = e0.elapsed_time(e1) / 10
        big_gbs = 816 * n_big / (big_ms * 1e-3) / 1e9
        del rows, cols, vals, xs_t, cs_t, bb
        result["hessian_stream"] = {
            "kernel": "contact_hessian_pt", "contacts": n_big, "ms": big_ms, "contacts_per_s": n_big / (big_ms * 1e-3),
            "achieved": big_gbs, "unit": "GB/s", "frac": big_gbs / HBM_PEAK_GBS,
            "traffic": measured_traffic("hessian_pt", STREAM_TRAFFIC_PROFILE, grid=n_big),
            "algorithmic_bytes_per_launch": 816 * n_big,
        }
        # the TOI kernel on synthetic CCD candidate streams (2^20 EE and 2^20
        # PT candidates, coordinates ~10, ~30 % impacts)
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
                "kernel": f"collision_time_{kind}", "candidates": n_ccd, "ms": ms,
                "candidates_per_s": n_ccd / (ms * 1e-3),
                "hbm": {"achieved": gbs, "unit": "GB/s", "peak": HBM_PEAK_GBS, "frac": gbs / HBM_PEAK_GBS,
                        "algorithmic_bytes_per_launch": 112 * n_ccd,
                        "traffic": measured_traffic(f"toi_{kind}", STREAM_TRAFFIC_PROFILE)},
                "fp64": None if flops is None else {"achieved": tflops, "unit": "TFLOP/s", "peak": FP64_PEAK_TFLOPS,
                                                    "frac": tflops / FP64_PEAK_TFLOPS, "flops_per_launch": flops,
                                                    "source": os.path.relpath(FP64_PROFILE, ROOT)},
            }
            del x0t, x1t, qt
        result["toi_stream"] = toi_stream

        # the other workloads, each on one state: the screw turning in its nut
        # (two affine bodies, world and per-body BVHs) and the 64-icosphere grid
        s3, q3, dq3, dx3, fb3 = scenes.screw_and_nut(seed=args.seed)
        result["screw_and_nut_step"] = dict(
            workload="screw (r=1, pitch 0.2, 4 turns, 512 columns x 32 rows per turn) turning at -pi rad/s "
                     "(h=1e-2) in a fixed 3-turn nut with 1e-3 radial clearance; 117,998 v, 353,092 e, 235,032 t; "
                     "thickness 5e-4, kappa 1e8, same-body pairs excluded; one state",
            **frozen_step(s3, dx3, q3, dq3, True, GRID_THICKNESS, GRID_KAPPA, 50, 5, dev, fixed_bodies=fb3))
        result["screw_and_nut_abd_reduction"] = abd_reduction(dev, args.seed)
        sg, vel = scenes.icosphere_grid(seed=args.seed)
        nb = sg.n_bodies
        qg = np.zeros((nb, 4, 3), np.float32)
        qg[:, 1:] = np.eye(3, dtype=np.float32)[None]
        dxg = (-GRID_H * vel[sg.body]).astype(np.float32)
        dqg = np.zeros((nb, 4, 3), np.float32)
        dqg[:, 0] = dxg[np.searchsorted(sg.body, np.arange(nb))]
        result["icosphere_grid_step"] = dict(
            workload="64 icospheres (642 v / 1280 t, r=0.5) on a 4x4x4 grid spaced 1.001, N(0,1) velocities, "
                     "h=1e-2, thickness 5e-4, kappa 1e8, same-body pairs excluded; one state",
            **frozen_step(sg, dxg, qg.reshape(-1, 3), dqg.reshape(-1, 3), True, GRID_THICKNESS, GRID_KAPPA, 100, 10,
                          dev))

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
    result["hbm_copy_gbs"] = 2 * copy_src.numel() * 4 / (copy_ms * 1e-3) / 1e9
    del copy_src, copy_dst

    if not args.no_newton:
        result["newton"] = newton_comparison(args.newton_seeds, args.newton_steps, dev)
        result["newton_screw_and_nut"] = newton_screw(dev, args.seed, 8)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        result["cpu_baseline"], result["precision"] = cpu_baseline_and_precision(states[0][0], states[0][1], first,
                                                                                 threads)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
