"""Train the learned Newton warm start on seeded cube drops and compare it
with plain Newton on held-out seeds.

  python -m ipc_warp_amd.train_warmstart --out model.npz [--train-seeds 0 1 2 3]
      [--steps 40] [--epochs 60] [--eval-seeds 100 101] [--eval-steps 30]
  python -m ipc_warp_amd.train_warmstart --model model.npz --eval-only

Trajectories come from plain Newton (ContactIntegrator, 1e-6 relative
energy tolerance); the network, its loss and Adam run in libipc_warp_amd.
Prints one line per step and a JSON summary at the end.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import warmstart


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=None)
    ap.add_argument("--model", default=None)
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--train-seeds", type=int, nargs="*", default=[0, 1, 2, 3, 4, 5])
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-seeds", type=int, nargs="*", default=[100, 101])
    ap.add_argument("--eval-steps", type=int, default=30)
    ap.add_argument("--cells", type=int, default=8)
    ap.add_argument("--save-data", default=None, help="write the collected (features, targets) to this .npz")
    ap.add_argument("--data", default=None, help="train on a saved .npz instead of collecting trajectories")
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("train_warmstart needs a GPU", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    summary = {}
    if args.eval_only:
        if not args.model:
            print("--eval-only needs --model", file=sys.stderr)
            return 2
        ws = warmstart.WarmStart.load(args.model, dev)
    else:
        t0 = time.time()
        Xs, Ts = [], []
        if args.data:
            z = np.load(args.data)
            Xs.append(torch.from_numpy(z["X"]).to(dev))
            Ts.append(torch.from_numpy(z["T"]).to(dev))
        for seed in ([] if args.data else args.train_seeds):
            sim = warmstart.drop_simulation(seed, dev, n=args.cells)

            def log(s, st, seed=seed):
                print(f"train seed {seed} step {s}: iters {st.iterations} conv {st.converged}", flush=True)

            X, T = warmstart.record_trajectory(sim, args.steps, log=log)
            Xs.append(X)
            Ts.append(T)
        X, T = torch.cat(Xs), torch.cat(Ts)
        if args.save_data:
            np.savez_compressed(args.save_data, X=X.cpu().numpy(), T=T.cpu().numpy())
        t1 = time.time()
        print(f"samples {X.shape[0]} in {t1 - t0:.1f} s", flush=True)

        def elog(e, loss):
            print(f"epoch {e}: loss {loss:.4e}", flush=True)

        mlp = warmstart.train(X, T, epochs=args.epochs, lr=args.lr, log=elog)
        torch.cuda.synchronize()
        summary["train"] = {"samples": int(X.shape[0]), "seeds": args.train_seeds, "steps": args.steps,
                            "epochs": args.epochs, "collect_s": t1 - t0, "train_s": time.time() - t1}
        ws = warmstart.WarmStart(mlp)
        if args.out:
            ws.save(args.out)
    evals = {}
    for seed in args.eval_seeds:
        def log(mode, s, st, n_cross, seed=seed):
            print(f"eval seed {seed} {mode} step {s}: iters {st.iterations} conv {st.converged} "
                  f"accepted {st.warm_start_accepted} crossings {n_cross}", flush=True)

        r = warmstart.compare_newton(seed, args.eval_steps, ws, dev, n=args.cells, log=log)
        evals[seed] = {m: {"mean_iterations": r[m]["mean_iterations"], "all_converged": all(r[m]["converged"]),
                           "max_intersections": max(r[m]["intersections"]), "iterations": r[m]["iterations"]}
                       for m in r}
    summary["eval"] = evals
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
