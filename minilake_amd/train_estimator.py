"""Train the learned-sizing network on the GPU and save its weights.

  python -m minilake_amd.train_estimator --tables 6000 --epochs 300

Training pairs come from the plain engine (exact group and join counts over
synthetic Zipf key columns generated in HBM); the network is trained with the
library's own forward/backward/Adam kernels. Prints the held-out error of the
prediction (in log2 units) and writes minilake_amd/estimator_weights.npy.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import sizing


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--tables", type=int, default=6000)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--min-rows", type=float, default=1e4)
    ap.add_argument("--max-rows", type=float, default=1e8)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=sizing.WEIGHTS_PATH)
    ap.add_argument("--data-out", default="")
    args = ap.parse_args(argv)
    from .gpu import GpuEngine

    def log(m):
        print(f"[train] {m}", flush=True)

    eng = GpuEngine()
    t0 = time.time()
    X, Y = sizing.training_set(eng, args.tables, seed=args.seed, min_rows=args.min_rows, max_rows=args.max_rows,
                               log=log)
    log(f"{args.tables} labelled tables in {time.time() - t0:.1f}s")
    if args.data_out:
        np.savez(args.data_out, X=X, Y=Y)
    n_val = max(1, args.tables // 10)
    Xt, Yt, Xv, Yv = X[n_val:], Y[n_val:], X[:n_val], Y[:n_val]
    mlp = sizing.Mlp(eng)
    t0 = time.time()
    mlp.train(Xt, Yt, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, log=log)
    log(f"trained in {time.time() - t0:.1f}s")
    pv = mlp.forward(Xv)
    err = np.abs(pv - Yv)
    stats = {"val_mae_log2_groups": float(err[:, 0].mean()), "val_mae_log2_join": float(err[:, 1].mean()),
             "val_p90_log2_groups": float(np.quantile(err[:, 0], 0.9)),
             "val_p90_log2_join": float(np.quantile(err[:, 1], 0.9)),
             "val_under_1.15x_groups": float(np.mean(pv[:, 0] + np.log2(1.15) < Yv[:, 0])),
             "tables": args.tables}
    np.save(args.out, mlp.params())
    log(json.dumps(stats))
    eng.close()


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    main()
