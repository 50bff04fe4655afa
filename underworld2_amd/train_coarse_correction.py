"""Train the learned coarse correction on the GPU and save it for ``learned.load_default``.

    python -m underworld2_amd.train_coarse_correction [--problems 64] [--res 64] [--level 16]
        [--kind 1] [--channels 32] [--steps 3000] [--out PATH]

Training problems are 64^3 free-slip boxes with eta = 10^(c * GRF), c ~ U(0, 6), correlation
length ~ U(0.1, 0.3) and 1-5 buoyant spheres.  The plain mixed-precision solver runs on each
one with vector capture at the network's level; FP64 level solves give the targets.  The last
10 % of problems are held out and their loss is stored with the model.
"""
import argparse
import os
import sys
import time

import numpy as np

from . import learned


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=64)
    ap.add_argument("--res", type=int, default=64)
    ap.add_argument("--level", type=int, default=16)
    ap.add_argument("--kind", type=int, default=1, choices=(0, 1))
    ap.add_argument("--channels", type=int, default=32)
    ap.add_argument("--depth", type=int, default=4)
    ap.add_argument("--samples", type=int, default=8)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--out", default=learned.DEFAULT_PATH)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    t0 = time.time()
    X, T, M = learned.build_training_set(a.res, (a.level,) * 3, a.kind, range(a.seed0, a.seed0 + a.problems),
                                         samples_per_problem=a.samples, every=4, log=log)
    log(f"training set {X.shape} from {a.problems} problems in {time.time() - t0:.1f}s")
    nval = max(1, int(round(0.1 * a.problems))) * a.samples
    ntr = len(X) - nval
    net = learned.ConvNet((a.level + 1,) * 3, channels=a.channels, depth=a.depth, seed=0)
    schedule = [1e-3, 1e-3, 1e-3, 5e-4, 2e-4, 1e-4]
    t1 = time.time()
    val = float("nan")
    for r, lr in enumerate(schedule):
        net.set_data(X[:ntr], T[:ntr], M[:ntr])
        losses = net.fit(a.steps // len(schedule), batch=a.batch, lr=lr, weight_decay=1e-4, seed=r)
        net.set_data(X[ntr:], T[ntr:], M[ntr:])
        val = net.evaluate(a.batch)
        log(f"stage {r}: lr {lr:.0e} train {losses[-50:].mean():.4f} held-out {val:.4f} ({time.time() - t1:.1f}s)")
    model = learned.LearnedCoarseCorrection(
        net.get_params(), (a.level,) * 3, a.channels, a.depth, mode="add" if a.kind == 1 else "replace",
        meta=dict(problems=a.problems, res=a.res, samples=len(X), held_out=nval, steps=a.steps, batch=a.batch,
                  heldout_loss=f"{val:.4f}", seeds=f"{a.seed0}-{a.seed0 + a.problems - 1}"))
    model.save(a.out)
    log(f"saved {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
