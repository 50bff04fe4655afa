"""Train the learned Schur stencil on the GPU and save it for ``learned.load_default_schur``.

    python -m underworld2_amd.train_schur_stencil [--problems 512] [--res 64] [--inclusions 0.25]
        [--samples 2] [--channels 16] [--layers 2] [--steps 3000] [--corr-len LO HI] [--out PATH]

Training problems are free-slip unit cubes: a fraction with eta = 10^(c * GRF) (c ~ U(0, 6),
correlation length ~ U(0.1, 0.3)) and 1-5 buoyant spheres, the rest (``--inclusions``) with 1-5
spherical viscosity inclusions of contrast 10^U(-4, 4) (sharp interfaces).  For each problem
the plain solver's operators give the triples (features of approxS, r^, e^) with
r^ = D^-1/2 G^T K^-1 G D^-1/2 e^ for random e^ (``StencilNet.add_samples``); the last 10 % of
problems are held out and their loss is stored with the model.
"""
import argparse
import os
import sys
import time

import numpy as np

from . import learned


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=512)
    ap.add_argument("--res", type=int, default=64)
    ap.add_argument("--inclusions", type=float, default=0.25)
    ap.add_argument("--samples", type=int, default=2)
    ap.add_argument("--channels", type=int, default=16)
    ap.add_argument("--layers", type=int, default=2)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--seed0", type=int, default=5000)
    ap.add_argument("--inner-rtol", type=float, default=1e-6)
    ap.add_argument("--corr-len", type=float, nargs=2, default=None, metavar=("LO", "HI"),
                    help="GRF correlation length range (default 0.1 0.3); longer lengths at a coarse training "
                         "resolution give the element-scale contrast of finer grids")
    ap.add_argument("--out", default=learned.SCHUR_PATH)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    net = learned.StencilNet(a.channels, a.layers, seed=0)
    t0 = time.time()
    n_incl = int(round(a.inclusions * a.problems))
    for k in range(a.problems):
        seed = a.seed0 + k
        # interleave the two families so the held-out tail holds both
        if n_incl and (k % max(1, int(round(a.problems / n_incl)))) == 0:
            st, v, p, desc = learned.inclusion_problem(a.res, seed)
        elif a.corr_len is not None:
            ell = np.random.default_rng(seed + 10**6).uniform(*a.corr_len)
            st, v, p, desc = learned.random_problem(a.res, seed, corr_len=ell)
        else:
            st, v, p, desc = learned.random_problem(a.res, seed)
        net.add_samples(st, a.samples, seed=seed, inner_rtol=a.inner_rtol)
        del st, v, p
        if k % 16 == 0 or k == a.problems - 1:
            log(f"problem {k + 1}/{a.problems} {desc}: {net.sample_count} samples ({time.time() - t0:.1f}s)")
    n = net.sample_count
    nval = max(1, int(round(0.1 * a.problems))) * a.samples
    ntr = n - nval
    base = net.evaluate(ntr, nval)
    log(f"training set {n} samples ({ntr} train, {nval} held out), untrained held-out loss {base:.4f}")
    schedule = [2e-3, 2e-3, 1e-3, 5e-4, 2e-4, 1e-4]
    t1 = time.time()
    val = base
    for r, lr in enumerate(schedule):
        losses = net.fit(a.steps // len(schedule), batch=a.batch, lr=lr, weight_decay=0.0, seed=r, ntrain=ntr)
        val = net.evaluate(ntr, nval)
        log(f"stage {r}: lr {lr:.0e} train {losses[-50:].mean():.4f} held-out {val:.4f} ({time.time() - t1:.1f}s)")
    model = learned.LearnedSchurStencil(
        net.get_params(), a.channels, a.layers,
        meta=dict(problems=a.problems, res=a.res, inclusions=n_incl, samples=n, held_out=nval, steps=a.steps,
                  batch=a.batch, heldout_loss=f"{val:.4f}", untrained_loss=f"{base:.4f}",
                  seeds=f"{a.seed0}-{a.seed0 + a.problems - 1}",
                  corr_len="0.1-0.3" if a.corr_len is None else f"{a.corr_len[0]}-{a.corr_len[1]}"))
    model.save(a.out)
    log(f"saved {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
