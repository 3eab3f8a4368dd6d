"""Fitting driver: recovers per-agent secretion rates, D and lambda from the
substrate fields after `steps` steps, by gradient descent through the adjoint.

    python -m biofvm_amd.fit [--n 128] [--steps 500] [--agents 20000] [--iters 20]

The target comes from the case's hidden parameters (``workloads.differentiation_128``);
the fit starts from perturbed values and minimises
J = sum (rho_N - target)^2 / 2 over log D, log lambda and log secretion with
Adam.  Every forward and adjoint step runs in the library's HIP kernels.
"""
import argparse
import json
import time

import numpy as np
import torch

from ._capi import FP32, FP64
from .autograd import LodSteps, Model
from . import workloads as W


def make_problem(n=128, steps=500, n_agents=20_000, seed=6, precision=FP32, checkpoint_every=50):
    case = W.differentiation_128(steps=steps, seed=seed, n=n, n_agents=n_agents)
    model = Model(case, precision, checkpoint_every)
    rho0, D, lam, sec, upt = model.params()
    with torch.no_grad():
        target = LodSteps.apply(rho0, D, lam, sec, upt, model).clone()
    return model, target


def fit(model, target, iters=20, lr=0.05, perturb=0.3, seed=0, log=print):
    """Adam on (log D, log lambda, log secretion) from parameters perturbed by
    a factor exp(U[-perturb, perturb]); returns the loss history and the
    relative errors of the recovered parameters."""
    rho0, D, lam, sec, upt = model.params()
    rng = np.random.default_rng(seed)
    def jitter(x):
        return x * torch.from_numpy(np.exp(rng.uniform(-perturb, perturb, tuple(x.shape))))
    logD = torch.log(jitter(D)).requires_grad_()
    logL = torch.log(jitter(lam)).requires_grad_()
    logS = torch.log(jitter(sec).clamp_min(1e-6)).requires_grad_()
    opt = torch.optim.Adam([logD, logL, logS], lr=lr)
    hist = []
    for it in range(iters):
        opt.zero_grad()
        out = LodSteps.apply(rho0, logD.exp(), logL.exp(), logS.exp(), upt, model)
        loss = 0.5 * ((out - target) ** 2).sum()
        loss.backward()
        opt.step()
        hist.append(float(loss))
        if log:
            log(f"iter {it:3d}  loss {float(loss):.6e}  "
                f"D rel err {float(((logD.exp() - D).abs() / D).max()):.3e}  "
                f"lambda rel err {float(((logL.exp() - lam).abs() / lam).max()):.3e}")
    errs = {"D": float(((logD.exp() - D).abs() / D).max()),
            "lambda": float(((logL.exp() - lam).abs() / lam).max()),
            "secretion_median": float(((logS.exp() - sec).abs() / sec.clamp_min(1e-6)).median())}
    return hist, errs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--agents", type=int, default=20_000)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--fp64", action="store_true")
    args = ap.parse_args()
    t0 = time.perf_counter()
    model, target = make_problem(args.n, args.steps, args.agents, precision=FP64 if args.fp64 else FP32)
    hist, errs = fit(model, target, iters=args.iters)
    print(json.dumps({"loss_first": hist[0], "loss_last": hist[-1], "errors": errs,
                      "seconds": time.perf_counter() - t0}))
    model.close()


if __name__ == "__main__":
    main()
