"""Train the ARGLE warm-start network on ground states the solver generates.

    python -m quantumvortices_amd.train_warmstart --samples 20000 --out fno.npz

1. Data: for each seeded configuration (16-128 vortices with zero net charge,
   0-8 Gaussian obstacles with V0 in [2, 20] g and width 2 xi, on N^2), the
   input planes come from vk_fno_features and the target is ARGLE relaxed to
   the residual tolerance from the Thomas-Fermi profile (warmstart.py states
   why that is plain ARGLE's fixed point).  Inputs and FP32 targets stay in
   device memory.
2. Training: mean-squared error on (Re psi - TF, Im psi), Adam with a cosine
   learning-rate schedule, the library's forward/backward kernels.  The
   number of optimizer steps is capped by a wall-time budget measured on
   the first steps.
3. Validation on held-out seeds: relative L2 error of the prediction and of
   the Thomas-Fermi profile against the relaxed state.
"""
import argparse
import json
import math
import sys
import time

import numpy as np


def log(*a):
    print(*a, flush=True)


def generate(W, G, torch, N, seeds, dt, tol, X, T, every=1000):
    g = G.GPE(W.solver_params(N, dt))
    f = W.FNO(N, batch=1)
    psi = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    t0 = time.perf_counter()
    steps = []
    for i, s in enumerate(seeds):
        cfg = W.random_config(int(s))
        n, r = W.ground_state(g, f, cfg, psi, tol=tol)
        if not r <= tol:
            raise RuntimeError("ground state %d (seed %d) stopped at residual %.3e" % (i, s, r))
        f.features(cfg, g.params, X[i])
        g.psi_to(psi)
        T[i, 0].copy_(psi.real)
        T[i, 0].sub_(X[i, 2])            # target: correction to the Thomas-Fermi plane
        T[i, 1].copy_(psi.imag)
        torch.cuda.synchronize()          # psi is rewritten by the library's own streams next
        steps.append(n)
        if (i + 1) % every == 0:
            log("  data %d/%d  %.1f s  mean ARGLE steps %.0f" % (i + 1, len(seeds), time.perf_counter() - t0,
                                                               np.mean(steps[-every:])))
    f.close()
    g.close()
    return float(np.mean(steps)), time.perf_counter() - t0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--val", type=int, default=64)
    ap.add_argument("--epochs", type=float, default=4.0)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-min", type=float, default=1e-5)
    ap.add_argument("--train-seconds", type=float, default=480.0)
    ap.add_argument("--dt", type=float, default=0.08)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    import torch
    from . import gpe as G
    from . import warmstart as W
    if not torch.cuda.is_available():
        sys.exit("train_warmstart needs a ROCm GPU")
    N, S, V = a.N, a.samples, a.val
    dev = "cuda"
    X = torch.empty((S + V, 3, N, N), dtype=torch.float32, device=dev)
    T = torch.empty((S + V, 2, N, N), dtype=torch.float32, device=dev)
    seeds = np.arange(S + V) + 100000 * (a.seed + 1)
    log("generating %d + %d ARGLE ground states at %d^2 (dt %g, tol %g)" % (S, V, N, a.dt, a.tol))
    mean_steps, gen_s = generate(W, G, torch, N, seeds, a.dt, a.tol, X, T)
    log("data: %.1f s, %.2f ms/sample, mean %.0f ARGLE steps" % (gen_s, 1e3 * gen_s / (S + V), mean_steps))

    B = a.batch
    f = W.FNO(N, batch=B)
    f.init_params(a.seed)
    per_epoch = S // B
    planned = int(a.epochs * per_epoch)
    gen = torch.Generator(device=dev)
    gen.manual_seed(a.seed)
    perm = torch.randperm(S, device=dev, generator=gen)
    xb = torch.empty((B, 3, N, N), dtype=torch.float32, device=dev)
    tb = torch.empty((B, 2, N, N), dtype=torch.float32, device=dev)
    t0 = time.perf_counter()
    losses = []
    step = 0
    total = planned
    while step < total:
        k = step % per_epoch
        if k == 0 and step > 0:
            perm = torch.randperm(S, device=dev, generator=gen)
        idx = perm[k * B:(k + 1) * B]
        torch.index_select(X, 0, idx, out=xb)
        torch.index_select(T, 0, idx, out=tb)
        torch.cuda.synchronize()
        lr = a.lr_min + 0.5 * (a.lr - a.lr_min) * (1 + math.cos(math.pi * step / total))
        losses.append(f.train_step(xb, tb, lr))
        step += 1
        if step == 20:
            per_step = (time.perf_counter() - t0) / 20
            total = max(20, min(planned, int(a.train_seconds / per_step)))
            log("train: %.2f ms/step (batch %d), %d steps planned (%.2f epochs)" %
                (1e3 * per_step, B, total, total / per_epoch))
        if step % 500 == 0:
            log("  step %d/%d  loss %.3e  lr %.2e  %.0f s" % (step, total, np.mean(losses[-500:]), lr,
                                                             time.perf_counter() - t0))
    train_s = time.perf_counter() - t0

    # validation: relative L2 error of prediction and of the TF profile
    # against the relaxed state psi = (T0 + TF, T1)
    errs, tf_errs = [], []
    for i in range(S, S + V):
        y = torch.from_numpy(f.forward(X[i:i + 1].cpu().numpy())).to(dev)[0]
        t = T[i].clone()
        tf = X[i, 2]
        t[0] += tf
        y[0] += tf
        errs.append(float(torch.linalg.norm(y - t) / torch.linalg.norm(t)))
        tf_errs.append(float(torch.linalg.norm(tf - t[0]) / torch.linalg.norm(t)))
    info = dict(samples=S, val=V, steps=step, batch=B, epochs=step / per_epoch, train_seconds=round(train_s, 1),
                gen_seconds=round(gen_s, 1), mean_argle_steps=mean_steps, dt=a.dt, tol=a.tol,
                final_loss=float(np.mean(losses[-200:])), val_rel_err=float(np.mean(errs)),
                val_tf_rel_err=float(np.mean(tf_errs)), seed_base=int(seeds[0]), residual=True)
    log(json.dumps(info))
    f.residual = True
    f.save(a.out, meta=info)
    log("saved %s" % a.out)
    return f, info


if __name__ == "__main__":
    main()
