"""What the trained warm-start network gains: on a held-out 512^2
configuration with obstacles, ARGLE from the network's prediction reaches residual 1e-10
and the same relaxed state (grand potential) as ARGLE from the Thomas-Fermi
profile, in no more steps, while plain ARGLE from the Pade product ansatz is
still far from the tolerance after many times the warm start's wall time."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "quantumvortices_amd", "data", "fno_argle_512.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_warm_start_gain():
    import torch
    from quantumvortices_amd import gpe as G
    from quantumvortices_amd import warmstart as W
    assert torch.cuda.is_available()
    fno = W.FNO.load(WEIGHTS, batch=1)
    assert fno.residual and fno.info["samples"] >= 20000
    g = G.GPE(W.solver_params(512, W.DT_MAX))
    psi = torch.empty((512, 512), dtype=torch.complex128, device="cuda")
    ref = torch.empty_like(psi)
    # outside the training seeds; with obstacles (without them the
    # Thomas-Fermi profile psi = 1 is already the exact ground state)
    seed = next(s for s in range(900007, 900100) if len(W.random_config(s)["obs"]) >= 2)
    cfg = W.random_config(seed)
    say("case warm start: nv=%d nobs=%d V0=%.2f" % (len(cfg["xs"]), len(cfg["obs"]), cfg["V0"]))
    W.warm_start(g, fno, cfg, psi)                       # untimed first pass
    nw, rw, tw = W.warm_start(g, fno, cfg, psi)
    om_w = W.grand_potential(g.diagnostics_gpe())
    g.psi_to(ref)
    fno.predict_psi(psi)
    torch.cuda.synchronize()
    raw = float(torch.linalg.norm(psi - ref) / torch.linalg.norm(ref))
    torch.cuda.synchronize()
    nt, rt, tt = W.tf_start(g, fno, cfg, psi)
    om_t = W.grand_potential(g.diagnostics_gpe())
    fno.tf_guess(g.params, psi)
    torch.cuda.synchronize()
    raw_tf = float(torch.linalg.norm(psi - ref) / torch.linalg.norm(ref))
    torch.cuda.synchronize()
    cap = 20000
    npl, rpl, tpl = W.plain_start(g, cfg, max_steps=cap, check_every=2000)
    say("  warm %d steps %.1f ms res %.2e | tf %d steps %.1f ms | plain %d steps %.0f ms res %.2e" %
        (nw, 1e3 * tw, rw, nt, 1e3 * tt, npl, 1e3 * tpl, rpl))
    say("  raw error: network %.2e  Thomas-Fermi %.2e;  Omega warm %.12e tf %.12e" % (raw, raw_tf, om_w, om_t))
    assert rw <= 1e-10 and rt <= 1e-10
    assert abs(om_w - om_t) <= 1e-10 * abs(om_t)
    assert raw < 2e-3 and raw < 0.1 * raw_tf
    assert nw <= nt
    assert rpl > 1e-10 and npl == cap           # plain ARGLE is still relaxing ...
    assert tpl > 3.0 * tw                        # ... after more than 3x the warm start's wall time
    fno.close()
    g.close()
