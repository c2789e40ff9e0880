"""Spatio-temporal spectra on the GPU: snapshots of kpsi recorded on the
device (sts_gpe, src/diagnostics.f90:636-695) and the CompSTS spectra
(Run/CompSTS/driver.f90:134-265) against the numpy oracle, for every
supported number of snapshots; and the Bogoliubov dispersion of a weakly
perturbed uniform condensate read off the spectra."""
import math
import sys

import numpy as np
import pytest

import sts_oracle as S
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _vortex_gpe(N, dt=1e-3):
    from quantumvortices_amd import gpe as G
    p = G.params_from_namelist(N, dt, 2.5)
    xs, ys, qs = G.lattice_vortices(4)
    dx = 2 * np.pi / N
    g = G.GPE(p)
    g.set_psi(G.ic_vortices_host(N, xs + 0.31 * dx, ys + 0.17 * dx, qs, p.xi), sync=True)
    return g


def _check_spectra(sp2, sp1, ref2, ref1, nsh):
    # shells past the retained disk are empty in the oracle (kpsi is dealiased)
    assert np.all(ref2[:, nsh:] == 0.0)
    e2 = rel_l2(sp2, ref2[:, :nsh])
    e1 = rel_l2(sp1, ref1)
    m2 = np.max(np.abs(sp2 - ref2[:, :nsh])) / np.max(np.abs(ref2))
    say("  rel-L2 spec2d %.3e (max %.3e)  spec1d %.3e" % (e2, m2, e1))
    assert e2 < 1e-12 and e1 < 1e-12 and m2 < 1e-12


def test_record_and_spectra_vs_oracle():
    """Each snapshot is the handle's kpsi at record time, bit for bit; the
    spectra match the oracle run on the host copies."""
    from quantumvortices_amd import gpe as G
    N, nt = 64, 64
    say("case sts N=%d nt=%d (host snapshots)" % (N, nt))
    g = _vortex_gpe(N)
    sts = G.STS(g, nt)
    host = []
    for n in range(nt):
        g.gpe_rk4(1)
        sts.record(time=(n + 1) * 1e-3)
        host.append(g.kpsi)
    host = np.array(host)
    for n in (0, nt // 2, nt - 1):
        assert np.array_equal(sts.snapshot(n), host[n])
    assert np.array_equal(sts.x0(), host[:, 0, :])
    sp2, sp1 = sts.compute()
    assert sp2.shape == (nt, sts.nshells) and sp1.shape == (nt, N)
    ref2, ref1 = S.sts_spectra(host)
    _check_spectra(sp2, sp1, ref2, ref1, sts.nshells)
    sts.close()
    g.close()


@pytest.mark.parametrize("N,nt", [(64, 32), (128, 128), (64, 256), (32, 512), (32, 1024), (32, 2048),
                                  (64, 4096)])
def test_spectra_all_lengths(N, nt):
    from quantumvortices_amd import gpe as G
    say("case sts N=%d nt=%d" % (N, nt))
    g = _vortex_gpe(N)
    sts = G.STS(g, nt)
    for n in range(nt):
        g.gpe_rk4(1)
        sts.record()
    sp2, sp1 = sts.compute()
    snaps = np.array([sts.snapshot(n) for n in range(nt)])
    ref2, ref1 = S.sts_spectra(snaps)
    _check_spectra(sp2, sp1, ref2, ref1, sts.nshells)
    # reset and record again: the same run gives the same spectra
    sts.reset()
    assert sts.count == 0
    sts.close()
    g.close()


def test_errors():
    from quantumvortices_amd import gpe as G
    say("case sts errors")
    g = _vortex_gpe(32)
    with pytest.raises(Exception):
        G.STS(g, 48)          # not a power of two
    with pytest.raises(Exception):
        G.STS(g, 8192)        # above the largest supported length
    sts = G.STS(g, 32)
    for _ in range(31):
        sts.record()
    with pytest.raises(Exception):
        sts.compute()         # needs all snapshots
    sts.record()
    with pytest.raises(Exception):
        sts.record()          # all taken
    with pytest.raises(Exception):
        sts.snapshot(32)
    sts.compute()
    sts.close()
    g.close()


def test_bogoliubov_dispersion():
    """A uniform condensate (rho = 1, omega_chem = g) with a small random
    perturbation: kpsi(k, t) oscillates at the Bogoliubov frequency
    w_B(k) = sqrt(alpha k^2 (alpha k^2 + 2 g)).  i dpsi/dt = H psi makes the
    particle branch e^{-i w_B t}, which the forward DFT places at negative
    frequency indices: the peak of every shell k must lie at
    -w in [w_B(k - 1/2), w_B(k + 1/2)] (the shell's span), to a bin."""
    from quantumvortices_amd import gpe as G
    N, dt, every, nt = 64, 5e-3, 10, 512
    say("case sts Bogoliubov N=%d nt=%d" % (N, nt))
    p = G.params_from_namelist(N, dt, 2.5)
    rng = np.random.default_rng(11)
    psi = 1.0 + 1e-3 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    g = G.GPE(p)
    g.set_psi(psi, sync=True)
    sts = G.STS(g, nt)
    for n in range(nt):
        g.gpe_rk4(every)
        sts.record((n + 1) * every * dt)
    sp2, _ = sts.compute()
    dw = 2 * math.pi / (nt * every * dt)
    a, gs = p.alpha, p.g

    def wb(k):
        e = a * k * k
        return math.sqrt(e * (e + 2 * gs))

    assert wb(N / 3) < math.pi / (every * dt)      # below the sampling Nyquist
    bad = []
    for k in range(4, 20):
        w = int(np.argmax(sp2[:, k]))
        ws = w if w < nt // 2 else w - nt
        om = -ws * dw
        lo, hi = wb(k - 0.5) - 1.01 * dw, wb(k + 0.5) + 1.01 * dw
        say("  shell %2d: peak at w=%4d  omega %.3f  w_B span [%.3f, %.3f]" % (k, ws, om, wb(k - 0.5), wb(k + 0.5)))
        if not lo <= om <= hi:
            bad.append(k)
    assert not bad, bad
    sts.close()
    g.close()
