"""Row-pass variants of the psi-row path (N >= 2048) and their autotuner:
the tuner restores the state it started from (kpsi and particle state, bit
for bit), keeps the fastest variant pair, and a handle set to that pair by
hand continues exactly as the tuned one; every variant agrees with the
others to roundoff."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available()
    from quantumvortices_amd import gpe
    return gpe


def make(G, N, npart):
    kw = dict(gp_N_obj=npart, ppotV0=10.0, ppotd=2.0, gp_mass_part=5.0) if npart else {}
    g = G.GPE(G.params_from_namelist(N, 2.5e-4, 2.5, **kw))
    xs, ys, qs = G.random_vortices(32, 21)
    g.set_ic_vortices(xs, ys, qs)
    if npart:
        g.initialize_particles(np.random.default_rng(6).uniform(0, 2 * np.pi, (npart, 2)))
    return g


def snapshot(g, npart):
    out = [g.kpsi]
    if npart:
        out += list(g.particles())
    return out


@pytest.mark.parametrize("N,npart", [(2048, 8), (4096, 0)])
def test_autotune_rows(G, N, npart):
    say("case autotune_rows N=%d particles=%d" % (N, npart))
    g = make(G, N, npart)
    g.gpe_rk4(1)
    before = snapshot(g, npart)
    ch, ms = g.autotune_rows(2)
    say("  ms/step [p1][p3]:", ms.tolist(), "chosen", ch)
    after = snapshot(g, npart)
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    assert g.row_variants() == ch
    assert ms[0, 0] > 0 and ms[0, 1] > 0 and ms[1, 0] > 0 and ms[1, 1] > 0
    if N != 4096:
        assert np.all(ms[2] == -1.0)         # the prefetching P1 is 4096^2 only
    assert ms[ch] == ms[ms > 0].min()
    g.gpe_rk4(3)
    tuned = snapshot(g, npart)
    g.close()
    h = make(G, N, npart)
    h.gpe_rk4(1)
    h.set_row_variants(*ch)
    h.gpe_rk4(3)
    for a, b in zip(tuned, snapshot(h, npart)):
        assert np.array_equal(a, b)
    h.close()


def test_row_variants_agree(G):
    """Every variant pair at 2048^2 with particles gives the same 3 steps to
    roundoff (the pairs order sums differently)."""
    say("case row variants agree, N=2048 with 8 particles")
    res = []
    for p1 in (0, 1):
        for p3 in (0, 1):
            g = make(G, 2048, 8)
            g.set_row_variants(p1, p3)
            g.gpe_rk4(3)
            res.append(snapshot(g, 8))
            g.close()
    for r in res[1:]:
        assert rel_l2(r[0], res[0][0]) < 1e-13
        assert rel_l2(r[1], res[0][1]) < 1e-13
        assert rel_l2(r[2], res[0][2]) < 1e-10    # velocities ~1e-13 after 3 steps from rest
    g = make(G, 2048, 8)
    from quantumvortices_amd import _native as nat
    with pytest.raises(nat.VKError):
        g.set_row_variants(2, 0)             # no prefetching P1 at 2048^2
    g.close()
