"""Kernel variants of one field: the row-pass variants of the psi-row path
(N >= 2048) and their autotuner, and the transform shape of small fields.
Row variants and tuner:
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


@pytest.mark.parametrize("N", [256, 512, 1024])
def test_single_field_shape(G, oracle_mod, N):
    """A single field of N <= 1024 stepped in its second transform shape (8
    points per thread): within roundoff of the default shape, also with
    psi reads and diagnostics between steps (which rebuild the y-inverse in
    the default layout); within 1e-11 of the oracle; the shape can be set
    back; fields with particles have one shape."""
    say("case single-field shape 1, N=%d" % N)
    p = G.params_from_namelist(N, 1e-3, 2.5)
    xs, ys, qs = G.random_vortices(16, 40 + N % 7)
    a = G.GPE(p)
    a.set_ic_vortices(xs, ys, qs)
    psi0 = a.psi
    a.gpe_rk4(6)
    ref = a.psi
    a.close()
    b = G.GPE(p)
    b.set_ic_vortices(xs, ys, qs)
    b.set_shape(1)
    assert b.shape() == 1
    b.gpe_rk4(3)
    _ = b.psi
    _ = b.diagnostics_gpe()
    b.gpe_rk4(3)
    assert rel_l2(b.psi, ref) < 1e-13
    c = G.GPE(p)
    c.set_ic_vortices(xs, ys, qs)
    c.set_shape(1)
    c.gpe_rk4(6)
    assert rel_l2(c.psi, b.psi) < 1e-13
    c.set_shape(0)
    c.gpe_rk4(1)
    b.gpe_rk4(1)
    assert rel_l2(c.psi, b.psi) < 1e-13
    b.close()
    c.close()
    if N == 256:
        O = oracle_mod
        st = O.State(O.make_params(N, 1e-3, 2.5))
        st.set_psi(psi0, sync=True)
        st.rk4(6)
        assert rel_l2(ref, st.psi) < 1e-11
    from quantumvortices_amd import _native as nat
    q = G.GPE(G.params_from_namelist(N, 1e-3, 2.5, gp_N_obj=2, ppotV0=10.0, ppotd=2.0, gp_mass_part=5.0))
    with pytest.raises(nat.VKError):
        q.set_shape(1)
    q.close()
