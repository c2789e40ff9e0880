"""Batched ensemble of unequal realizations: every schedule (one launch per
pass for all sizes, per size, per realization, or one persistent launch per
step; graph replay or direct) gives each member exactly its single-field
result, and members match the oracle."""
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def members(G):
    sizes = [128, 256, 128, 512, 256, 1024, 128]
    ps = [G.params_from_namelist(N, 1e-3, 2.5) for N in sizes]
    ics = []
    for r, N in enumerate(sizes):
        xs, ys, qs = G.random_vortices(8 + 2 * r, seed=100 + r)
        ics.append((xs, ys, qs))
    return sizes, ps, ics


def single_runs(G, ps, ics, steps):
    out = []
    for p, (xs, ys, qs) in zip(ps, ics):
        g = G.GPE(p)
        g.set_ic_vortices(xs, ys, qs)
        g.gpe_rk4(steps)
        out.append(g.psi)
        g.close()
    return out


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available()
    from quantumvortices_amd import gpe
    return gpe


@pytest.mark.parametrize("mode,graph", [(0, True), (0, False), (1, True), (2, True), (2, False), (3, True), (3, False)])
def test_ensemble_equals_single(G, mode, graph):
    say("case ensemble mode=%d graph=%s" % (mode, graph))
    sizes, ps, ics = members(G)
    ref = single_runs(G, ps, ics, 5)
    e = G.Ensemble(ps)
    for r, (xs, ys, qs) in enumerate(ics):
        e.set_ic_vortices(r, xs, ys, qs)
    e.set_mode(mode, graph)
    e.rk4(2)
    _ = e.psi(0)            # eager work between graph replays
    e.rk4(3)
    for r in range(len(sizes)):
        assert np.array_equal(e.psi(r), ref[r]), r
    e.close()


def test_ensemble_diagnostics_all(G):
    """All members' diagnostics in one call (concurrent pipelines, one wait)
    equal the per-member calls bit for bit, shape-1 members included.
    Stepping after them continues bitwise as without them for shape-0
    members; a shape-1 member's y-inverse is rebuilt in its own layout
    after a single-field access (a different kernel), so it continues to
    roundoff."""
    say("case ensemble diagnostics_all")
    sizes, ps, ics = members(G)
    e = G.Ensemble(ps)
    for r, (xs, ys, qs) in enumerate(ics):
        e.set_ic_vortices(r, xs, ys, qs)
    before = e.diagnostics_all()           # before any step: the initial fields
    for r in range(len(sizes)):
        scal, spec = e.diagnostics(r)
        assert np.array_equal(before[r][0], scal) and np.array_equal(before[r][1], spec), r
    e.set_shape(512, 1)
    e.rk4(2)
    allr = e.diagnostics_all()
    for r in range(len(sizes)):
        scal, spec = e.diagnostics(r)
        assert np.array_equal(allr[r][0], scal), r
        assert np.array_equal(allr[r][1], spec), r
    e.rk4(1)
    a = [e.psi(r) for r in range(len(sizes))]
    e.close()
    e = G.Ensemble(ps)
    for r, (xs, ys, qs) in enumerate(ics):
        e.set_ic_vortices(r, xs, ys, qs)
    e.set_shape(512, 1)
    e.rk4(3)
    for r, N in enumerate(sizes):
        if N == 512:
            assert rel_l2(e.psi(r), a[r]) < 1e-13, r
        else:
            assert np.array_equal(e.psi(r), a[r]), r
    e.close()


def test_ensemble_autotune_restores_state(G):
    say("case ensemble autotune")
    sizes, ps, ics = members(G)
    ref = single_runs(G, ps, ics, 4)
    e = G.Ensemble(ps)
    for r, (xs, ys, qs) in enumerate(ics):
        e.set_ic_vortices(r, xs, ys, qs)
    ms = e.autotune(2)
    say("  ms/step modes x {graph, direct}:", ms.tolist(), "chosen", e.mode())
    assert np.all(ms > 0)
    e.rk4(4)
    for r in range(len(sizes)):
        assert np.array_equal(e.psi(r), ref[r]), r
    e.close()


def test_ensemble_vs_oracle(G, oracle_mod):
    say("case ensemble vs oracle")
    O = oracle_mod
    sizes = [128, 256]
    ps = [G.params_from_namelist(N, 1e-3, 2.5) for N in sizes]
    e = G.Ensemble(ps)
    psi0 = []
    for r, N in enumerate(sizes):
        xs, ys, qs = G.random_vortices(10, seed=7 + r)
        a = G.ic_vortices_host(N, xs, ys, qs, ps[r].xi)
        psi0.append(a)
        e.set_psi(r, a)
    e.rk4(10)
    for r, N in enumerate(sizes):
        st = O.State(O.make_params(N, 1e-3, 2.5))
        st.set_psi(psi0[r], sync=True)
        st.rk4(10)
        err = rel_l2(e.psi(r), st.psi)
        say("  N=%d rel-L2 %.3e" % (N, err))
        assert err < 1e-11
        scal, spec = e.diagnostics(r)
        s2, sp2, _ = st.diagnostics()
        assert abs(scal[4] - s2[4]) < 1e-11 * s2[4]
        for i in range(6):
            assert rel_l2(spec[i], sp2[i]) < 1e-11
    e.close()


def test_ensemble_384_mix(G):
    """The 384-member workload itself (128 x 128^2, 128 x 256^2, 128 x 512^2,
    8-64 zero-net-charge vortices, seed = member id), autotuned schedule:
    after 5 RK4 steps, sampled members of every size equal their single-field
    runs bit for bit, and every member's norm is conserved to the scheme's
    accuracy."""
    say("case ensemble 384 members")
    sizes = [128] * 128 + [256] * 128 + [512] * 128
    ps = [G.params_from_namelist(N, 1e-3, 2.5) for N in sizes]
    ics = []
    for r in range(len(sizes)):
        rng = np.random.default_rng(r)
        nv = 2 * int(rng.integers(4, 33))
        ics.append(G.random_vortices(nv, seed=r))
    e = G.Ensemble(ps)
    for r, (xs, ys, qs) in enumerate(ics):
        e.set_ic_vortices(r, xs, ys, qs)
    n0 = {r: e.diagnostics(r)[0][4] for r in (0, 200, 383)}
    e.autotune(2)
    say("  schedule", e.mode())
    e.rk4(5)
    sample = [0, 1, 77, 127, 128, 129, 200, 255, 256, 300, 382, 383]
    for r in sample:
        g = G.GPE(ps[r])
        xs, ys, qs = ics[r]
        g.set_ic_vortices(xs, ys, qs)
        g.gpe_rk4(5)
        assert np.array_equal(e.psi(r), g.psi), r
        g.close()
    for r, nn in n0.items():
        n1 = e.diagnostics(r)[0][4]
        assert abs(n1 - nn) < 1e-11 * nn, r
    e.close()


def test_ensemble_384_persistent(G):
    """The 384-member mix stepped by the persistent launch (all resident
    workgroups, grid barriers between the 16 passes), graph-replayed: sampled
    members of every size equal their single-field runs bit for bit."""
    say("case ensemble 384 members, persistent step")
    sizes = [128] * 128 + [256] * 128 + [512] * 128
    ps = [G.params_from_namelist(N, 1e-3, 2.5) for N in sizes]
    ics = [G.random_vortices(2 * int(np.random.default_rng(r).integers(4, 33)), seed=r) for r in range(len(sizes))]
    e = G.Ensemble(ps)
    for r, (xs, ys, qs) in enumerate(ics):
        e.set_ic_vortices(r, xs, ys, qs)
    e.set_mode(3, True)
    e.rk4(3)
    for r in (0, 100, 127, 128, 200, 255, 256, 301, 383):
        g = G.GPE(ps[r])
        xs, ys, qs = ics[r]
        g.set_ic_vortices(xs, ys, qs)
        g.gpe_rk4(3)
        assert np.array_equal(e.psi(r), g.psi), r
        g.close()
    e.close()


def test_ensemble_shapes(G, oracle_mod):
    """Members of sizes 256..1024 in their second shape (8 points per
    thread, radix-8 stages, half-width panels): within roundoff of the
    single-field runs and within 1e-11 of the oracle; members left in
    shape 0 stay bit-identical; reading a member between steps, the
    per-member launch mode and the persistent step keep the state
    consistent; the shape autotuner
    restores the state and reports both timings."""
    say("case ensemble shapes")
    sizes, ps, ics = members(G)
    ref = single_runs(G, ps, ics, 5)
    e = G.Ensemble(ps)
    for r, (xs, ys, qs) in enumerate(ics):
        e.set_ic_vortices(r, xs, ys, qs)
    for N in (256, 512, 1024):
        e.set_shape(N, 1)
        assert e.shape(N) == 1
    e.set_mode(0, True)
    e.rk4(2)
    _ = e.psi(3)              # a shape-1 member read between steps
    e.set_mode(2, False)
    e.rk4(1)                  # per-member launches use the single-field layout
    e.set_mode(3, True)
    e.rk4(1)                  # the persistent step over shape-1 members
    e.set_mode(0, True)
    e.rk4(1)
    for r, N in enumerate(sizes):
        a = e.psi(r)
        if N == 128:
            assert np.array_equal(a, ref[r]), r
        else:
            err = rel_l2(a, ref[r])
            say("  member %d N=%d shape 1: rel-L2 vs single %.2e" % (r, N, err))
            assert err < 1e-13, r
    O = oracle_mod
    r = 5                     # N = 1024
    st = O.State(O.make_params(sizes[r], 1e-3, 2.5))
    xs, ys, qs = ics[r]
    g = G.GPE(ps[r])
    g.set_ic_vortices(xs, ys, qs)
    st.set_psi(g.psi, sync=True)
    g.close()
    st.rk4(5)
    assert rel_l2(e.psi(r), st.psi) < 1e-11
    e.close()
    # autotuner: state restored, timings reported
    e = G.Ensemble(ps)
    for r, (xs, ys, qs) in enumerate(ics):
        e.set_ic_vortices(r, xs, ys, qs)
    res = e.autotune_shapes(2)
    say("  autotune_shapes", res)
    assert res[128][0] == 0 and res[128][1] == -1.0
    for N in (256, 512, 1024):
        assert res[N][1] > 0 and res[N][2] > 0 and e.shape(N) == res[N][0]
    e.rk4(5)
    for r, N in enumerate(sizes):
        a = e.psi(r)
        if e.shape(N) == 0:
            assert np.array_equal(a, ref[r]), r
        else:
            assert rel_l2(a, ref[r]) < 1e-13, r
    e.close()
    # the tuned shapes are remembered for later ensembles until forgotten
    e = G.Ensemble(ps)
    assert all(e.shape(N) == res[N][0] for N in (256, 512, 1024))
    e.close()
    from quantumvortices_amd import _native as nat
    nat.lib().vk_ens_forget_shapes()
    e = G.Ensemble(ps)
    assert all(e.shape(N) == 0 for N in (256, 512, 1024))
    e.close()
