"""GPU parity: the HIP library (through the C ABI) against the CPU oracle and
numpy on the same seeded inputs.  Tolerance: FP64 fields to relative L2
<= 1e-11 (the bar the reference's FP64 path is held to)."""
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-11


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available()
    from quantumvortices_amd import gpe
    return gpe


def rand_field(N, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))


@pytest.mark.parametrize("N", [32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_fft2d_vs_numpy(G, N):
    say("case fft2d N=%d" % N)
    p = G.params_from_namelist(N, 1e-3, 2.5)
    g = G.GPE(p)
    a = rand_field(N, N)
    k = g.fft(a, -1)
    ref = np.fft.fft2(a).T          # k-space [kx, ky]
    assert rel_l2(k, ref) < 1e-13
    back = g.fft(k, +1)
    assert rel_l2(back, a) < 1e-13
    g.close()


@pytest.mark.parametrize("N", [32, 128, 1024])
def test_set_get_psi(G, N):
    say("case set/get N=%d" % N)
    p = G.params_from_namelist(N, 1e-3, 2.5)
    g = G.GPE(p)
    a = rand_field(N, 7)
    g.set_psi(a, sync=False)
    assert rel_l2(g.psi, a) < 1e-13
    assert rel_l2(g.kpsi, np.fft.fft2(a).T) < 1e-13
    g.close()


def _ic(G, N, nv_lattice=4, xi_in=2.5):
    # lattice shifted off the grid points: a core exactly on a grid point
    # makes |psi| roundoff-sized there and w = Im(psi* grad psi)/|psi|
    # ill-conditioned for any implementation
    p = G.params_from_namelist(N, 1e-3, xi_in)
    xs, ys, qs = G.lattice_vortices(nv_lattice)
    dx = 2 * np.pi / N
    return G.ic_vortices_host(N, xs + 0.31 * dx, ys + 0.17 * dx, qs, p.xi)


@pytest.mark.parametrize("N,steps", [(32, 20), (64, 20), (128, 100), (256, 10), (512, 4)])
def test_rk4_vs_oracle(G, oracle_mod, N, steps):
    say("case rk4 N=%d steps=%d" % (N, steps))
    O = oracle_mod
    dt = 1e-3
    psi0 = _ic(G, N)
    g = G.GPE(G.params_from_namelist(N, dt, 2.5))
    g.set_psi(psi0, sync=True)
    g.gpe_rk4(steps)
    st = O.State(O.make_params(N, dt, 2.5))
    st.set_psi(psi0, sync=True)
    st.rk4(steps)
    e1, e2 = rel_l2(g.psi, st.psi), rel_l2(g.kpsi, st.kpsi)
    say("  rel-L2 psi %.3e kpsi %.3e" % (e1, e2))
    assert e1 < TOL and e2 < TOL
    g.close()


@pytest.mark.parametrize("N,steps", [(64, 10), (128, 20)])
def test_argle_vs_oracle(G, oracle_mod, N, steps):
    say("case argle N=%d steps=%d" % (N, steps))
    O = oracle_mod
    dt = 1e-2
    psi0 = _ic(G, N)
    g = G.GPE(G.params_from_namelist(N, dt, 2.5))
    g.set_psi(psi0, sync=False)
    r_gpu = g.ARGLE(nsteps=steps)
    st = O.State(O.make_params(N, dt, 2.5))
    st.set_psi(psi0, sync=False)
    r_orc = st.argle(steps)
    e = rel_l2(g.kpsi, st.kpsi)
    say("  rel-L2 kpsi %.3e residual gpu %.6e oracle %.6e" % (e, r_gpu, r_orc))
    assert e < TOL
    assert abs(r_gpu - r_orc) <= 1e-9 * abs(r_orc)
    g.close()


@pytest.mark.parametrize("case", ["rk4", "rk4_particles", "argle_obstacles"])
def test_large_n_vs_oracle(G, oracle_mod, case):
    """N = 2048: the single-field path keeps x-space psi rows between the
    row passes (P3 reads them instead of transforming Hpsi again)."""
    N = 2048
    say("case large-N %s N=%d" % (case, N))
    O = oracle_mod
    psi0 = _ic(G, N, nv_lattice=4)
    # next to vortex cores, where the density gradient gives O(1) forces
    xs, ys, _ = G.lattice_vortices(4)
    pos = np.stack([xs[[0, 5, 10, 15]] + 0.04, ys[[0, 5, 10, 15]] - 0.03], axis=1)
    if case == "rk4":
        dt, steps, kw, okw = 1e-3, 2, {}, {}
    elif case == "rk4_particles":
        dt, steps = 1e-3, 2
        kw = dict(gp_N_obj=4, ppotV0=10.0, ppotd=2.0, gp_mass_part=5.0)
        okw = dict(np_=4, V0_in=10.0, d_in=2.0, mass=5.0)
    else:
        dt, steps = 2e-2, 3
        kw = dict(gp_N_obj=2, ppotV0=10.0, ppotd=2.0, gp_mass_part=1.0)
        okw = dict(np_=2, V0_in=10.0, d_in=2.0, mass=1.0)
        pos = pos[:2]
    g = G.GPE(G.params_from_namelist(N, dt, 2.5, **kw))
    st = O.State(O.make_params(N, dt, 2.5, **okw))
    sync = case != "argle_obstacles"
    g.set_psi(psi0, sync=sync)
    st.set_psi(psi0, sync=sync)
    if kw:
        g.initialize_particles(pos)
        st.part(0)[...] = pos
        st.particle_potential()
    if case == "argle_obstacles":
        r, ro = g.ARGLE(nsteps=steps), st.argle(steps)
        assert abs(r - ro) <= 1e-9 * abs(ro)
    else:
        g.gpe_rk4(steps)
        st.rk4(steps)
    e = rel_l2(g.kpsi, st.kpsi)
    say("  rel-L2 kpsi %.3e" % e)
    assert e < TOL
    if case == "rk4_particles":
        gp, gv, _ = g.particles()
        say("  |vel| %.3e rel-L2 pos %.3e vel %.3e" % (np.abs(gv).max(), rel_l2(gp, st.part(0)),
                                                      rel_l2(gv, st.part(1))))
        assert np.abs(gv).max() > 1e-12           # the particles do move
        assert rel_l2(gp, st.part(0)) < TOL and rel_l2(gv, st.part(1)) < 1e-10
    g.close()


@pytest.mark.parametrize("N", [64, 128])
def test_diagnostics_vs_oracle(G, oracle_mod, N):
    say("case diagnostics N=%d" % N)
    O = oracle_mod
    dt = 1e-3
    psi0 = _ic(G, N)
    g = G.GPE(G.params_from_namelist(N, dt, 2.5))
    g.set_psi(psi0, sync=True)
    g.gpe_rk4(3)
    d = g.diagnostics_gpe()
    st = O.State(O.make_params(N, dt, 2.5))
    st.set_psi(psi0, sync=True)
    st.rk4(3)
    scal, spec, _ = st.diagnostics()
    say("  scal gpu", d["scal"])
    say("  scal orc", scal)
    # energies and norm: relative; momentum is zero by the lattice's
    # symmetry (roundoff-sized), so it is held to an absolute bound on the
    # energy scale
    for i in (0, 1, 2, 3, 4):
        assert abs(d["scal"][i] - scal[i]) <= 1e-11 * abs(scal[i])
    escale = abs(scal[0]) + abs(scal[1])
    for i in (5, 6):
        assert abs(d["scal"][i] - scal[i]) <= 1e-12 * escale
    for i in range(6):
        e = rel_l2(d["spec"][i], spec[i])
        say("  spectrum %d rel-L2 %.3e" % (i + 1, e))
        assert e < 1e-11
    g.close()


@pytest.mark.parametrize("N,tg,vx,vy,nobs,steps", [(64, 1, 0.0, 0.0, 0, 10), (128, 0, 0.7, -0.4, 0, 10),
                                                     (128, 1, 0.3, 0.2, 0, 10), (64, 1, 0.5, 0.0, 2, 8),
                                                     (2048, 1, 0.3, -0.2, 0, 3)])
def test_advective_argle_vs_oracle(G, oracle_mod, N, tg, vx, vy, nobs, steps):
    """ARGLE with the Taylor-Green advection and/or counterflow terms
    (GPE_ADVECTION / GPE_COUNTERFLOW builds of src/gpe.f90:5-224)."""
    say("case advective argle N=%d tg=%d v=(%g,%g) obstacles=%d" % (N, tg, vx, vy, nobs))
    O = oracle_mod
    dt = 1e-2
    psi0 = _ic(G, N)
    kw, okw = {}, {}
    if nobs:
        kw = dict(gp_N_obj=nobs, ppotV0=10.0, ppotd=2.0, gp_mass_part=1.0)
        okw = dict(np_=nobs, V0_in=10.0, d_in=2.0, mass=1.0)
    g = G.GPE(G.params_from_namelist(N, dt, 2.5, **kw))
    g.set_advection(bool(tg), vx, vy)
    g.set_psi(psi0, sync=False)
    st = O.State(O.make_params(N, dt, 2.5, adv=tg, vnx=vx, vny=vy, **okw))
    st.set_psi(psi0, sync=False)
    if nobs:
        pos = np.array([[1.1, 2.3], [4.4, 0.7]])[:nobs]
        g.initialize_particles(pos)
        st.part(0)[...] = pos
        st.particle_potential()
    r, ro = g.ARGLE(nsteps=steps), st.argle(steps)
    e = rel_l2(g.kpsi, st.kpsi)
    say("  rel-L2 kpsi %.3e residual %.6e / %.6e" % (e, r, ro))
    assert e < TOL and abs(r - ro) <= 1e-9 * abs(ro)
    # the advection changes the result: not the plain update
    st0 = O.State(O.make_params(N, dt, 2.5, **okw))
    st0.set_psi(psi0, sync=False)
    if nobs:
        st0.part(0)[...] = pos
        st0.particle_potential()
    st0.argle(steps)
    assert rel_l2(st0.kpsi, st.kpsi) > 1e-6
    g.close()


def test_reference_cpu_case_128(G, oracle_mod):
    """The reference's own CPU-runnable configuration end to end: 128^2,
    g = 1, xi = 2.5 dx, 4x4 lattice of alternating vortices (Pade cores),
    RK4 dt = 1e-3 for 1000 steps with the diagnostics every 100 steps.  At
    every checkpoint the field, the energies and all six spectra match the
    oracle, and norm and energy drift match it too."""
    N, dt = 128, 1e-3
    say("case reference 128^2 x 1000 steps")
    O = oracle_mod
    psi0 = _ic(G, N)
    g = G.GPE(G.params_from_namelist(N, dt, 2.5))
    g.set_psi(psi0, sync=True)
    st = O.State(O.make_params(N, dt, 2.5))
    st.set_psi(psi0, sync=True)
    first = None
    for blk in range(10):
        g.gpe_rk4(100)
        st.rk4(100)
        d = g.diagnostics_gpe()
        scal, spec, _ = st.diagnostics()
        e_psi = rel_l2(g.kpsi, st.kpsi)
        e_spec = max(rel_l2(d["spec"][i], spec[i]) for i in range(6))
        say("  step %4d rel-L2 kpsi %.2e  worst spectrum %.2e  etot %.12e  norm %.12e"
            % (100 * (blk + 1), e_psi, e_spec, d["etot"], d["norm"]))
        assert e_psi < TOL and e_spec < TOL
        for i in (0, 1, 2, 3, 4):
            assert abs(d["scal"][i] - scal[i]) <= 1e-11 * abs(scal[i])
        if first is None:
            first = (d["norm"], d["etot"], scal[4], scal[3])
    # drift over the run is the scheme's, identical on both sides
    dn_g, dn_o = d["norm"] - first[0], scal[4] - first[2]
    de_g, de_o = d["etot"] - first[1], scal[3] - first[3]
    say("  drift after 900 steps: norm %.3e / %.3e  etot %.3e / %.3e" % (dn_g, dn_o, de_g, de_o))
    assert abs(dn_g - dn_o) <= 1e-12 * first[0] and abs(de_g - de_o) <= 1e-12 * abs(first[1])
    g.close()


def test_config_1024_slice(G, oracle_mod):
    """The 1024^2 configuration's pipeline on a slice the oracle finishes in
    seconds: 64 seeded random vortices with zero net charge (Pade ansatz),
    ARGLE, then RK4 dt = 5e-4, then the energies and all six spectra."""
    N = 1024
    say("case 1024^2: ARGLE 10 steps, RK4 4 steps, diagnostics")
    O = oracle_mod
    xs, ys, qs = G.random_vortices(64, seed=7)
    p = G.params_from_namelist(N, 1e-2, 2.5)
    psi0 = G.ic_vortices_host(N, xs, ys, qs, p.xi, nimg=3)
    g = G.GPE(p)
    g.set_psi(psi0, sync=False)
    st = O.State(O.make_params(N, 1e-2, 2.5))
    st.set_psi(psi0, sync=False)
    r, ro = g.ARGLE(nsteps=10), st.argle(10)
    assert abs(r - ro) <= 1e-9 * abs(ro)
    kpsi = g.kpsi
    assert rel_l2(kpsi, st.kpsi) < TOL
    # real time from the relaxed state, as the reference restarts from ARGLE
    # output (dealias + psi = IFFT(kpsi))
    q = G.params_from_namelist(N, 5e-4, 2.5)
    h = G.GPE(q)
    h.set_kpsi(np.ascontiguousarray(kpsi))
    st2 = O.State(O.make_params(N, 5e-4, 2.5))
    st2.kpsi[...] = kpsi
    st2.ic_sync()
    h.gpe_rk4(4)
    st2.rk4(4)
    e = rel_l2(h.kpsi, st2.kpsi)
    d = h.diagnostics_gpe()
    scal, spec, _ = st2.diagnostics()
    e_spec = max(rel_l2(d["spec"][i], spec[i]) for i in range(6))
    say("  rel-L2 kpsi %.2e  worst spectrum %.2e" % (e, e_spec))
    assert e < TOL and e_spec < TOL
    for i in (0, 1, 2, 3, 4):
        assert abs(d["scal"][i] - scal[i]) <= 1e-11 * abs(scal[i])
    g.close()
    h.close()


def test_conservation_4096():
    """Full size, where the oracle is too slow: 20 RK4 steps of the 4096^2
    workload (256 random vortices, dt = 2.5e-4) conserve norm and energy to
    the scheme's accuracy, and momentum is conserved on the energy scale."""
    from quantumvortices_amd import gpe as G
    N = 4096
    say("case conservation 4096^2")
    xs, ys, qs = G.random_vortices(256, seed=3)
    g = G.GPE(G.params_from_namelist(N, 2.5e-4, 2.5))
    g.set_ic_vortices(xs, ys, qs, nimg=3)
    d0 = g.diagnostics_gpe()
    g.gpe_rk4(20)
    d1 = g.diagnostics_gpe()
    dn = abs(d1["norm"] - d0["norm"]) / d0["norm"]
    de = abs(d1["etot"] - d0["etot"]) / abs(d0["etot"])
    say("  norm %.15e -> %.15e (rel %.2e)  etot %.12e -> %.12e (rel %.2e)  mom %.2e %.2e"
        % (d0["norm"], d1["norm"], dn, d0["etot"], d1["etot"], de, d1["momx"], d1["momy"]))
    dm = abs(d1["momx"] - d0["momx"]) + abs(d1["momy"] - d0["momy"])
    say("  momentum change %.2e" % dm)
    assert dn < 1e-12 and de < 1e-9
    assert dm < 1e-9 * (abs(d0["etot"]) + abs(d0["momx"]) + abs(d0["momy"]))
    g.close()


# ppotd=2 (width 5 dx): potential and force evaluated in real space inside
# the nonlinear pass; ppotd=1 (width 2.5 dx): spectral translation path
@pytest.mark.parametrize("ppotd", [2.0, 1.0])
def test_particles_rk4_vs_oracle(G, oracle_mod, ppotd):
    N, dt, steps, npart = 64, 1e-3, 5, 3
    say("case particles rk4 N=%d np=%d steps=%d ppotd=%g" % (N, npart, steps, ppotd))
    O = oracle_mod
    psi0 = _ic(G, N)
    rng = np.random.default_rng(4)
    pos = rng.uniform(0, 2 * np.pi, (npart, 2))
    vel = rng.uniform(-0.1, 0.1, (npart, 2))
    fext = rng.uniform(-0.1, 0.1, (npart, 2))
    kw = dict(gp_N_obj=npart, ppotV0=10.0, ppotd=ppotd, gp_mass_part=5.0)
    p = G.params_from_namelist(N, dt, 2.5, **kw)
    g = G.GPE(p)
    g.set_psi(psi0, sync=True)
    g.initialize_particles(pos, vel, fext)
    g.gpe_rk4(steps)
    op = O.make_params(N, dt, 2.5, np_=npart, V0_in=10.0, d_in=ppotd, mass=5.0)
    st = O.State(op)
    st.set_psi(psi0, sync=True)
    st.part(0)[...] = pos
    st.part(1)[...] = vel
    st.part(2)[...] = fext
    st.particle_potential()
    st.rk4(steps)
    gp, gv, _ = g.particles()
    e = rel_l2(g.psi, st.psi)
    say("  rel-L2 psi %.3e pos %.3e vel %.3e" % (e, rel_l2(gp, st.part(0)), rel_l2(gv, st.part(1))))
    assert e < TOL
    assert rel_l2(gp, st.part(0)) < TOL
    assert rel_l2(gv, st.part(1)) < 1e-10
    d = g.diagnostics_gpe()
    scal, spec, pscal = st.diagnostics()
    say("  frc gpu", d["frc"].ravel(), "orc", st.part(3).ravel())
    say("  pscal gpu", [d["eppot"], d["eobj_kin"], d["esrep"], *d["momobj"]], "orc", list(pscal))
    assert rel_l2(d["frc"], st.part(3)) < 1e-10
    assert abs(d["eppot"] - pscal[0]) <= 1e-11 * abs(pscal[0])
    # particle kinetic energy, short-range energy and momentum
    # (src/diagnostics.f90:210-239)
    assert abs(d["eobj_kin"] - pscal[1]) <= 1e-10 * abs(pscal[1])
    assert abs(d["esrep"] - pscal[2]) <= 1e-11 * abs(pscal[2])
    assert np.abs(d["momobj"] - pscal[3:5]).max() <= 1e-10 * np.abs(pscal[3:5]).max()
    assert pscal[1] > 0 and pscal[2] > 0
    g.close()


def test_particles_wide_potential_full_window(G, oracle_mod):
    """Real-space potential path with a wide potential (d = 0.3 at 256^2):
    the periodic coordinate saturates near the half box, so the 10 d window
    would cut V off at ~1e-14 V0; the window then spans the whole period."""
    N, dt, steps, npart = 256, 1e-3, 3, 3
    O = oracle_mod
    ppotd = 0.3 / (2.5 * 2 * np.pi / N)
    say("case particles wide potential N=%d d=0.3 (ppotd=%.3f)" % (N, ppotd))
    psi0 = _ic(G, N)
    pos = np.array([[1.2, 0.4], [3.3, 2.9], [5.1, 5.6]])
    vel = np.array([[0.05, -0.02], [0.0, 0.03], [-0.04, 0.01]])
    kw = dict(gp_N_obj=npart, ppotV0=1.0, ppotd=ppotd, gp_mass_part=50.0)
    g = G.GPE(G.params_from_namelist(N, dt, 2.5, **kw))
    g.set_psi(psi0, sync=True)
    g.initialize_particles(pos, vel)
    g.gpe_rk4(steps)
    st = O.State(O.make_params(N, dt, 2.5, np_=npart, V0_in=1.0, d_in=ppotd, mass=50.0))
    st.set_psi(psi0, sync=True)
    st.part(0)[...] = pos
    st.part(1)[...] = vel
    st.particle_potential()
    st.rk4(steps)
    gp, gv, _ = g.particles()
    e = rel_l2(g.psi, st.psi)
    say("  rel-L2 psi %.3e pos %.3e vel %.3e" % (e, rel_l2(gp, st.part(0)), rel_l2(gv, st.part(1))))
    assert e < TOL and rel_l2(gp, st.part(0)) < TOL and rel_l2(gv, st.part(1)) < 1e-10
    g.close()


@pytest.mark.parametrize("ppotd", [2.0, 1.0])
def test_particles_argle_vs_oracle(G, oracle_mod, ppotd):
    N, dt, steps, npart = 64, 1e-2, 8, 2
    say("case particles argle N=%d np=%d ppotd=%g" % (N, npart, ppotd))
    O = oracle_mod
    psi0 = _ic(G, N)
    pos = np.array([[1.1, 2.3], [4.4, 0.7]])
    p = G.params_from_namelist(N, dt, 2.5, gp_N_obj=npart, ppotV0=10.0, ppotd=ppotd, gp_mass_part=5.0)
    g = G.GPE(p)
    g.set_psi(psi0, sync=False)
    g.initialize_particles(pos)
    r = g.ARGLE(nsteps=steps)
    st = O.State(O.make_params(N, dt, 2.5, np_=npart, V0_in=10.0, d_in=ppotd, mass=5.0))
    st.set_psi(psi0, sync=False)
    st.part(0)[...] = pos
    st.particle_potential()
    ro = st.argle(steps)
    e = rel_l2(g.kpsi, st.kpsi)
    say("  rel-L2 kpsi %.3e residual %.6e / %.6e" % (e, r, ro))
    assert e < TOL
    assert abs(r - ro) <= 1e-9 * abs(ro)
    g.close()


def test_prefetching_p1_equals_plain_p1(monkeypatch):
    """At 4096^2 the RK4 P1 runs as a persistent kernel that prefetches the
    next row pair into LDS when VK_P1_PREFETCH=1 (the default is the plain
    row pass).
    Both compute the same transforms, so 5 RK4 steps with particles agree to
    roundoff (the two kernels may contract multiply-adds differently)."""
    from quantumvortices_amd import gpe as G
    N = 4096
    say("case prefetching P1 vs plain P1 at 4096^2")
    kw = dict(gp_N_obj=8, ppotV0=10.0, ppotd=2.0, gp_mass_part=5.0)
    xs, ys, qs = G.random_vortices(64, seed=5)
    pos = np.random.default_rng(6).uniform(0, 2 * np.pi, (8, 2))
    out = []
    for on in ("1", "0"):
        monkeypatch.setenv("VK_P1_PREFETCH", on)
        g = G.GPE(G.params_from_namelist(N, 2.5e-4, 2.5, **kw))
        g.set_ic_vortices(xs, ys, qs, nimg=3)
        g.initialize_particles(pos)
        g.gpe_rk4(5)
        out.append((g.psi, g.particles()[0]))
        g.close()
    e = rel_l2(out[0][0], out[1][0])
    ep = rel_l2(out[0][1], out[1][1])
    say("  rel-L2 psi %.3e  positions %.3e" % (e, ep))
    assert e < 1e-14 and ep < 1e-14
