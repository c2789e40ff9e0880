"""The 4096^2 particle workload against the oracle.

Field: 256 random-sign vortices (seed 3), the first 4 of the workload's
seeded particle positions (seed 4), V0 = 10 g, d = 2 xi (the real-space
potential path inside the nonlinear row pass), mass 5x the displaced fluid.
Two RK4 steps on the GPU, with the persistent prefetching P1 and with the
plain P1 (VK_P1_PREFETCH=1 / 0), and with the default row-pass shapes (one
row per workgroup in P1 and P3, 4-column panels), two rows in both, or one
row in both set explicitly (VK_P1_TR1, VK_P3_TR1), each against the same two
oracle steps
(src/gpe.f90:310-494, src/gpe_particle.f90:186-531), then the diagnostics
(energies, six spectra, particle forces and particle energies,
src/diagnostics.f90:54-533).  The oracle runs on the host's cores."""
import os
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-11
NPART, STEPS = 4, 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cfg():
    import torch
    assert torch.cuda.is_available()
    from quantumvortices_amd import workloads as W
    return W


@pytest.fixture(scope="module")
def oracle_run(cfg, oracle_mod):
    """psi0 from the device-built initial condition, then the oracle's two
    steps and diagnostics (computed once for both GPU variants)."""
    W = cfg
    O = oracle_mod
    mass = W.MASS_FACTOR * W.displaced_fluid(4096)
    say("case cfg4096: mass %.6e, building the initial field on the device" % mass)
    g, info = W.particle_workload(4096, 2.5e-4, npart=NPART, mass=mass)
    psi0 = g.psi
    g.close()
    O.set_threads(min(16, len(os.sched_getaffinity(0))))
    op = O.make_params(4096, 2.5e-4, W.XI_IN, np_=NPART, V0_in=W.V0_IN, d_in=W.D_IN, mass=mass)
    st = O.State(op)
    st.set_psi(psi0, sync=True)
    st.part(0)[...] = info["pos"]
    st.particle_potential()
    say("  oracle: %d RK4 steps at 4096^2 with %d particles" % (STEPS, NPART))
    st.rk4(STEPS)
    say("  oracle: diagnostics")
    scal, spec, pscal = st.diagnostics()
    return dict(mass=mass, info=info, psi=st.psi.copy(), kpsi=st.kpsi.copy(), pos=st.part(0).copy(),
                vel=st.part(1).copy(), frc=st.part(3).copy(), scal=scal, spec=spec, pscal=pscal)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("prefetch,tr1", [("0", None), ("0", "0"), ("1", "0"), ("0", "1")])
def test_cfg4096_particles_vs_oracle(cfg, oracle_run, monkeypatch, prefetch, tr1):
    W = cfg
    R = oracle_run
    monkeypatch.setenv("VK_P1_PREFETCH", prefetch)
    for k in ("VK_P1_TR1", "VK_P3_TR1"):
        if tr1 is None:
            monkeypatch.delenv(k, raising=False)     # the default row-pass shapes
        else:
            monkeypatch.setenv(k, tr1)
    say("case cfg4096 GPU, VK_P1_PREFETCH=%s VK_P1_TR1=VK_P3_TR1=%s" % (prefetch, tr1))
    g, _ = W.particle_workload(4096, 2.5e-4, npart=NPART, mass=R["mass"])
    g.gpe_rk4(STEPS)
    e_psi, e_k = rel_l2(g.psi, R["psi"]), rel_l2(g.kpsi, R["kpsi"])
    gp, gv, _ = g.particles()
    e_pos, e_vel = rel_l2(gp, R["pos"]), rel_l2(gv, R["vel"])
    say("  rel-L2 psi %.3e kpsi %.3e pos %.3e vel %.3e |vel| %.3e" % (e_psi, e_k, e_pos, e_vel, np.abs(gv).max()))
    # the particles do move: mass 5x the displaced fluid (~7e-3) and forces
    # of ~1e-9 give velocities of ~1e-11 after two steps
    assert np.abs(R["vel"]).max() > 1e-12
    assert e_psi < TOL and e_k < TOL
    assert e_pos < TOL and e_vel < 1e-10
    d = g.diagnostics_gpe()
    e_frc = rel_l2(d["frc"], R["frc"])
    e_spec = max(rel_l2(d["spec"][i], R["spec"][i]) for i in range(6))
    say("  forces rel-L2 %.3e worst spectrum %.3e" % (e_frc, e_spec))
    say("  pscal gpu", [d["eppot"], d["eobj_kin"], d["esrep"], *d["momobj"]])
    say("  pscal orc", list(R["pscal"]))
    assert e_frc < 1e-10 and e_spec < TOL
    for i in (0, 1, 2, 3, 4):
        assert abs(d["scal"][i] - R["scal"][i]) <= 1e-11 * abs(R["scal"][i])
    ps = R["pscal"]
    assert abs(d["eppot"] - ps[0]) <= 1e-11 * abs(ps[0])
    assert abs(d["eobj_kin"] - ps[1]) <= 1e-10 * abs(ps[1])
    assert abs(d["esrep"] - ps[2]) <= 1e-11 * abs(ps[2]) + 1e-300
    mscale = np.abs(ps[3:5]).max()
    assert np.abs(d["momobj"] - ps[3:5]).max() <= 1e-10 * mscale
    g.close()
