"""Restart equals continuous: a run checkpointed to the reference's wf and
particle files halfway and restarted in a fresh handle (restart_gpe ->
kpsi = FFT(psi), dealiased at the first RK4 entry, as the reference's
gpe_rk4 does) ends where the uninterrupted run ends, to the roundoff of one
FFT round trip."""
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("N,nsplit", [(128, 1), (256, 4)])
def test_restart_equals_continuous(tmp_path, N, nsplit):
    from quantumvortices_amd import gpe as G
    say("case restart N=%d nsplit=%d" % (N, nsplit))
    dt, half = 1e-3, 50
    p = G.params_from_namelist(N, dt, 2.5)
    xs, ys, qs = G.lattice_vortices(4)
    dx = 2 * np.pi / N
    psi0 = G.ic_vortices_host(N, xs + 0.31 * dx, ys + 0.17 * dx, qs, p.xi)
    a = G.GPE(p)
    a.set_psi(psi0, sync=True)
    a.gpe_rk4(2 * half)
    ref = a.psi
    b = G.GPE(p)
    b.set_psi(psi0, sync=True)
    b.gpe_rk4(half)
    b.io_fields(tmp_path, 1, half * dt, nsplit=nsplit)
    c = G.GPE(p)
    icp, t, dtr = c.restart_gpe(tmp_path, 1, nsplit_orig=nsplit)
    assert (icp, t, dtr) == (1, half * dt, dt)
    c.gpe_rk4(half)
    e = rel_l2(c.psi, ref)
    say("  rel-L2 psi restarted vs continuous %.3e" % e)
    assert e < 1e-13
    for g in (a, b, c):
        g.close()


def test_restart_with_particles(tmp_path):
    from quantumvortices_amd import gpe as G
    import shutil
    N, dt, half, npart = 64, 1e-3, 10, 3
    say("case restart with particles N=%d" % N)
    kw = dict(gp_N_obj=npart, ppotV0=10.0, ppotd=2.0, gp_mass_part=5.0)
    p = G.params_from_namelist(N, dt, 2.5, **kw)
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, 2 * np.pi, (npart, 2))
    vel = rng.uniform(-0.1, 0.1, (npart, 2))
    fext = rng.uniform(-0.1, 0.1, (npart, 2))
    xs, ys, qs = G.lattice_vortices(2)
    psi0 = G.ic_vortices_host(N, xs + 0.31 * 2 * np.pi / N, ys + 0.17 * 2 * np.pi / N, qs, p.xi)
    a = G.GPE(p)
    a.set_psi(psi0, sync=True)
    a.initialize_particles(pos, vel, fext)
    a.gpe_rk4(2 * half)
    pa, va, _ = a.particles()
    b = G.GPE(p)
    b.set_psi(psi0, sync=True)
    b.initialize_particles(pos, vel, fext)
    b.gpe_rk4(half)
    b.io_fields(tmp_path, 2, half * dt)
    b.gppart_checkpoint(tmp_path, 0, 2, half * dt)
    for name in ("gppos", "gpvel", "gpf0ext"):    # the driver's restart/part copies
        shutil.copy(tmp_path / ("%s_ckp.dat_nrst0" % name), tmp_path / ("%s.dat_nrst0" % name))
    c = G.GPE(p)
    c.restart_gpe(tmp_path, 2)
    c.initialize_particles_from(tmp_path, 0)
    c.gpe_rk4(half)
    pc, vc, _ = c.particles()
    e = rel_l2(c.psi, a.psi)
    say("  rel-L2 psi %.3e  pos %.3e  vel %.3e" % (e, rel_l2(pc, pa), rel_l2(vc, va)))
    assert e < 1e-13 and rel_l2(pc, pa) < 1e-13 and rel_l2(vc, va) < 1e-10
    for g in (a, b, c):
        g.close()
