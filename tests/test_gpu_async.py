"""Host calls made while steps are still queued on the device: the C ABI's
setters and getters order themselves after the handle's outstanding work
(the asynchronous step entry point returns as soon as the launches are
queued; the particle update runs on a second stream)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available()
    from quantumvortices_amd import gpe
    return gpe


def _run(G, sync_before_set):
    """20 queued RK4 steps at 2048^2 with 8 real-space particles, then new
    particle positions and velocities, then 3 more steps."""
    N, npart = 2048, 8
    p = G.params_from_namelist(N, 2.5e-4, 2.5, gp_N_obj=npart, ppotV0=10.0, ppotd=2.0, gp_mass_part=1e-2)
    g = G.GPE(p)
    xs, ys, qs = G.random_vortices(32, 11)
    g.set_ic_vortices(xs, ys, qs)
    rng = np.random.default_rng(5)
    g.initialize_particles(rng.uniform(0, 2 * np.pi, (npart, 2)))
    pos2 = rng.uniform(0, 2 * np.pi, (npart, 2))
    vel2 = rng.uniform(-0.05, 0.05, (npart, 2))
    g.rk4_async(20)
    if sync_before_set:
        _ = g.psi                       # drains the handle's work
    g.initialize_particles(pos2, vel2)
    rp, rv, _ = g.particles()
    g.gpe_rk4(3)
    out = (rp, rv, g.psi, g.particles())
    g.close()
    return pos2, vel2, out


def test_particles_set_after_queued_steps(G):
    """vk_gpe_particles_set right after vk_gpe_rk4_steps_async: the state
    read back is the one set (the queued steps' particle updates do not
    overwrite it), and the run continues exactly as when the host waits
    for the queued steps before setting."""
    say("case particles_set after 20 queued steps, N=2048, 8 particles")
    pos2, vel2, (rp, rv, psi_a, parts_a) = _run(G, False)
    assert np.array_equal(rp, pos2) and np.array_equal(rv, vel2)
    _, _, (_, _, psi_b, parts_b) = _run(G, True)
    assert np.array_equal(psi_a, psi_b)
    for a, b in zip(parts_a, parts_b):
        assert np.array_equal(a, b)
