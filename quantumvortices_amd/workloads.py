"""Seeded synthetic workloads of the 2D GPE path (no dataset, no checkpoint).

The particle workload: 4096^2 periodic box L = 2 pi, g = 1, xi = 2.5 dx,
density 1, 256 random-sign vortices (seed 3, zero net charge, Padé cores),
inertial particles at seeded uniform positions (seed 4) with the Gaussian
potential V0 = 10 g, width d = 2 xi, and mass 5x the fluid one particle
displaces, RK4 dt = 2.5e-4.  The reference takes gp_mass_part as a raw
namelist input (src/driver.f90:74); here it is derived from the ground state.
"""
import math

import numpy as np

from . import gpe as G

XI_IN = 2.5          # xi in units of dx
V0_IN = 10.0         # ppotV0 in units of g
D_IN = 2.0           # ppotd in units of xi
MASS_FACTOR = 5.0


def displaced_fluid(N, xi_in=XI_IN, V0_in=V0_IN, d_in=D_IN, n_solve=512, tol=1e-10):
    """Integral of (1 - rho_GS) over the box for one particle at rest, where
    rho_GS is the ARGLE ground state (omega = g, far-field density 1) around
    the particle's potential, on an N^2 grid with xi = xi_in dx.

    In units of xi the problem does not depend on N once the box is many
    healing lengths wide (the depletion decays exponentially), so the ground
    state is relaxed on an n_solve^2 grid with the same xi/dx and d/xi and the
    area integral is scaled by (dx_N/dx_solve)^2 = (n_solve/N)^2."""
    p = G.params_from_namelist(n_solve, 0.02, xi_in, gp_N_obj=1, ppotV0=V0_in, ppotd=d_in, gp_mass_part=1.0)
    g = G.GPE(p)
    try:
        g.set_psi(np.ones((n_solve, n_solve), dtype=np.complex128), sync=False)
        g.initialize_particles(np.array([[math.pi, math.pi]]))
        steps, res = g.argle_until(tol=tol, max_steps=200000, check_every=200)
        if not res <= tol:
            raise RuntimeError("displaced_fluid: ARGLE did not reach %g (residual %g)" % (tol, res))
        rho = np.abs(g.psi) ** 2
    finally:
        g.close()
    dx = 2.0 * math.pi / n_solve
    return float(np.sum(1.0 - rho) * dx * dx) * (n_solve / N) ** 2


def particle_positions(npart, seed=4, L=2.0 * math.pi):
    return np.random.default_rng(seed).uniform(0.0, L, (npart, 2))


def particle_workload(N=4096, dt=2.5e-4, npart=32, nv=256, seed_v=3, seed_p=4, mass=None, sync=True):
    """GPE handle holding the seeded vortex field (built on the device) and,
    for npart > 0, the particles.  Returns (handle, info) with info holding
    the vortex list, positions and the particle parameters."""
    xs, ys, qs = G.random_vortices(nv, seed_v)
    kw = {}
    pos = None
    if npart > 0:
        if mass is None:
            mass = MASS_FACTOR * displaced_fluid(N)
        kw = dict(gp_N_obj=npart, ppotV0=V0_IN, ppotd=D_IN, gp_mass_part=mass)
        pos = particle_positions(32, seed_p)[:npart] if npart <= 32 else particle_positions(npart, seed_p)
    p = G.params_from_namelist(N, dt, XI_IN, **kw)
    g = G.GPE(p)
    g.set_ic_vortices(xs, ys, qs, nimg=3, sync=sync)
    if npart > 0:
        g.initialize_particles(pos)
    return g, dict(xs=xs, ys=ys, qs=qs, pos=pos, kw=kw, mass=mass)
