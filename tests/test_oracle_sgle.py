"""Pins the oracle's SGLE step (GPE_SGLE, src/gpe.f90:149-217, 226-266) and
its ARGLE with a caller-supplied advection velocity, by what the
mathematics fixes:

  noise term      at g = 0, V = 0, nuN = 0 one step is linear:
                  kpsi' = (kpsi (1 + omega dt) + sqrt(dt) P_G[F(sigma xi/sqrt(dvol))])
                          / (1 + alpha k^2 dt),  sigma = sqrt(oneybeta/(2 alpha))
                  (numpy FFT of the supplied noise; the noise is dealiased,
                  the rest is not)
  zero noise      oneybeta = 0, nuN = 0: SGLE equals plain ARGLE bit for bit
  chemical pot.   uniform |psi|^2 = rho0: omega' = omega - dt nuN (rho0 - 1)
                  (norm = rho0 L^2, Nparticle = L^2), used in the same step
  ktrunc          modes with k^2 > ktrunc^2 are zero after a step
  velocity field  the Taylor-Green field supplied as an array equals the
                  built-in Taylor-Green term bit for bit; the imposed
                  vortices' current u = 2 alpha rho0 grad(theta0) keeps them
                  in place while plain ARGLE lets them drift
"""
import math

import numpy as np

from conftest import rel_l2


def kwn_np(N):
    i = np.arange(N)
    return np.where(i < N // 2, i, i - N)


def mask(N):
    k = kwn_np(N)
    return 9 * (k[:, None] ** 2 + k[None, :] ** 2) >= N * N


def field(N, seed):
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(N) / N
    X, Y = np.meshgrid(x, x)
    psi = np.ones((N, N), complex)
    for _ in range(4):
        kx, ky = rng.integers(-6, 7, 2)
        psi += 0.2 * (rng.standard_normal() + 1j * rng.standard_normal()) * np.exp(1j * (kx * X + ky * Y))
    return psi


def noise(N, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))


def test_sgle_noise_term_closed_form(oracle_mod):
    O = oracle_mod
    N, dt, beta = 32, 1e-2, 0.3
    p = O.make_params(N, dt, 2.5, sgle=1, oneybeta=beta, nuN=0.0, chempot0=0.7)
    p.g = 0.0
    st = O.State(p)
    psi = field(N, 1)
    st.set_psi(psi, sync=False)
    k0 = st.kpsi.copy()
    xi = noise(N, 2)
    st.set_noise(xi)
    st.argle(1)
    sigma = math.sqrt(beta / (2 * p.alpha))
    nk = np.fft.fft2(sigma * xi / math.sqrt(p.dvol)).T * math.sqrt(dt)
    nk[mask(N)] = 0
    k = kwn_np(N)
    k2 = (k[:, None] ** 2 + k[None, :] ** 2).astype(float)
    expect = (k0 * (1 + 0.7 * dt) + nk) / (1 + p.alpha * k2 * dt)
    assert rel_l2(st.kpsi, expect) < 1e-13
    assert st.omega == 0.7                                      # nuN = 0
    # the noise is really dealiased and really there
    assert np.abs(st.kpsi[mask(N)] - (k0 * (1 + 0.7 * dt) / (1 + p.alpha * k2 * dt))[mask(N)]).max() < 1e-12
    assert rel_l2(st.kpsi, expect - nk / (1 + p.alpha * k2 * dt)) > 1e-3


def test_sgle_zero_noise_is_argle(oracle_mod):
    O = oracle_mod
    N, dt = 32, 1e-2
    psi = field(N, 3)
    a = O.State(O.make_params(N, dt, 2.5))
    b = O.State(O.make_params(N, dt, 2.5, sgle=1, oneybeta=0.0, nuN=0.0, chempot0=a.p.omega))
    a.set_psi(psi, sync=False)
    b.set_psi(psi, sync=False)
    b.set_noise(noise(N, 4))
    ra, rb = a.argle(3), b.argle(3)
    assert np.array_equal(a.kpsi, b.kpsi) and ra == rb


def test_sgle_chemical_potential_update(oracle_mod):
    O = oracle_mod
    N, dt, nuN, rho0, om = 32, 1e-2, 2.5, 1.21, 0.9
    p = O.make_params(N, dt, 2.5, sgle=1, oneybeta=0.0, nuN=nuN, chempot0=om)
    p.g = 0.0
    st = O.State(p)
    st.set_psi(np.full((N, N), math.sqrt(rho0), complex), sync=False)
    k00 = st.kpsi[0, 0]
    st.argle(1)
    om1 = om - dt * nuN * (rho0 - 1.0)
    assert abs(st.omega - om1) < 1e-14
    # the updated omega is the one the same step uses (k = 0: factor 1 + omega' dt)
    assert abs(st.kpsi[0, 0] / k00 - (1 + om1 * dt)) < 1e-14
    # and the sign: a norm above L^2 lowers omega
    assert st.omega < om


def test_sgle_ktrunc(oracle_mod):
    O = oracle_mod
    N, dt, kt = 32, 1e-2, 5.5
    st = O.State(O.make_params(N, dt, 2.5, sgle=1, oneybeta=0.1, nuN=0.1, chempot0=1.0, ktrunc=kt))
    st.set_psi(field(N, 5), sync=False)
    st.set_noise(noise(N, 6))
    st.argle(1)
    k = kwn_np(N)
    k2 = k[:, None] ** 2 + k[None, :] ** 2
    assert np.abs(st.kpsi[k2 > kt * kt]).max() == 0.0
    assert np.abs(st.kpsi[k2 <= kt * kt]).min() > 0.0


def test_supplied_velocity_equals_taylor_green(oracle_mod):
    O = oracle_mod
    N, dt = 32, 1e-2
    psi = field(N, 7)
    x = 2 * np.pi * np.arange(N) / N
    X, Y = np.meshgrid(x, x)
    a = O.State(O.make_params(N, dt, 2.5, adv=1, vnx=0.2, vny=-0.1))
    b = O.State(O.make_params(N, dt, 2.5, adv=2, vnx=0.2, vny=-0.1))
    b.set_velocity(np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y))
    a.set_psi(psi, sync=False)
    b.set_psi(psi, sync=False)
    a.argle(2)
    b.argle(2)
    assert np.array_equal(a.kpsi, b.kpsi)


def vortex_velocity(N, xs, ys, qs, alpha, nimg=3):
    """u = 2 alpha grad(theta) of the periodic product phase, by central
    differences of the phase of the product e^{i theta} (independent of the
    analytic gradient the library uses)."""
    from quantumvortices_amd import gpe as G
    h = 1e-5
    x = 2 * np.pi * np.arange(N) / N
    X, Y = np.meshgrid(x, x)

    def ph(dx, dy):
        z = np.ones((N, N), complex)
        for xv, yv, q in zip(xs, ys, qs):
            a = X + dx - xv
            for m in range(-nimg, nimg + 1):
                b = Y + dy - yv - 2 * np.pi * m
                w = np.sin(0.5 * a) * np.cosh(0.5 * b) + 1j * np.cos(0.5 * a) * np.sinh(0.5 * b)
                w /= np.abs(w)
                z *= w if q > 0 else np.conj(w)
        lin = sum(q * xv for xv, q in zip(xs, qs))
        return z * np.exp(-1j * lin * (Y + dy) / (2 * np.pi))

    gx = np.angle(ph(h, 0) / ph(-h, 0)) / (2 * h)
    gy = np.angle(ph(0, h) / ph(0, -h)) / (2 * h)
    return 2 * alpha * gx, 2 * alpha * gy


def _core(a, xv, yv, dx, half=6):
    """grid minimum of |psi| in a window around (xv, yv): (value, x, y)"""
    N = a.shape[0]
    i0, j0 = int(round(xv / dx)), int(round(yv / dx))
    best = (np.inf, 0.0, 0.0)
    for j in range(j0 - half, j0 + half + 1):
        for i in range(i0 - half, i0 + half + 1):
            v = a[j % N, i % N]
            if v < best[0]:
                best = (v, i * dx, j * dx)
    return best


def test_imposed_vortex_velocity_keeps_vortices(oracle_mod):
    """ARGLE with the imposed vortices' current as the advection field,
    u = 2 alpha rho0 grad(theta0) (rho0, theta0: density and phase of the
    Pade product ansatz; u is smooth and vanishes in the cores), relaxes the
    cores in place: the phase winds by q around each imposed position and
    |psi| has a core there, while plain ARGLE lets the pair drift together."""
    O = oracle_mod
    from quantumvortices_amd import gpe as G
    N, dt = 64, 2e-2
    xs = np.array([1.5, 4.7])
    ys = np.array([3.1, 3.3])
    qs = np.array([1, -1], dtype=np.int32)
    p = O.make_params(N, dt, 2.5, adv=2)
    psi0 = G.ic_vortices_host(N, xs, ys, qs, p.xi)
    ux, uy = vortex_velocity(N, xs, ys, qs, p.alpha)
    r0 = np.abs(psi0) ** 2
    st = O.State(p)
    st.set_velocity(ux * r0, uy * r0)
    st.set_psi(psi0, sync=False)
    r = st.argle(1000)
    psi = O.fft2d(st.kpsi, +1)
    plain = O.State(O.make_params(N, dt, 2.5))
    plain.set_psi(psi0, sync=False)
    plain.argle(1000)
    psi_p = O.fft2d(plain.kpsi, +1)
    dx = 2 * np.pi / N

    def winding(f, xv, yv, rad=4):
        i0, j0 = int(round(xv / dx)), int(round(yv / dx))
        path = [(i0 + a, j0 - rad) for a in range(-rad, rad)] + [(i0 + rad, j0 + a) for a in range(-rad, rad)] + \
               [(i0 - a, j0 + rad) for a in range(-rad, rad)] + [(i0 - rad, j0 - a) for a in range(-rad, rad)]
        ph = np.array([np.angle(f[j % N, i % N]) for i, j in path])
        d = np.diff(np.concatenate([ph, ph[:1]]))
        return int(round(np.sum((d + np.pi) % (2 * np.pi) - np.pi) / (2 * np.pi)))

    for xv, yv, q in zip(xs, ys, qs):
        assert winding(psi, xv, yv) == q
        v, xc, yc = _core(np.abs(psi), xv, yv, dx)
        assert v < 0.3 and math.hypot(xc - xv, yc - yv) < 1.5 * dx
        # plain ARGLE: the pair attracts and the cores leave their places
        vp, xp, yp = _core(np.abs(psi_p), xv, yv, dx)
        assert math.hypot(xp - xv, yp - yv) > 3 * dx
    assert r < 1e-3


def test_philox_known_answers():
    """The generator's reference restatement against the published
    Philox4x32-10 known-answer vectors (Random123 kat_vectors)."""
    from philox_ref import philox4x32_10
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, out in kat:
        got = philox4x32_10(*[np.array([v], np.uint32) for v in ctr], *key)
        assert tuple(int(g[0]) for g in got) == out


def test_vortex_velocity_host_matches_phase_gradient():
    """The library's imposed vortex velocity (analytic gradient of the
    product phase, times the Padé density) against central differences of
    the phase, away from the cores."""
    from quantumvortices_amd import gpe as G
    N = 64
    rng = np.random.default_rng(3)
    xs, ys = rng.uniform(0, 2 * np.pi, 6), rng.uniform(0, 2 * np.pi, 6)
    qs = np.array([1, -1, 1, -1, 1, -1], dtype=np.int32)
    xi, alpha = 2.5 * 2 * np.pi / N, 0.0376
    ux, uy = G.vortex_velocity_host(N, xs, ys, qs, xi, alpha)
    fx, fy = vortex_velocity(N, xs, ys, qs, alpha)
    rho0 = np.abs(G.ic_vortices_host(N, xs, ys, qs, xi)) ** 2
    x = 2 * np.pi * np.arange(N) / N
    X, Y = np.meshgrid(x, x)
    far = np.ones((N, N), bool)
    for xv, yv in zip(xs, ys):
        dx = (X - xv + np.pi) % (2 * np.pi) - np.pi
        dy = (Y - yv + np.pi) % (2 * np.pi) - np.pi
        far &= dx * dx + dy * dy > (3 * xi) ** 2
    sc = np.abs(fx * rho0)[far].max()
    assert np.abs(ux - fx * rho0)[far].max() < 1e-7 * sc
    assert np.abs(uy - fy * rho0)[far].max() < 1e-7 * sc
    assert np.all(np.isfinite(ux)) and np.all(np.isfinite(uy))
