"""Pins the CPU oracle (oracle/vk_oracle.c) before the GPU tests trust it.
The reference ships no tests, fixtures or golden vectors for this
path and its transform library (FFTW3-MPI) is absent, so every oracle
function is pinned by what the mathematics fixes, each check chosen so that a
plausible mistake (dropped term, wrong sign, index or scale, transposed
operand) fails it:

  orc_fft2d          brute-force DFT (N<=16), numpy.fft golden fixtures,
                     delta<->constant, plane wave, Parseval, inverse o forward
  orc_kwn/mask/shell integer identities over whole grids
  orc_set_grid/constants  survey table values, alpha = xi^2, g = 1
  orc_rk4_step       g=0 exact per-mode RK4 polynomial (with and without
                     dissipation), plane-wave exact solution, uniform state
                     stationary, norm and energy conservation
  orc_argle_step     g=0 exact per-mode factor and residual, uniform fixed
                     point, energy decrease
  orc_diagnostics    plane-wave closed forms for every energy, momentum and
                     spectrum; Helmholtz orthogonality; pure-gradient field;
                     Parseval of the occupation spectrum; shell placement
  particles          peak/symmetry/Gaussian integral of V, exact grid shift
                     of the translated potential, zero force in uniform
                     fluid, force by an independent route (analytic gradient
                     of V against the spectral derivative of rho), Newton's
                     third law for the short-range force, exact free and
                     uniformly accelerated particle motion
"""
import math
import os

import numpy as np
import pytest

from conftest import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))


def kwn_np(N):
    i = np.arange(N)
    return np.where(i < N // 2, i, i - N)


# ---------------------------------------------------------------- FFT -------

def brute_dft2(a, sign):
    N = a.shape[0]
    n = np.arange(N)
    W = np.exp(sign * 2j * np.pi * np.outer(n, n) / N)
    return a, W


@pytest.mark.parametrize("N", [4, 8, 16])
def test_fft_bruteforce(oracle_mod, N):
    O = oracle_mod
    rng = np.random.default_rng(N)
    a = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    # forward: out[kx, ky] = sum_{y,x} a[y,x] exp(-2 pi i (kx x + ky y)/N)
    out = np.zeros((N, N), complex)
    for kx in range(N):
        for ky in range(N):
            ph = np.exp(-2j * np.pi * (kx * np.arange(N)[None, :] + ky * np.arange(N)[:, None]) / N)
            out[kx, ky] = np.sum(a * ph)
    assert rel_l2(O.fft2d(a, -1), out) < 1e-14
    # inverse: x[y, x] = N^-2 sum_{kx,ky} K[kx,ky] exp(+2 pi i (kx x + ky y)/N)
    K = a
    xs = np.zeros((N, N), complex)
    for y in range(N):
        for x in range(N):
            ph = np.exp(2j * np.pi * (np.arange(N)[:, None] * x + np.arange(N)[None, :] * y) / N)
            xs[y, x] = np.sum(K * ph) / N ** 2
    assert rel_l2(O.fft2d(K, +1), xs) < 1e-14


@pytest.mark.parametrize("N", [8, 16, 32])
def test_fft_golden(oracle_mod, N):
    O = oracle_mod
    g = np.load(os.path.join(HERE, "golden", "fft_N%d.npz" % N))
    assert rel_l2(O.fft2d(g["x"], -1), g["fwd"]) < 1e-14
    assert rel_l2(O.fft2d(g["inv_in"], +1), g["inv"]) < 1e-14


def test_fft_identities(oracle_mod):
    O = oracle_mod
    N = 64
    d = np.zeros((N, N), complex)
    d[0, 0] = 1.0
    assert np.allclose(O.fft2d(d, -1), 1.0, atol=1e-15)
    # plane wave exp(i(kx x + ky y)) -> N^2 at [kx, ky]
    x = 2 * np.pi * np.arange(N) / N
    kx, ky = 5, -7
    pw = np.exp(1j * (kx * x[None, :] + ky * x[:, None]))
    K = O.fft2d(pw, -1)
    ref = np.zeros((N, N), complex)
    ref[kx % N, ky % N] = N * N
    assert np.abs(K - ref).max() < 1e-10
    rng = np.random.default_rng(3)
    a = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    Ka = O.fft2d(a, -1)
    assert abs(np.sum(np.abs(Ka) ** 2) / N ** 2 - np.sum(np.abs(a) ** 2)) < 1e-10 * np.sum(np.abs(a) ** 2)
    assert rel_l2(O.fft2d(Ka, +1), a) < 1e-14


def test_twiddles(oracle_mod):
    # against extended-precision sin/cos (float64 exp(-2 pi i k/N) itself
    # carries ~angle*eps error); every entry within one ulp
    pi_ld = np.longdouble("3.141592653589793238462643383279502884")
    for N in (4, 64, 4096):
        tw = oracle_mod.twiddles(N)
        ang = 2 * pi_ld * np.arange(N, dtype=np.longdouble) / N
        ref = np.cos(ang).astype(np.float64) - 1j * np.sin(ang).astype(np.float64)
        assert np.abs(tw.real - ref.real).max() <= 1.2e-16
        assert np.abs(tw.imag - ref.imag).max() <= 1.2e-16


# --------------------------------------------------- grid and constants -----

# nshell = int(1.414 N / 2) + 1 with 1.414 N in single precision
# (src/mod_grid.f90:63-101), worked out by hand
NSHELL_EXPECTED = {128: 91, 256: 181, 512: 362, 1024: 724, 4096: 2896}


@pytest.mark.parametrize("N", sorted(NSHELL_EXPECTED))
def test_grid_constants(oracle_mod, N):
    O = oracle_mod
    p = O.make_params(N, 1e-3, 2.5)
    assert p.nshell == NSHELL_EXPECTED[N]
    assert abs(p.kalias - N / 3.0) == 0.0
    assert abs(p.dx - 2 * math.pi / N) < 1e-18
    assert abs(p.alpha - p.xi ** 2) < 1e-15 * p.xi ** 2      # alpha = xi^2 for g = 1, rho = 1
    assert abs(p.g - 1.0) < 1e-14 and p.omega == p.g
    assert np.array_equal(O.kwn(N), kwn_np(N))               # Nyquist -> -N/2
    assert O.kwn(N)[N // 2] == -N // 2
    # the largest |k| on the grid falls in shell nshell
    assert O.shell(N // 2, N // 2) == p.nshell


@pytest.mark.parametrize("N", [32, 128, 512])
def test_mask_identity(oracle_mod, N):
    O = oracle_mod
    p = O.make_params(N, 1e-3, 2.5)
    k = kwn_np(N)
    for k1 in k:
        row = [O.mask(p, int(k1), int(k2)) for k2 in k]
        assert row == [int(9 * (int(k1) ** 2 + int(k2) ** 2) >= N * N) for k2 in k]


def test_mask_identity_4096_boundary(oracle_mod):
    O = oracle_mod
    N = 4096
    p = O.make_params(N, 1e-3, 2.5)
    for k1 in range(0, N // 2, 7):
        m = int(math.isqrt(max(0, (N * N - 1) // 9 - k1 * k1))) if 9 * k1 * k1 < N * N else 0
        for k2 in (m - 1, m, m + 1, m + 2):
            if k2 < 0:
                continue
            assert O.mask(p, k1, k2) == int(9 * (k1 * k1 + k2 * k2) >= N * N)


def test_shell_integer_identity(oracle_mod):
    # nint(sqrt(s)) == m  <=>  m(m-1) < s <= m(m+1), for every s <= 2*2048^2
    s = np.arange(0, 2 * 2048 ** 2 + 1, dtype=np.int64)
    m = np.rint(np.sqrt(s.astype(np.float64))).astype(np.int64)  # ties cannot occur for integers
    ok = (s == 0) | ((m * (m - 1) < s) & (s <= m * (m + 1)))
    assert ok.all()
    rng = np.random.default_rng(0)
    for k1, k2 in rng.integers(-2048, 2048, size=(2000, 2)):
        ss = int(k1) ** 2 + int(k2) ** 2
        mm = int(math.isqrt(ss))
        while mm * (mm + 1) < ss:
            mm += 1
        while mm > 0 and (mm - 1) * mm >= ss:
            mm -= 1
        assert oracle_mod.shell(int(k1), int(k2)) == mm


# ------------------------------------------------------------------- RK4 ----

def _state(O, N, dt, g=None, disp=0.0):
    p = O.make_params(N, dt, 2.5, disp=disp)
    if g is not None:
        p.g = g
    return O.State(p), p


def _random_dealiased(O, p, seed):
    N = p.N
    rng = np.random.default_rng(seed)
    K = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) * N
    k = kwn_np(N)
    m = 9 * (k[:, None] ** 2 + k[None, :] ** 2) >= N * N
    K[m] = 0
    return K


def _kk(N):
    k = kwn_np(N).astype(float)
    return k[:, None] ** 2 + k[None, :] ** 2       # [kx, ky]


@pytest.mark.parametrize("disp", [0.0, 0.3])
def test_rk4_linear_closed_form(oracle_mod, disp):
    O = oracle_mod
    N, dt, n = 32, 2e-2, 3
    st, p = _state(O, N, dt, g=0.0, disp=disp)
    K0 = _random_dealiased(O, p, 1)
    st.kpsi[...] = K0
    st.ic_sync()
    st.rk4(n)
    pref = complex(-disp, -1.0)                      # (1 - i disp)/i
    z = pref * (p.alpha * _kk(N) - p.omega) * dt
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    assert rel_l2(st.kpsi, K0 * R ** n) < 1e-13


def test_rk4_plane_wave_exact(oracle_mod):
    O = oracle_mod
    N, dt, n, A = 32, 1e-2, 4, 0.7
    st, p = _state(O, N, dt)
    x = p.dx * np.arange(N)
    kx, ky = 3, -2
    st.set_psi(A * np.exp(1j * (kx * x[None, :] + ky * x[:, None])), sync=True)
    st.rk4(n)
    # every RK stage stays a single mode a(t) e^{ik.x}: the step reduces to
    # classical RK4 on  da/dt = -i (alpha k^2 - omega + g |a|^2) a
    f = lambda a: -1j * (p.alpha * (kx * kx + ky * ky) - p.omega + p.g * abs(a) ** 2) * a
    a = complex(A)
    for _ in range(n):
        k1 = f(a); k2 = f(a + dt / 2 * k1); k3 = f(a + dt / 2 * k2); k4 = f(a + dt * k3)
        a = a + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    ref = np.zeros((N, N), complex)
    ref[kx % N, ky % N] = a * N * N
    assert rel_l2(st.kpsi, ref) < 1e-13


def test_rk4_uniform_stationary(oracle_mod):
    O = oracle_mod
    st, p = _state(O, 32, 1e-2)
    st.set_psi(np.ones((32, 32), complex), sync=True)
    st.rk4(5)
    assert np.abs(st.psi - 1.0).max() < 1e-14


def _vortex_psi(N, xi, seed=2):
    # smooth generic state: Padé vortex pair built in numpy (input only)
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(N) / N
    X, Y = np.meshgrid(x, x)
    psi = np.ones((N, N), complex)
    for (x0, y0, q) in [(2.0, 3.1, 1), (4.2, 3.3, -1)]:
        dx = np.angle(np.exp(1j * (X - x0)))
        dy = np.angle(np.exp(1j * (Y - y0)))
        r2 = (dx * dx + dy * dy) / xi ** 2
        rho = r2 * (11 / 32 + 11 / 384 * r2) / (1 + r2 / 3 + 11 / 384 * r2 * r2)
        psi *= np.sqrt(rho) * np.exp(1j * q * np.arctan2(np.sin(Y - y0), np.sin(X - x0)))
    return psi * (1 + 0.01 * rng.standard_normal((N, N)))


def test_rk4_conservation(oracle_mod):
    O = oracle_mod
    N, dt = 64, 1e-3
    st, p = _state(O, N, dt)
    st.set_psi(_vortex_psi(N, 5 * p.xi), sync=True)
    s0, _, _ = st.diagnostics()
    st.rk4(20)
    s1, _, _ = st.diagnostics()
    assert abs(s1[4] - s0[4]) < 1e-10 * s0[4]          # norm
    assert abs(s1[3] - s0[3]) < 1e-7 * abs(s0[3])      # etot = egrad + eintr


# ----------------------------------------------------------------- ARGLE ----

def test_argle_linear_closed_form(oracle_mod):
    O = oracle_mod
    N, dt, n = 32, 5e-2, 3
    st, p = _state(O, N, dt, g=0.0)
    rng = np.random.default_rng(5)
    K0 = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) * N   # not dealiased
    st.kpsi[...] = K0
    r = st.argle(n)
    f = (1 + p.omega * dt) / (1 + p.alpha * _kk(N) * dt)
    assert rel_l2(st.kpsi, K0 * f ** n) < 1e-13
    prev = K0 * f ** (n - 1)
    r_ref = np.linalg.norm(prev * (f - 1)) / (dt * np.linalg.norm(prev))
    assert abs(r - r_ref) < 1e-12 * r_ref


def _adv_state(O, N, dt, adv=0, vnx=0.0, vny=0.0):
    p = O.make_params(N, dt, 2.5, adv=adv, vnx=vnx, vny=vny)
    p.g = 0.0
    return O.State(p), p


def test_argle_counterflow_closed_form(oracle_mod):
    """-i v.grad psi on a plane wave e^{i(m x + n y)} is (vnx m + vny n) psi:
    with g = 0 every mode keeps its own factor
    (1 + dt (omega + vnx kx + vny ky)) / (1 + alpha k^2 dt)."""
    O = oracle_mod
    N, dt, vnx, vny = 32, 5e-2, 0.7, -0.4
    st, p = _adv_state(O, N, dt, vnx=vnx, vny=vny)
    rng = np.random.default_rng(9)
    K0 = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) * N
    st.kpsi[...] = K0
    st.argle(2)
    k = kwn_np(N).astype(float)
    f = (1 + dt * (p.omega + vnx * k[:, None] + vny * k[None, :])) / (1 + p.alpha * _kk(N) * dt)
    assert rel_l2(st.kpsi, K0 * f ** 2) < 1e-13


def test_argle_taylor_green_uniform(oracle_mod):
    """On psi = 1 the Taylor-Green term is -|u|^2/(4 alpha) with
    |u|^2 = (1 - cos 2x cos 2y)/2: one step moves only the (0,0) and
    (+-2,+-2) modes, by exact amounts."""
    O = oracle_mod
    N, dt = 32, 1e-2
    st, p = _adv_state(O, N, dt, adv=1)
    st.set_psi(np.ones((N, N), complex), sync=False)
    st.argle(1)
    K = st.kpsi
    a, n2 = p.alpha, N * N
    exp = np.zeros((N, N), complex)
    exp[0, 0] = n2 + dt * (p.omega * n2 - n2 / (8 * a))
    for kx in (2, -2):
        for ky in (2, -2):
            exp[kx % N, ky % N] = dt * n2 / (32 * a) / (1 + 8 * a * dt)
    assert np.abs(K - exp).max() < 1e-10 * n2


def test_argle_taylor_green_plane_wave(oracle_mod):
    """psi = e^{ix}: u.grad psi with the analytic gradient i e^{ix} x-hat
    (the oracle takes it spectrally); fixes the sign of -i, the velocity
    components and the 1/(4 alpha)."""
    O = oracle_mod
    N, dt = 32, 1e-2
    st, p = _adv_state(O, N, dt, adv=1)
    x = p.dx * np.arange(N)
    X, Y = x[None, :], x[:, None]
    psi = np.exp(1j * X) * np.ones((N, 1))
    st.set_psi(psi, sync=False)
    st.argle(1)
    ux, uy = np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y)
    nl = -1j * (ux * (1j * psi)) - 0.25 / p.alpha * (ux ** 2 + uy ** 2) * psi   # [y, x]
    NL = np.fft.fft2(nl).T                                                      # [kx, ky]
    P = np.fft.fft2(psi).T
    exp = (P + dt * (NL + p.omega * P)) / (1 + p.alpha * _kk(N) * dt)
    assert rel_l2(st.kpsi, exp) < 1e-13


def test_argle_uniform_fixed_point(oracle_mod):
    O = oracle_mod
    st, p = _state(O, 32, 1e-2)
    st.set_psi(np.ones((32, 32), complex), sync=False)
    r = st.argle(3)
    assert np.abs(st.kpsi[0, 0] - 32 * 32) < 1e-11 * 32 * 32
    assert r < 1e-13


def test_argle_energy_decreases(oracle_mod):
    O = oracle_mod
    N = 64
    st, p = _state(O, N, 1e-2)
    st.set_psi(_vortex_psi(N, 5 * p.xi), sync=False)
    e = []
    for _ in range(5):
        st.argle(5)
        st.ic_sync()
        e.append(st.diagnostics()[0][3] - st.diagnostics()[0][1] / 2)
    # imaginary-time relaxation lowers the grand-canonical energy
    assert all(b < a for a, b in zip(e, e[1:]))


# ----------------------------------------------------------- diagnostics ----

def test_diagnostics_plane_wave(oracle_mod):
    O = oracle_mod
    N, A = 64, 0.8
    st, p = _state(O, N, 1e-3)
    x = p.dx * np.arange(N)
    kx, ky = 4, 3
    st.set_psi(A * np.exp(1j * (kx * x[None, :] + ky * x[:, None])), sync=True)
    scal, spec, _ = st.diagnostics()
    a, g, L2 = p.alpha, p.g, p.L ** 2
    k2 = kx * kx + ky * ky
    assert abs(scal[0] - 2 * a * a * A * A * k2) < 1e-12           # egrad
    assert abs(scal[1] - 2 * a * p.omega * A * A) < 1e-12          # echem
    assert abs(scal[2] - a * g * A ** 4) < 1e-12                   # eintr
    assert abs(scal[4] - A * A * L2) < 1e-11                       # norm
    assert abs(scal[5] - 2 * a * A * A * kx) < 1e-12               # momx
    assert abs(scal[6] - 2 * a * A * A * ky) < 1e-12               # momy
    sh = round(math.sqrt(k2))                                     # shell 5
    w2 = (2 * a * A) ** 2 * k2                                    # |w|^2, w = 2 alpha A k
    assert abs(spec[0, 0] - 0.5 * w2) < 1e-12 and abs(spec[0, 1:]).max() < 1e-14
    assert abs(spec[1, 0] - 0.5 * w2) < 1e-12                      # all incompressible (k=0)
    assert abs(spec[2]).max() < 1e-14                              # no compressible part
    assert abs(spec[3]).max() < 1e-14                              # grad|psi| = 0
    assert abs(spec[4, 0] - A ** 4 * g * a) < 1e-12                # interaction
    assert abs(spec[5, sh] - A * A) < 1e-13 and abs(np.delete(spec[5], sh)).max() < 1e-14


def test_diagnostics_helmholtz(oracle_mod):
    O = oracle_mod
    N = 64
    st, p = _state(O, N, 1e-3)
    st.set_psi(_vortex_psi(N, 5 * p.xi), sync=True)
    _, spec, _ = st.diagnostics()
    # orthogonal split: |w|^2 = |w_c|^2 + |w_i|^2 shell by shell
    assert np.abs(spec[0] - spec[1] - spec[2]).max() < 1e-12 * spec[0].max()
    # Parseval of the occupation spectrum: sum |psi_k|^2/N^4 = sum |psi|^2 / N^2
    assert abs(spec[5].sum() - np.sum(np.abs(st.psi) ** 2) / N ** 2) < 1e-12 * spec[5].sum()


def test_diagnostics_pure_gradient(oracle_mod):
    O = oracle_mod
    N = 64
    st, p = _state(O, N, 1e-3)
    x = p.dx * np.arange(N)
    theta = 0.5 * np.cos(x[None, :]) + 0.3 * np.sin(2 * x[:, None])
    st.set_psi(np.exp(1j * theta), sync=True)
    _, spec, _ = st.diagnostics()
    # w = 2 alpha grad(theta): incompressible part vanishes
    assert spec[1].sum() < 1e-20 * spec[2].sum()


# -------------------------------------------------------------- particles ---

def _pstate(O, N, V0=10.0, d=2.0, np_=1, mass=5.0, dt=1e-3):
    p = O.make_params(N, dt, 2.5, np_=np_, V0_in=V0, d_in=d, mass=mass)
    return O.State(p), p


def test_particle_potential_shape(oracle_mod):
    O = oracle_mod
    N = 1024      # d = 5 dx: the x^5 term of the periodic coordinate is ~1e-8 of the integral
    st, p = _pstate(O, N)
    st.set_psi(np.ones((N, N), complex), sync=True)
    st.particle_potential()
    V = O.fft2d(st.vobj_k, +1).real
    assert abs(V[0, 0] - p.V0) < 1e-12 * p.V0                     # peak at the origin
    assert np.abs(V - V.T).max() < 1e-13 * p.V0                    # x <-> y
    assert np.abs(V[1:, 1:] - V[1:, 1:][::-1, ::-1]).max() < 1e-13 * p.V0   # x -> -x
    # Gaussian integral V0 * 2 pi d^2 (width d << L)
    assert abs(V.sum() * p.dvol - p.V0 * 2 * math.pi * p.d ** 2) < 1e-7 * p.V0 * 2 * math.pi * p.d ** 2


def test_particle_translation_grid_shift(oracle_mod):
    O = oracle_mod
    N = 64
    st, p = _pstate(O, N)
    st.set_psi(np.ones((N, N), complex), sync=True)
    st.particle_potential()
    V = O.fft2d(st.vobj_k, +1)
    i, j = 9, 23
    st.part(0)[0] = [i * p.dx, j * p.dx]
    st.translate_all()
    assert np.abs(st.trans_v - np.roll(V, (j, i), axis=(0, 1))).max() < 1e-13 * p.V0


def _analytic_gradV(p, q, N):
    # gradient of V(x - q) with the periodic coordinate of the potential
    x = p.dx * np.arange(N)
    def coord(a):
        s = np.sin(0.5 * a)
        X = 2 * (s + s ** 3 / 6 + s ** 5 / 40)
        dX = (1 + s * s / 2 + s ** 4 / 8) * np.cos(0.5 * a)
        return X, dX
    X, dX = coord(x[None, :] - q[0])
    Y, dY = coord(x[:, None] - q[1])
    V = p.V0 * np.exp(-(X * X + Y * Y) / (2 * p.d ** 2))
    return -V * X * dX / p.d ** 2, -V * Y * dY / p.d ** 2


def test_particle_force_uniform_and_independent_route(oracle_mod):
    O = oracle_mod
    N = 64
    st, p = _pstate(O, N)
    st.set_psi(np.ones((N, N), complex), sync=True)
    st.particle_potential()
    st.part(0)[0] = [1.3, 4.1]
    f = st.particle_force(0)
    assert np.abs(f).max() < 1e-12                                 # uniform fluid: no force
    psi = _vortex_psi(N, 5 * p.xi)
    st.set_psi(psi, sync=True)
    f = st.particle_force(0)
    # independent route: F_j = 2 alpha dvol sum_x rho_d * (d_j V)(x - q)
    rho = np.abs(st.psi) ** 2
    k = kwn_np(N)
    m = 9 * (k[:, None] ** 2 + k[None, :] ** 2) >= N * N
    R = np.fft.fft2(rho).T
    R[m] = 0
    rho_d = np.fft.ifft2(R.T).real
    gx, gy = _analytic_gradV(p, st.part(0)[0], N)
    fr = 2 * p.alpha * p.dvol * np.array([np.sum(rho_d * gx), np.sum(rho_d * gy)])
    assert np.abs(f - fr).max() < 1e-10 * np.abs(fr).max()


def test_shortrange_third_law(oracle_mod):
    O = oracle_mod
    st, p = _pstate(O, 32, np_=5)
    rng = np.random.default_rng(7)
    st.part(0)[...] = rng.uniform(0, 2 * np.pi, (5, 2))
    st.shortrange()
    f = np.array([st.particle_force(i) for i in range(5)])  # fluid part 0 (psi = 0 state)
    assert np.abs(f.sum(axis=0)).max() < 1e-12 * np.abs(f).max()


@pytest.mark.parametrize("fext", [0.0, 0.7])
def test_particle_free_and_accelerated_motion(oracle_mod, fext):
    O = oracle_mod
    N, dt, n, mass = 32, 1e-2, 7, 3.0
    st, p = _pstate(O, N, V0=0.0, np_=1, mass=mass, dt=dt)
    st.set_psi(np.ones((N, N), complex), sync=True)
    st.particle_potential()
    q0, v0 = np.array([1.0, 2.0]), np.array([0.3, -0.2])
    st.part(0)[0] = q0
    st.part(1)[0] = v0
    st.part(2)[0] = [fext, -fext]
    st.rk4(n)
    t = n * dt
    a = np.array([fext, -fext]) / mass
    assert np.abs(st.part(0)[0] - (q0 + v0 * t + 0.5 * a * t * t)).max() < 1e-14
    assert np.abs(st.part(1)[0] - (v0 + a * t)).max() < 1e-14


def test_ic_sync_dealiased(oracle_mod):
    O = oracle_mod
    N = 32
    st, p = _state(O, N, 1e-3)
    rng = np.random.default_rng(9)
    st.set_psi(rng.standard_normal((N, N)) + 0j, sync=True)
    k = kwn_np(N)
    m = 9 * (k[:, None] ** 2 + k[None, :] ** 2) >= N * N
    assert np.abs(st.kpsi[m]).max() == 0.0
    assert rel_l2(O.fft2d(st.kpsi, +1), st.psi) < 1e-15
