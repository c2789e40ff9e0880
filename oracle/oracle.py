"""ctypes view of the CPU oracle (oracle/vk_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the reference results the tests
compare against (or the timed CPU baseline), never as the product path.

Arrays follow the reference's Fortran views written as C arrays:
x-space psi[y, x] (x fastest) and k-space kpsi[kx, ky] (ky fastest, the
FFTW_MPI_TRANSPOSED_OUT layout of src/mod_fft.f90:143-146).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libvkoracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int),
        ("L", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("dvol", ctypes.c_double),
        ("kalias", ctypes.c_double),
        ("nshell", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("xi", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("g", ctypes.c_double),
        ("omega", ctypes.c_double),
        ("disp", ctypes.c_double),
        ("np", ctypes.c_int),
        ("V0", ctypes.c_double),
        ("d", ctypes.c_double),
        ("mass", ctypes.c_double),
        ("force_dealias", ctypes.c_int),
        ("adv", ctypes.c_int),
        ("vnx", ctypes.c_double),
        ("vny", ctypes.c_double),
        ("sgle", ctypes.c_int),
        ("oneybeta", ctypes.c_double),
        ("nuN", ctypes.c_double),
        ("ktrunc", ctypes.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.POINTER
        dp = P(ctypes.c_double)
        L.orc_set_grid.argtypes = [ctypes.c_int, P(OrcParams)]
        L.orc_constants.argtypes = [ctypes.c_double] * 4 + [P(OrcParams)]
        L.orc_kwn.argtypes = [ctypes.c_int, P(ctypes.c_int)]
        L.orc_mask.argtypes = [P(OrcParams), ctypes.c_int, ctypes.c_int]
        L.orc_shell.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_twiddles.argtypes = [ctypes.c_int, dp]
        L.orc_fft2d.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_create.argtypes = [P(OrcParams)]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for f in ("orc_psi", "orc_kpsi", "orc_vobj_k", "orc_trans_v"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = dp
        L.orc_part.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_part.restype = dp
        for f in ("orc_sync_psi", "orc_ic_sync", "orc_rk4_step",
                  "orc_particle_potential", "orc_translate_all", "orc_shortrange"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.orc_argle_step.argtypes = [ctypes.c_void_p, dp]
        L.orc_rhs_only.argtypes = [ctypes.c_void_p, dp]
        L.orc_diagnostics.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.orc_particle_force.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.orc_set_noise.argtypes = [ctypes.c_void_p, dp]
        L.orc_set_velocity.argtypes = [ctypes.c_void_p, dp, dp]
        L.orc_omega.argtypes = [ctypes.c_void_p]
        L.orc_omega.restype = ctypes.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n):
    lib().orc_set_threads(int(n))


def make_params(N, dt, xi_in=2.5, rho_avg=1.0, disp=0.0, csound=None,
                np_=0, V0_in=0.0, d_in=0.0, mass=1.0, force_dealias=1, adv=0, vnx=0.0, vny=0.0,
                sgle=0, oneybeta=0.0, nuN=0.0, chempot0=None, ktrunc=0.0):
    """Grid + constants as the reference derives them (src/mod_grid.f90:63-101,
    src/driver.f90:360-393).  Default: g=1 with xi = xi_in*dx, i.e.
    csound = sqrt(2)*xi so that alpha = xi^2."""
    p = OrcParams()
    lib().orc_set_grid(N, ctypes.byref(p))
    if csound is None:
        csound = np.sqrt(2.0) * xi_in * p.dx
    lib().orc_constants(xi_in, csound, rho_avg, disp, ctypes.byref(p))
    p.dt = dt
    p.np = np_
    p.V0 = V0_in * p.g        # ppotV0 = ppotV0*gstrength
    p.d = d_in * p.xi         # ppotd = ppotd*xi
    p.mass = mass
    p.force_dealias = force_dealias
    p.adv = adv
    p.vnx = vnx
    p.vny = vny
    p.sgle = sgle
    p.oneybeta = oneybeta
    p.nuN = nuN
    p.ktrunc = ktrunc
    if sgle and chempot0 is not None:
        p.omega = chempot0      # omega_chem = chempot0 under GPE_SGLE (src/driver.f90:373)
    return p


def kwn(N):
    k = np.zeros(N, dtype=np.int32)
    lib().orc_kwn(N, k.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return k


def mask(p, k1, k2):
    return lib().orc_mask(ctypes.byref(p), int(k1), int(k2))


def shell(k1, k2):
    return lib().orc_shell(int(k1), int(k2))


def twiddles(N):
    tw = np.zeros(N, dtype=np.complex128)
    lib().orc_twiddles(N, _dp(tw))
    return tw


def fft2d(a, direction):
    """direction -1: x-space [y,x] -> k-space [kx,ky]; +1: inverse, /N^2."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    out = np.empty_like(a)
    lib().orc_fft2d(a.shape[0], _dp(a), _dp(out), int(direction))
    return out


class State:
    """One oracle state (the reference's module globals for one run)."""

    def __init__(self, params):
        self.p = params
        self.N = params.N
        self.h = lib().orc_create(ctypes.byref(params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def _view(self, fn):
        ptr = getattr(lib(), fn)(self.h)
        buf = np.ctypeslib.as_array(ptr, shape=(self.N * self.N * 2,))
        return buf.view(np.complex128).reshape(self.N, self.N)

    @property
    def psi(self):
        return self._view("orc_psi")

    @property
    def kpsi(self):
        return self._view("orc_kpsi")

    @property
    def vobj_k(self):
        return self._view("orc_vobj_k")

    @property
    def trans_v(self):
        return self._view("orc_trans_v")

    def part(self, which):
        n = max(self.p.np, 1)
        ptr = lib().orc_part(self.h, which)
        return np.ctypeslib.as_array(ptr, shape=(n, 2))

    def set_psi(self, psi, sync=True):
        self.psi[...] = psi
        lib().orc_sync_psi(self.h)
        if sync:
            lib().orc_ic_sync(self.h)

    def ic_sync(self):
        lib().orc_ic_sync(self.h)

    def rk4(self, n=1):
        for _ in range(n):
            lib().orc_rk4_step(self.h)

    def argle(self, n=1):
        r = np.zeros(1)
        for _ in range(n):
            lib().orc_argle_step(self.h, _dp(r))
        return float(r[0])

    def rhs(self):
        out = np.zeros((self.N, self.N), dtype=np.complex128)
        lib().orc_rhs_only(self.h, _dp(out))
        return out

    def diagnostics(self):
        ns = self.p.nshell
        scal = np.zeros(8)
        spec = np.zeros(6 * (ns + 1))
        pscal = np.zeros(5)
        lib().orc_diagnostics(self.h, _dp(scal), _dp(spec), _dp(pscal))
        return scal, spec.reshape(6, ns + 1), pscal

    def particle_potential(self):
        lib().orc_particle_potential(self.h)

    def translate_all(self):
        lib().orc_translate_all(self.h)

    def particle_force(self, ip):
        f = np.zeros(2)
        lib().orc_particle_force(self.h, int(ip), _dp(f))
        return f

    def shortrange(self):
        lib().orc_shortrange(self.h)

    def set_noise(self, noise):
        a = np.ascontiguousarray(noise, dtype=np.complex128)
        assert a.shape == (self.N, self.N)
        lib().orc_set_noise(self.h, _dp(a))

    def set_velocity(self, ux, uy):
        ux = np.ascontiguousarray(ux, dtype=np.float64)
        uy = np.ascontiguousarray(uy, dtype=np.float64)
        lib().orc_set_velocity(self.h, _dp(ux), _dp(uy))

    @property
    def omega(self):
        return lib().orc_omega(self.h)
