"""ARGLE warm start from a Fourier neural operator (no reference counterpart:
the reference starts ARGLE from the Pade product ansatz of
src/initial_config.f90 and relaxes with src/gpe.f90:5-224).

Relaxation problem.  ARGLE here is the reference's plain update (no
advection, counterflow or noise term), so vortex-antivortex pairs drift
together and annihilate: its fixed point for a zero-net-charge vortex map
and a static obstacle potential is the vortex-free real ground state
psi_gs(V) of  alpha lap psi - g |psi|^2 psi - V psi + omega psi = 0.
ARGLE reaches that state quickly from the real Thomas-Fermi profile
sqrt(max(0, (omega - V)/g)) (a real field stays real, so no phase mode is
excited), and those relaxed states are the training targets.  The network
sees the vortex charge map and the potential; the warm start is ARGLE from
its prediction, the fallback is plain ARGLE from the Pade product ansatz of
the same vortex map.  Both stop at the same residual
||kpsi_new - kpsi_old|| / (dt ||kpsi_old||) <= tol and are compared by the
functional ARGLE decreases, the grand potential
Omega = egrad + eintr + eppot - echem (diagnostics_gpe's terms).

Everything on the warm-start path runs in the library: the input planes
(vk_fno_features), the network (vk_fno_*), the state upload and ARGLE
(vk_gpe_*).  torch is used only as a holder of device memory for datasets.
"""
import json
import math
import time

import numpy as np

from . import _native as nat
from . import gpe as G

VNORM_G = 20.0          # potential plane = V / (20 g)
SIGMA_XI = 2.0          # vortex blob width, units of xi


class FNO:
    """Device-resident network: lift, `layers` Fourier layers of `width`
    channels keeping modes kx in [0, modes), ky in [-modes, modes),
    projection through `proj` channels to (Re psi, Im psi).  Inputs: vortex
    charge map, potential / (20 g), Thomas-Fermi profile."""

    def __init__(self, N, width=32, modes=16, layers=4, proj=64, cin=3, batch=8):
        self.N, self.width, self.modes, self.layers, self.proj, self.cin, self.batch = \
            int(N), int(width), int(modes), int(layers), int(proj), int(cin), int(batch)
        h = nat.ctypes.c_void_p()
        nat.check(nat.lib().vk_fno_create(nat.ctypes.byref(h), self.N, self.cin, self.width, self.modes,
                                          self.layers, self.proj, self.batch))
        self.h = h
        self.nparams = int(nat.lib().vk_fno_nparams(self.h))
        offs = np.zeros(3 * self.layers + 8, dtype=np.int64)
        k = nat.lib().vk_fno_layout(self.h, offs.ctypes.data_as(nat.lp))
        names = ["Wp", "bp"] + sum([["R%d" % l, "W%d" % l, "b%d" % l] for l in range(self.layers)], []) + \
                ["Q1", "bq1", "Q2", "bq2"]
        assert k == len(names) + 1
        self.blocks = {n: (int(offs[i]), int(offs[i + 1] - offs[i])) for i, n in enumerate(names)}

    def close(self):
        if getattr(self, "h", None) and nat is not None:
            nat.lib().vk_fno_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def dims(self):
        return (self.cin, self.width, self.modes, self.layers, self.proj)

    def init_params(self, seed=0):
        """Deterministic initialisation: uniform(+-1/sqrt(fan_in)) for the
        pointwise maps, uniform(0, 1/C^2) magnitudes with random phases for
        the spectral weights, zero biases."""
        rng = np.random.default_rng(seed)
        p = np.zeros(self.nparams, dtype=np.float32)
        C = self.width

        def uni(name, fan):
            o, n = self.blocks[name]
            p[o:o + n] = rng.uniform(-1, 1, n) / math.sqrt(fan)

        uni("Wp", self.cin)
        for l in range(self.layers):
            o, n = self.blocks["R%d" % l]
            mag = rng.uniform(0, 1.0 / (C * C), n // 2)
            ph = rng.uniform(0, 2 * np.pi, n // 2)
            p[o:o + n:2] = mag * np.cos(ph)
            p[o + 1:o + n:2] = mag * np.sin(ph)
            uni("W%d" % l, C)
        uni("Q1", C)
        uni("Q2", self.proj)
        self.set_params(p)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        nat.check(nat.lib().vk_fno_set_params(self.h, nat.fptr(p)))

    def get_params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        nat.check(nat.lib().vk_fno_get_params(self.h, nat.fptr(p)))
        return p

    def grad(self):
        g = np.empty(self.nparams, dtype=np.float32)
        nat.check(nat.lib().vk_fno_get_grad(self.h, nat.fptr(g)))
        return g

    def forward(self, x):
        """x: float32 (nb, cin, N, N) numpy -> (nb, 2, N, N)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        nb = x.shape[0]
        y = np.empty((nb, 2, self.N, self.N), dtype=np.float32)
        nat.check(nat.lib().vk_fno_forward(self.h, nat.fptr(x), nat.fptr(y), nb))
        return y

    def train_step(self, x, tgt, lr, nb=None):
        """x, tgt: float32 arrays or device tensors, (nb, cin|2, N, N)."""
        nb = int(nb if nb is not None else x.shape[0])
        loss = np.zeros(1)
        nat.check(nat.lib().vk_fno_train_step(self.h, nat.fptr(x), nat.fptr(tgt), nb, float(lr), nat.ptr(loss)))
        return float(loss[0])

    def features(self, cfg, params, out=None):
        """Input planes of one configuration, built on the GPU and kept in the
        handle (also copied into `out` when given)."""
        xs, ys, qs = (np.ascontiguousarray(cfg[k], dtype=t) for k, t in
                      (("xs", np.float64), ("ys", np.float64), ("qs", np.int32)))
        pos = np.ascontiguousarray(cfg["obs"], dtype=np.float64).reshape(-1, 2)
        nv, nobs = len(xs), len(pos)
        sigma = SIGMA_XI * params.xi
        V0 = cfg["V0"] * params.g
        d = cfg["d"] * params.xi
        o = None if out is None else nat.fptr(out)
        nat.check(nat.lib().vk_fno_features(self.h, nv, nat.ptr(xs), nat.ptr(ys), qs.ctypes.data_as(nat.ip),
                                            sigma, nobs, nat.ptr(pos) if nobs else None, V0, d,
                                            VNORM_G * params.g, params.omega, params.g, o))

    def tf_guess(self, params, psi):
        """Thomas-Fermi profile of the potential plane into psi (complex128
        [y, x], host array or device tensor)."""
        nat.check(nat.lib().vk_fno_tf_guess(self.h, params.omega, params.g, VNORM_G * params.g, nat.dptr(psi)))

    residual = False    # trained on psi - TF: the prediction adds the TF plane back

    def predict_psi(self, psi, x=None, real=True):
        """Network output for the planes in the handle (or x) as complex128
        [y, x] into psi (host array or device tensor).  real: keep Re psi
        only, as the warm start does (the target ground state is real)."""
        flags = (1 if real else 0) | (2 if self.residual else 0)
        nat.check(nat.lib().vk_fno_predict_psi(self.h, None if x is None else nat.fptr(x), nat.dptr(psi), flags))

    # ---- persistence: BF16 weights + dimensions ---------------------------------
    def save(self, path, meta=None):
        p = self.get_params()
        bits = (p.view(np.uint32) >> 16).astype(np.uint16)
        info = dict(N=self.N, width=self.width, modes=self.modes, layers=self.layers, proj=self.proj,
                    cin=self.cin, **(meta or {}))
        np.savez(path, bf16=bits, info=np.array(json.dumps(info)))

    @staticmethod
    def load(path, batch=1):
        z = np.load(path, allow_pickle=False)
        info = json.loads(str(z["info"]))
        f = FNO(info["N"], info["width"], info["modes"], info["layers"], info["proj"], info["cin"], batch)
        p = (z["bf16"].astype(np.uint32) << 16).view(np.float32)
        f.set_params(p)
        f.info = info
        f.residual = bool(info.get("residual", False))
        return f


# ---- configurations ---------------------------------------------------------------

def random_config(seed, nv_range=(16, 128), nobs_range=(0, 8), V0_range=(2.0, 20.0), d_xi=2.0):
    """Seeded vortex map (even count, zero net charge) and obstacles: V0 in
    units of g, d in units of xi, positions uniform on the 2 pi box."""
    rng = np.random.default_rng(seed)
    nv = 2 * int(rng.integers(nv_range[0] // 2, nv_range[1] // 2 + 1))
    qs = np.array([1] * (nv // 2) + [-1] * (nv // 2), dtype=np.int32)
    rng.shuffle(qs)
    L = 2 * np.pi
    nobs = int(rng.integers(nobs_range[0], nobs_range[1] + 1))
    return dict(xs=rng.uniform(0, L, nv), ys=rng.uniform(0, L, nv), qs=qs,
                obs=rng.uniform(0, L, (nobs, 2)), V0=float(rng.uniform(*V0_range)), d=float(d_xi))


def solver_params(N, dt, xi=2.5, np_cap=8):
    """GPE parameters for ARGLE at N^2 (g = 1, rho = 1) with room for np_cap
    static obstacles."""
    p = G.params_from_namelist(N, dt, xi, gp_N_obj=np_cap, ppotV0=1.0, ppotd=2.0, gp_mass_part=1.0)
    return p


DT_MAX = 0.08           # ARGLE step for potentials up to ~19 g
DT_STAB = 1.5           # dt <= DT_STAB / (V_max - omega): explicit potential term stays contractive


def peak_potential(cfg, xi):
    """Largest sum of obstacle Gaussians at an obstacle centre (periodic
    minimum-image distances), in units of g: overlapping obstacles add up.
    For two obstacles closer than 2d the true peak lies between them and is
    up to ~6% higher; DT_STAB = 1.5 against the bound 2 covers that."""
    pos = np.asarray(cfg["obs"], dtype=np.float64).reshape(-1, 2)
    if len(pos) == 0:
        return 0.0
    L = 2 * np.pi
    dd = pos[:, None, :] - pos[None, :, :]
    dd -= L * np.round(dd / L)
    r2 = np.sum(dd * dd, axis=-1)
    d = cfg["d"] * xi
    return float(cfg["V0"] * np.max(np.sum(np.exp(-r2 / (2 * d * d)), axis=1)))


def step_for(cfg, params):
    """ARGLE step for a configuration: DT_MAX unless overlapping obstacles
    push the peak potential up (the fixed point does not depend on dt)."""
    vmax = peak_potential(cfg, params.xi) * params.g - params.omega
    return DT_MAX if vmax <= 0 else min(DT_MAX, DT_STAB / vmax)


def place_obstacles(g, cfg):
    p = g.params
    g.set_obstacles(cfg["obs"], cfg["V0"] * p.g, cfg["d"] * p.xi)
    dt = step_for(cfg, p)
    if dt != p.dt:
        g.set_dt(dt)


def relax(g, tol, max_steps, check_every):
    t0 = time.perf_counter()
    steps, res = g.argle_until(tol, max_steps, check_every)
    return steps, res, time.perf_counter() - t0


def ground_state(g, fno, cfg, psi_dev, tol=1e-10, max_steps=200000, check_every=25):
    """Training target: ARGLE from the Thomas-Fermi profile of the
    configuration's potential.  Leaves the relaxed state in g."""
    place_obstacles(g, cfg)
    fno.features(cfg, g.params)
    fno.tf_guess(g.params, psi_dev)
    g.set_psi_from(psi_dev, sync=True)
    return g.argle_until(tol, max_steps, check_every)


def warm_start(g, fno, cfg, psi_dev, tol=1e-10, max_steps=10 ** 6, check_every=25):
    """Warm start: features + network (real part) + upload + ARGLE; wall time
    of the whole."""
    t0 = time.perf_counter()
    place_obstacles(g, cfg)
    fno.features(cfg, g.params)
    fno.predict_psi(psi_dev)
    g.set_psi_from(psi_dev, sync=True)
    steps, res = g.argle_until(tol, max_steps, check_every)
    return steps, res, time.perf_counter() - t0


def plain_start(g, cfg, tol=1e-10, max_steps=10 ** 6, check_every=500, nimg=3):
    """Fallback: Pade product ansatz of the vortex map + ARGLE."""
    t0 = time.perf_counter()
    place_obstacles(g, cfg)
    g.set_ic_vortices(cfg["xs"], cfg["ys"], cfg["qs"], nimg=nimg, sync=True)
    steps, res = g.argle_until(tol, max_steps, check_every)
    return steps, res, time.perf_counter() - t0


def tf_start(g, fno, cfg, psi_dev, tol=1e-10, max_steps=10 ** 6, check_every=25):
    """Reference point without learning: ARGLE from the Thomas-Fermi profile."""
    t0 = time.perf_counter()
    place_obstacles(g, cfg)
    fno.features(cfg, g.params)
    fno.tf_guess(g.params, psi_dev)
    g.set_psi_from(psi_dev, sync=True)
    steps, res = g.argle_until(tol, max_steps, check_every)
    return steps, res, time.perf_counter() - t0


def grand_potential(d):
    """Omega = egrad + eintr + eppot - echem: the functional ARGLE decreases
    (src/diagnostics.f90 energy terms, same normalisation)."""
    return d["egrad"] + d["eintr"] + d.get("eppot", 0.0) - d["echem"]


def compare(fno, g, cfgs, tol=1e-10, plain_cap=100000, check_every=25, plain_check=1000, log=print):
    """Warm start vs plain ARGLE (Pade ansatz) vs Thomas-Fermi start on each
    configuration; wall times include everything from the first call
    (features, network, upload) to the step that meets the residual."""
    import torch
    N = g.N
    psi = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    pred = torch.empty_like(psi)
    out = []
    for k, cfg in enumerate(cfgs):
        row = dict(nv=len(cfg["xs"]), nobs=len(cfg["obs"]), V0=cfg["V0"])
        n, r, t = warm_start(g, fno, cfg, psi, tol, check_every=check_every)
        d = g.diagnostics_gpe()
        row.update(warm_steps=n, warm_res=r, warm_sec=t, warm_omega=grand_potential(d), warm_etot=d["etot"])
        g.psi_to(psi)
        fno.predict_psi(pred, real=False)
        row["raw_rel_err"] = float(torch.linalg.norm(pred - psi) / torch.linalg.norm(psi))
        fno.predict_psi(pred)
        row["raw_rel_err_real"] = float(torch.linalg.norm(pred - psi) / torch.linalg.norm(psi))
        torch.cuda.synchronize()
        n, r, t = tf_start(g, fno, cfg, psi, tol, check_every=check_every)
        d = g.diagnostics_gpe()
        row.update(tf_steps=n, tf_res=r, tf_sec=t, tf_omega=grand_potential(d))
        g.psi_to(pred)
        fno.tf_guess(g.params, psi)
        row["tf_raw_rel_err"] = float(torch.linalg.norm(psi - pred) / torch.linalg.norm(pred))
        torch.cuda.synchronize()
        n, r, t = plain_start(g, cfg, tol, max_steps=plain_cap, check_every=plain_check)
        d = g.diagnostics_gpe()
        row.update(plain_steps=n, plain_res=r, plain_sec=t, plain_omega=grand_potential(d),
                   plain_converged=bool(r <= tol))
        log(json.dumps({"config": k, **{a: (round(b, 6) if isinstance(b, float) else b) for a, b in row.items()}}))
        out.append(row)
    return out
