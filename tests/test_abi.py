"""CPU checks of the C ABI (no compute calls without a GPU): the library
loads, exports every function include/vikshobha.h declares, derives the
reference's constants bit-for-bit like the oracle, and fails loudly (no CPU
fallback) when asked to compute without a ROCm device."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "vikshobha.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vk_\w+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    from quantumvortices_amd import _native as nat
    L = nat.lib()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
        assert n in nat.SIGNATURES, "binding missing for " + n


@pytest.mark.parametrize("N", [32, 128, 1024, 4096])
def test_constants_match_oracle(oracle_mod, N):
    from quantumvortices_amd import gpe as G
    p = G.params_from_namelist(N, 2.5e-4, 2.5)
    o = oracle_mod.make_params(N, 2.5e-4, 2.5)
    assert p.alpha == o.alpha and p.g == o.g and p.omega == o.omega and p.xi == o.xi
    assert G.nshell(N) == o.nshell
    pp = G.params_from_namelist(N, 2.5e-4, 2.5, gp_N_obj=3, ppotV0=10.0, ppotd=2.0, gp_mass_part=5.0)
    op = oracle_mod.make_params(N, 2.5e-4, 2.5, np_=3, V0_in=10.0, d_in=2.0, mass=5.0)
    assert pp.V0 == op.V0 and pp.d == op.d and pp.mass == op.mass


def test_ic_host_vortices():
    from quantumvortices_amd import gpe as G
    N = 128
    p = G.params_from_namelist(N, 1e-3, 2.5)
    xs, ys, qs = G.lattice_vortices(2)
    dx = 2 * np.pi / N
    psi = G.ic_vortices_host(N, xs + 0.3 * dx, ys + 0.2 * dx, qs, p.xi, nimg=3)
    a = np.abs(psi)
    # density 1 far from the cores, a zero near each core, unit winding
    assert abs(a.max() - 1.0) < 0.05
    ix, iy = int(round(xs[0] / dx)), int(round(ys[0] / dx))
    assert a[iy, ix] < 0.2
    loop = [psi[iy - 4, ix - 4 + i] for i in range(8)] + [psi[iy - 4 + i, ix + 4] for i in range(8)] + \
           [psi[iy + 4, ix + 4 - i] for i in range(8)] + [psi[iy + 4 - i, ix - 4] for i in range(8)]
    ph = np.angle(np.array(loop + loop[:1]))
    wind = np.sum(np.angle(np.exp(1j * np.diff(ph)))) / (2 * np.pi)
    assert abs(abs(wind) - 1.0) < 1e-9
    # periodic: the field continues smoothly across the box edges
    assert np.abs(psi[0, :] - psi[-1, :]).max() < 0.1
    assert np.abs(psi[:, 0] - psi[:, -1]).max() < 0.1


def test_compute_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from quantumvortices_amd import gpe as G
    from quantumvortices_amd._native import VKError
    with pytest.raises(VKError, match="no ROCm device"):
        G.GPE(G.params_from_namelist(64, 1e-3, 2.5))


def test_bench_rejects_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
