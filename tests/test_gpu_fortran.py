"""The Fortran driver (ISO_C_BINDING layer with the reference's subroutine
names) against the Python path over the same C ABI and against the oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, rel_l2

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "quantumvortices_amd", "lib", "driver_check")


def test_fortran_driver(tmp_path, oracle_mod):
    if not os.path.exists(BIN):
        pytest.fail("Fortran check driver not built (quantumvortices_amd/build.py)")
    print("case fortran driver N=64 10 steps", flush=True)
    out = tmp_path / "f.bin"
    r = subprocess.run([BIN, str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    N, ns = 64, 46
    raw = np.fromfile(out, dtype=np.uint8)
    psi = np.frombuffer(raw[:N * N * 16], dtype=np.complex128).reshape(N, N)
    kpsi = np.frombuffer(raw[N * N * 16:2 * N * N * 16], dtype=np.complex128).reshape(N, N)
    sc = np.frombuffer(raw[2 * N * N * 16:2 * N * N * 16 + 7 * 8], dtype=np.float64)
    o = 2 * N * N * 16 + 7 * 8
    spec = np.frombuffer(raw[o:o + 6 * (ns + 1) * 8], dtype=np.float64).reshape(6, ns + 1)
    o += 6 * (ns + 1) * 8
    from quantumvortices_amd import gpe as G
    p = G.params_from_namelist(N, 1e-3, 2.5)
    assert G.nshell(N) == ns
    xs, ys, qs = G.lattice_vortices(4)
    dx = 2 * np.pi / N
    psi0 = G.ic_vortices_host(N, xs + 0.31 * dx, ys + 0.17 * dx, qs, p.xi)
    g = G.GPE(p)
    g.set_psi(psi0, sync=True)
    g.gpe_rk4(10)
    d = g.diagnostics_gpe()
    # same library, same calls: identical to the bit
    assert np.array_equal(psi, g.psi) and np.array_equal(kpsi, g.kpsi)
    assert np.array_equal(spec, d["spec"])
    st = oracle_mod.State(oracle_mod.make_params(N, 1e-3, 2.5))
    st.set_psi(psi0, sync=True)
    st.rk4(10)
    assert rel_l2(psi, st.psi) < 1e-11
    scal, _, _ = st.diagnostics()
    assert abs(sc[4] - scal[4]) < 1e-11 * scal[4] and abs(sc[3] - scal[3]) < 1e-11 * abs(scal[3])
    # sts_gpe + comp_sts: 32 snapshots, one per step
    nt = 32
    sts = G.STS(g, nt)
    for _ in range(nt):
        g.gpe_rk4(1)
        sts.record()
    s2, s1 = sts.compute()
    f2 = np.frombuffer(raw[o:o + nt * sts.nshells * 8], dtype=np.float64).reshape(nt, sts.nshells)
    f1 = np.frombuffer(raw[o + nt * sts.nshells * 8:], dtype=np.float64).reshape(nt, N)
    assert np.array_equal(f2, s2) and np.array_equal(f1, s1)
    sts.close()
    g.close()
