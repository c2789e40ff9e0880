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
    r = subprocess.run([BIN, str(out), str(tmp_path)], capture_output=True, text=True, timeout=120)
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
    o += nt * sts.nshells * 8
    f1 = np.frombuffer(raw[o:o + nt * N * 8], dtype=np.float64).reshape(nt, N)
    o += nt * N * 8
    assert np.array_equal(f2, s2) and np.array_equal(f1, s1)
    # fft() work-array plans (src/mod_fft.f90:489-540): forward to the
    # transposed Fourier layout, unnormalised; inverse scaled by 1/N^2
    fb = N * N * 16
    cplx = lambda k: np.frombuffer(raw[o + k * fb:o + (k + 1) * fb], dtype=np.complex128).reshape(N, N)
    x, fnl, rt, fcm = cplx(0), cplx(1), cplx(2), cplx(3)
    assert o + 4 * fb == raw.size
    ref = np.fft.fft2(x).T                                # [kx, ky] rows: ky fastest
    assert rel_l2(fnl, ref) < 1e-13
    assert rel_l2(rt, x) < 1e-14
    assert rel_l2(fcm, np.fft.fft2(x * x).T) < 1e-13
    sts.close()
    g.close()


PART = os.path.join(ROOT, "quantumvortices_amd", "lib", "driver_part")


def _read_table(path):
    return np.array([[float(v) for v in line.split()] for line in open(path) if line.strip()])


def test_fortran_particle_driver(tmp_path, oracle_mod):
    """The reference's particle run through the Fortran layer
    (src/driver.f90:460-508, src/solve.f90:105-136): argument-less
    initialize_fft on the namelist globals, restart_gpe, particle_potential
    and initialize_particles from restart/part files, 20 RK4 steps with
    diagnostics_gpe every 5, gppart_checkpoint.  The final state matches the
    oracle; the text files have the reference's names, formats and values."""
    if not os.path.exists(PART):
        pytest.fail("Fortran particle driver not built (quantumvortices_amd/build.py)")
    print("case fortran particle driver N=64 np=3 20 steps", flush=True)
    from quantumvortices_amd import gpe as G
    N, npart, nrst = 64, 3, 7
    p = G.params_from_namelist(N, 1e-3, 2.5)
    xs, ys, qs = G.lattice_vortices(4)
    dx = 2 * np.pi / N
    psi0 = G.ic_vortices_host(N, xs + 0.31 * dx, ys + 0.17 * dx, qs, p.xi)
    rng = np.random.default_rng(11)
    pos = rng.uniform(0, 2 * np.pi, (npart, 2))
    vel = rng.uniform(-0.1, 0.1, (npart, 2))
    fext = rng.uniform(-0.1, 0.1, (npart, 2))
    for d in ("restart/wf", "restart/part", "data", "spectra"):
        (tmp_path / d).mkdir(parents=True)
    G.write_wf(str(tmp_path / "restart/wf"), 0, 0.0, 1e-3, psi0)
    # the reference's restart files, list-directed 'iopoint ptime a b' lines
    for name, a in (("gppos", pos), ("gpvel", vel), ("gpf0ext", fext)):
        with open(tmp_path / "restart/part" / ("%s.dat_nrst%d" % (name, nrst)), "w") as f:
            for i in range(npart):
                f.write("  0  0.0000000000000000  %.17E  %.17E\n" % (a[i, 0], a[i, 1]))
    r = subprocess.run([PART, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(tmp_path / "final.bin", dtype=np.uint8)
    nb = N * N * 16
    psi = np.frombuffer(raw[:nb], dtype=np.complex128).reshape(N, N)
    rest = np.frombuffer(raw[2 * nb:], dtype=np.float64)
    fpos, fvel, ffrc = (rest[i * 2 * npart:(i + 1) * 2 * npart].reshape(npart, 2) for i in range(3))
    fps = rest[6 * npart:]
    O = oracle_mod
    st = O.State(O.make_params(N, 1e-3, 2.5, np_=npart, V0_in=10.0, d_in=2.0, mass=5.0))
    st.set_psi(psi0, sync=True)
    st.part(0)[...] = pos
    st.part(1)[...] = vel
    st.part(2)[...] = fext
    st.particle_potential()
    diag = []
    for blk in range(4):
        st.rk4(5)
        scal, spec, pscal = st.diagnostics()
        diag.append((5e-3 * (blk + 1), scal.copy(), spec.copy(), pscal.copy(), st.part(0).copy(),
                     st.part(1).copy(), st.part(3).copy()))
    assert rel_l2(psi, st.psi) < 1e-11
    assert rel_l2(fpos, st.part(0)) < 1e-11 and rel_l2(fvel, st.part(1)) < 1e-10
    assert rel_l2(ffrc, st.part(3)) < 1e-10
    ps = diag[-1][3]
    assert abs(fps[0] - ps[0]) <= 1e-11 * abs(ps[0]) and abs(fps[1] - ps[1]) <= 1e-10 * abs(ps[1])
    # text outputs: E15.6 values, one line per diagnostics call
    close = lambda a, b: np.all(np.abs(a - b) <= 6e-6 * np.maximum(np.abs(b), 1e-30) + 1e-300)
    data = tmp_path / "data"
    e1 = _read_table(data / "energy1.dat_nrst7")
    nrm = _read_table(data / "norm.dat_nrst7")
    mom = _read_table(data / "mom.dat_nrst7")
    en = _read_table(data / "energy.dat_nrst7")
    n0 = _read_table(data / "n0.dat_nrst7")
    eg = _read_table(data / "egpobj.dat_nrst7")
    mg = _read_table(data / "momgpobj.dat_nrst7")
    assert e1.shape == (4, 5) and nrm.shape == (4, 2) and mom.shape == (4, 3) and en.shape == (4, 6)
    for i, (t, scal, spec, pscal, ppos, pvel, pfrc) in enumerate(diag):
        assert close(e1[i], np.array([t, scal[0], scal[1], scal[2], scal[3]]))
        assert close(nrm[i], np.array([t, scal[4]]))
        assert abs(mom[i, 0] - t) < 1e-9 and np.all(np.abs(mom[i, 1:] - scal[5:7]) <= 1e-6 * abs(scal[0]))
        ns = spec.shape[1] - 1
        tot = [spec[0, 1:].sum(), spec[1, 1:].sum(), spec[2, 1:].sum(), spec[3, 1:].sum(), spec[4, :].sum()]
        assert close(en[i], np.array([t] + tot))
        assert close(n0[i], np.array([t, spec[5, 0]]))
        assert close(eg[i], np.array([t, pscal[0], pscal[1], pscal[2]]))
        assert np.all(np.abs(mg[i, 1:] - pscal[3:5]) <= 6e-6 * np.abs(pscal[3:5]).max())
        for k in range(npart):
            pp = _read_table(data / ("gpobj_pos_%d.dat_nrst7" % (k + 1)))
            ff = _read_table(data / ("gpobj_frc_%d.dat_nrst7" % (k + 1)))
            assert close(pp[i, 1:], ppos[k]) and np.all(np.abs(ff[i, 1:] - pfrc[k]) <= 6e-6 * np.abs(pfrc).max())
    # spectra/spec<spoint>.dat at jtime = 10, 20: shells 1..nshell
    for spoint, blk in ((1, 1), (2, 3)):
        sp = _read_table(tmp_path / "spectra" / ("spec%d.dat" % spoint))
        spec = diag[blk][2]
        assert sp.shape == (spec.shape[1] - 1, 7)
        assert np.array_equal(sp[:, 0], np.arange(1, spec.shape[1]))
        for c in range(6):
            assert np.all(np.abs(sp[:, 1 + c] - spec[c, 1:]) <= 6e-6 * np.abs(spec[c, 1:]) + 1e-300)
    # gppart_checkpoint: 17 significant digits, restart-exact
    ck = _read_table(data / "gppos_ckp.dat_nrst7")
    assert ck.shape == (npart, 4) and rel_l2(ck[:, 2:], st.part(0)) < 1e-11
    ckv = _read_table(data / "gpvel_ckp.dat_nrst7")
    assert rel_l2(ckv[:, 2:], st.part(1)) < 1e-10
    assert np.array_equal(_read_table(data / "gpf0ext_ckp.dat_nrst7")[:, 2:], fext)
