"""Host-side mirror of the reference's GPE module interface, over the C ABI.

The reference (Fortran) drives its step through argument-less subroutines on
module globals: gpe_rk4 (src/gpe.f90:310), ARGLE(time) (src/gpe.f90:5),
diagnostics_gpe(spoint,jtime,time) (src/diagnostics.f90:4),
fft(plan_name,direction) (src/mod_fft.f90:471), particle_potential /
initialize_particles (src/gpe_particle.f90:3,92).  Here one `GPE` object
holds the device-resident state (psi, kpsi, particles) and exposes the same
operations by the same names.  All compute runs in the HIP library; there is
no CPU path.
"""
import math

import numpy as np

from . import _native as nat

forward, inverse = -1, 1   # src/mod_fft.f90:48


def params_from_namelist(Nx, dt_fix, xi, csound=None, rho_avg=1.0, disp_factor=0.0,
                         gp_N_obj=0, ppotV0=0.0, ppotd=0.0, gp_mass_part=1.0,
                         force_dealias=1):
    """Build vk_gpe_params from the reference's namelist inputs
    (src/driver.f90:52-78): xi in units of dx, csound absolute, ppotV0 in
    units of g, ppotd in units of xi.  Derived constants follow
    src/driver.f90:360-393.  csound=None selects g = 1 (csound = sqrt(2)*xi)."""
    p = nat.Params()
    if csound is None:
        csound = math.sqrt(2.0) * xi * (2.0 * math.pi / Nx)
    nat.check(nat.lib().vk_gpe_params_init(p, int(Nx), float(dt_fix), float(xi), float(csound),
                                           float(rho_avg), float(disp_factor)))
    p.np = int(gp_N_obj)
    p.V0 = float(ppotV0) * p.g
    p.d = float(ppotd) * p.xi
    p.mass = float(gp_mass_part)
    p.force_dealias = int(force_dealias)
    return p


def nshell(N):
    return nat.lib().vk_nshell(int(N))


class GPE:
    """One device-resident GPE state (the reference's module globals)."""

    def __init__(self, params):
        self.params = params
        self.N = params.N
        self.nshell = nshell(self.N)
        h = nat.ctypes.c_void_p()
        nat.check(nat.lib().vk_gpe_create(nat.ctypes.byref(h), nat.ctypes.byref(params)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and nat is not None:
            nat.lib().vk_gpe_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- fields ------------------------------------------------------------
    def set_psi(self, psi, sync=True):
        """psi[y, x] complex128.  kpsi = FFT(psi); sync: dealias kpsi and
        psi = IFFT(kpsi) (restart_from_argle_data + ic_otoc_gpe)."""
        a = np.ascontiguousarray(psi, dtype=np.complex128)
        assert a.shape == (self.N, self.N)
        nat.check(nat.lib().vk_gpe_set_psi(self.h, nat.ptr(a), 1 if sync else 0))

    def set_kpsi(self, kpsi):
        a = np.ascontiguousarray(kpsi, dtype=np.complex128)
        assert a.shape == (self.N, self.N)
        nat.check(nat.lib().vk_gpe_set_kpsi(self.h, nat.ptr(a)))

    def set_ic_vortices(self, xs, ys, qs, nimg=3, sync=True):
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        ys = np.ascontiguousarray(ys, dtype=np.float64)
        qs = np.ascontiguousarray(qs, dtype=np.int32)
        nat.check(nat.lib().vk_gpe_set_ic_vortices(self.h, len(xs), nat.ptr(xs), nat.ptr(ys),
                                                   qs.ctypes.data_as(nat.ip), int(nimg), 1 if sync else 0))

    @property
    def psi(self):
        out = np.empty((self.N, self.N), dtype=np.complex128)
        nat.check(nat.lib().vk_gpe_get_psi(self.h, nat.ptr(out), None))
        return out

    @property
    def kpsi(self):
        out = np.empty((self.N, self.N), dtype=np.complex128)
        nat.check(nat.lib().vk_gpe_get_psi(self.h, None, nat.ptr(out)))
        return out

    # ---- the reference's operations ------------------------------------------
    def gpe_rk4(self, nsteps=1):
        """nsteps x gpe_rk4 (src/gpe.f90:310-494)."""
        nat.check(nat.lib().vk_gpe_rk4_steps(self.h, int(nsteps)))

    def ARGLE(self, time=None, nsteps=1):
        """nsteps x ARGLE(time) (src/gpe.f90:5-224); returns the residual
        ||kpsi_new - kpsi_old|| / (dt ||kpsi_old||) of the last step."""
        r = np.zeros(1)
        nat.check(nat.lib().vk_gpe_argle_steps(self.h, int(nsteps), nat.ptr(r)))
        return float(r[0])

    def argle_until(self, tol=1e-10, max_steps=10 ** 6, check_every=500):
        """ARGLE until residual <= tol; returns (steps, residual)."""
        n = nat.ctypes.c_int()
        r = np.zeros(1)
        nat.check(nat.lib().vk_gpe_argle_until(self.h, float(tol), int(max_steps), int(check_every),
                                               nat.ctypes.byref(n), nat.ptr(r)))
        return n.value, float(r[0])

    def diagnostics_gpe(self, spoint=0, jtime=0, time=0.0):
        """diagnostics_gpe (src/diagnostics.f90:4-571): returns a dict with
        egrad, echem, eintr, etot, norm, momx, momy, n0, spec (6, nshell+1)
        and, with particles, eppot, eobj_kin, esrep, momobj, frc."""
        scal = np.zeros(8)
        spec = np.zeros(6 * (self.nshell + 1))
        pscal = np.zeros(5)
        nat.check(nat.lib().vk_gpe_diagnostics(self.h, nat.ptr(scal), nat.ptr(spec), nat.ptr(pscal)))
        d = dict(zip(["egrad", "echem", "eintr", "etot", "norm", "momx", "momy", "n0"], scal.tolist()))
        d["spec"] = spec.reshape(6, self.nshell + 1)
        d["scal"] = scal
        if self.params.np > 0:
            d.update(eppot=pscal[0], eobj_kin=pscal[1], esrep=pscal[2], momobj=pscal[3:5].copy())
            d["frc"] = self.particles()[2]
        return d

    def fft(self, a, direction):
        """fft(plan_name, direction) conformance hook (src/mod_fft.f90:471-542)
        on an explicit array: forward x-space [y,x] -> k-space [kx,ky];
        inverse k-space -> x-space times 1/(Nx*Ny)."""
        a = np.ascontiguousarray(a, dtype=np.complex128)
        out = np.empty_like(a)
        nat.check(nat.lib().vk_fft2d(self.h, int(direction), nat.ptr(a), nat.ptr(out)))
        return out

    # ---- particles -----------------------------------------------------------
    def initialize_particles(self, pos, vel=None, fext=None):
        """particle_potential + initialize_particles (src/gpe_particle.f90:3-141);
        pos/vel/fext are (np, 2) arrays (x, y per particle)."""
        n = self.params.np
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 2)
        vel = np.zeros((n, 2)) if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 2)
        fext = np.zeros((n, 2)) if fext is None else np.ascontiguousarray(fext, dtype=np.float64).reshape(n, 2)
        self._fext = fext.copy()
        nat.check(nat.lib().vk_gpe_particles_set(self.h, nat.ptr(pos), nat.ptr(vel), nat.ptr(fext)))

    def set_obstacles(self, pos, V0, d):
        """Static obstacles for ARGLE: np = len(pos) particles (at most the np
        the handle was created with) at pos (np, 2), amplitude V0 and width d
        (absolute units), velocities cleared."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 2)
        n = pos.shape[0]
        nat.check(nat.lib().vk_gpe_set_obstacles(self.h, n, nat.ptr(pos) if n else None, float(V0), float(d)))
        self.params.np = n
        self.params.V0 = float(V0)
        self.params.d = float(d)

    def set_advection(self, taylor_green=False, vnx=0.0, vny=0.0):
        """ARGLE's advective terms: the Taylor-Green velocity (GPE_ADVECTION)
        and/or a constant counterflow (GPE_COUNTERFLOW, vnx, vny)."""
        nat.check(nat.lib().vk_gpe_set_advection(self.h, 1 if taylor_green else 0, float(vnx), float(vny)))

    def set_advection_field(self, ux=None, uy=None):
        """ARGLE advection by a supplied velocity field (x-space (N, N) real
        arrays, host or device) in place of the Taylor-Green one; None
        turns it off."""
        if ux is None:
            nat.check(nat.lib().vk_gpe_set_advection_field(self.h, None, None))
            return
        if not hasattr(ux, "data_ptr"):
            ux = np.ascontiguousarray(ux, dtype=np.float64)
            uy = np.ascontiguousarray(uy, dtype=np.float64)
        nat.check(nat.lib().vk_gpe_set_advection_field(self.h, nat.dptr(ux), nat.dptr(uy)))

    def set_vortex_advection(self, xs, ys, qs, nimg=3):
        """Imposed velocity of a vortex configuration (the current of its
        Padé product ansatz) as ARGLE's advection field."""
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        ys = np.ascontiguousarray(ys, dtype=np.float64)
        qs = np.ascontiguousarray(qs, dtype=np.int32)
        nat.check(nat.lib().vk_gpe_set_vortex_advection(self.h, len(xs), nat.ptr(xs), nat.ptr(ys),
                                                        qs.ctypes.data_as(nat.ip), int(nimg)))

    def set_sgle(self, on=True, oneybeta=0.0, nuN=0.0, chempot0=None, ktrunc=0.0, seed=0):
        """SGLE (GPE_SGLE build of ARGLE): noise strength oneybeta,
        chemical-potential rate nuN from chempot0 (default: the handle's
        omega), ktrunc > 0 truncation, generator seed."""
        c0 = self.params.omega if chempot0 is None else chempot0
        nat.check(nat.lib().vk_gpe_set_sgle(self.h, 1 if on else 0, float(oneybeta), float(nuN), float(c0),
                                            float(ktrunc), int(seed)))

    def sgle_step(self, noise):
        """One SGLE step with the caller's complex standard normals
        noise[y, x]; returns the residual."""
        a = noise if hasattr(noise, "data_ptr") else np.ascontiguousarray(noise, dtype=np.complex128)
        r = np.zeros(1)
        nat.check(nat.lib().vk_gpe_argle_step_noise(self.h, nat.dptr(a), nat.ptr(r)))
        return float(r[0])

    def sgle_noise(self, draw):
        """Draw `draw` of the SGLE generator (unscaled complex normals)."""
        out = np.empty((self.N, self.N), dtype=np.complex128)
        nat.check(nat.lib().vk_sgle_noise(self.h, int(draw), nat.ptr(out)))
        return out

    @property
    def chempot(self):
        o = np.zeros(1)
        nat.check(nat.lib().vk_gpe_get_chempot(self.h, nat.ptr(o)))
        return float(o[0])

    def set_dt(self, dt):
        """Replace dt_fix of the live state."""
        nat.check(nat.lib().vk_gpe_set_dt(self.h, float(dt)))
        self.params.dt = float(dt)

    def set_psi_from(self, psi, sync=True):
        """set_psi from a complex128 [y, x] array that may live on the GPU
        (anything with data_ptr(), e.g. a torch tensor) or on the host."""
        nat.check(nat.lib().vk_gpe_set_psi(self.h, nat.dptr(psi), 1 if sync else 0))

    def psi_to(self, psi):
        """Current psi as complex128 [y, x] into a host array or a device
        tensor (anything with data_ptr())."""
        nat.check(nat.lib().vk_gpe_get_psi(self.h, nat.dptr(psi), None))

    def particles(self):
        n = self.params.np
        pos, vel, frc = np.zeros((n, 2)), np.zeros((n, 2)), np.zeros((n, 2))
        nat.check(nat.lib().vk_gpe_particles_get(self.h, nat.ptr(pos), nat.ptr(vel), nat.ptr(frc)))
        return pos, vel, frc

    # ---- checkpoints (the reference's file formats) ------------------------------
    def io_fields(self, directory, iopoint, time, nsplit=1):
        """io_fields (src/io_fields.f90:1-62): psi to wf<iopoint>p<r>.dat,
        r = 1..nsplit, in <directory>."""
        write_wf(directory, iopoint, time, self.params.dt, self.psi, nsplit)

    def restart_gpe(self, directory, nrestart, nsplit_orig=1, sync=False):
        """restart_gpe (src/io_fields.f90:71-224): read checkpoint nrestart
        (nsplit_orig files) and set psi (kpsi = FFT(psi); sync also dealiases
        and re-synchronises psi, as restart_from_argle_data + ic_otoc_gpe).
        Returns (iopoint, time, dt) from the file header."""
        icp, time, dt, psi = read_wf(directory, nrestart, self.N, nsplit_orig)
        self.set_psi(psi, sync=sync)
        return icp, time, dt

    def gppart_checkpoint(self, directory, nrst, iopoint, ptime):
        """gppart_checkpoint (src/gpe_particle.f90:144-184): append positions,
        velocities and external forces to the *_ckp.dat_nrst<nrst> files."""
        n = self.params.np
        pos, vel, _ = self.particles()
        fext = np.ascontiguousarray(self._fext if getattr(self, "_fext", None) is not None else np.zeros((n, 2)))
        nat.check(nat.lib().vk_part_checkpoint(str(directory).encode(), int(nrst), int(iopoint), float(ptime), n,
                                               nat.ptr(np.ascontiguousarray(pos)), nat.ptr(np.ascontiguousarray(vel)),
                                               nat.ptr(fext)))

    def initialize_particles_from(self, directory, nrst):
        """initialize_particles (src/gpe_particle.f90:92-141) from
        gppos/gpvel/gpf0ext.dat_nrst<nrst> in <directory>."""
        n = self.params.np
        pos, vel, fext = np.zeros((n, 2)), np.zeros((n, 2)), np.zeros((n, 2))
        nat.check(nat.lib().vk_part_restart(str(directory).encode(), int(nrst), n, nat.ptr(pos), nat.ptr(vel),
                                            nat.ptr(fext)))
        self.initialize_particles(pos, vel, fext)
        return pos, vel, fext

    # ---- timing ----------------------------------------------------------------
    def rk4_async(self, nsteps):
        nat.check(nat.lib().vk_gpe_rk4_steps_async(self.h, int(nsteps)))

    @property
    def stream(self):
        return nat.lib().vk_gpe_stream(self.h)

    def pass_times(self, nsteps):
        k = nat.lib().vk_pass_kinds()
        ms = np.zeros(k)
        nat.check(nat.lib().vk_gpe_pass_times(self.h, int(nsteps), nat.ptr(ms), k))
        names = [nat.lib().vk_pass_name(i).decode() for i in range(k)]
        return dict(zip(names, ms.tolist()))

    def set_shape(self, var):
        """Transform shape of the RK4 step of a field of N in {256, 512,
        1024} without particles: 0 (default) or 1 (8 points per thread)."""
        nat.check(nat.lib().vk_gpe_set_shape(self.h, int(var)))

    def shape(self):
        v = nat.ctypes.c_int()
        nat.check(nat.lib().vk_gpe_get_shape(self.h, nat.ctypes.byref(v)))
        return v.value

    def set_row_variants(self, p1, p3):
        """Row-pass variants of the psi-row path (N >= 2048): p1 0 = TR rows
        per workgroup, 1 = one row, 2 = the prefetching persistent P1;
        p3 0 = TR rows, 1 = one row."""
        nat.check(nat.lib().vk_gpe_set_row_variants(self.h, int(p1), int(p3)))

    def row_variants(self):
        a, b = nat.ctypes.c_int(), nat.ctypes.c_int()
        nat.check(nat.lib().vk_gpe_get_row_variants(self.h, nat.ctypes.byref(a), nat.ctypes.byref(b)))
        return a.value, b.value

    def autotune_rows(self, nsteps=3):
        """Time every row-pass variant pair from the current state, restore
        the state and keep the fastest: ((p1, p3), ms/step [p1][p3], -1
        where a variant does not exist)."""
        ch = (nat.ctypes.c_int * 2)()
        ms = np.zeros(6)
        nat.check(nat.lib().vk_gpe_autotune_rows(self.h, int(nsteps), ch, nat.ptr(ms)))
        return (ch[0], ch[1]), ms.reshape(3, 2)

    def pass_bytes(self):
        k = nat.lib().vk_pass_kinds()
        names = [nat.lib().vk_pass_name(i).decode() for i in range(k)]
        return {names[i]: nat.lib().vk_gpe_pass_bytes(self.h, i) for i in range(k)}


def _fortran_e(x, w=15, d=6):
    """Fortran Ew.d edit descriptor (0.dddddddE+xx, right-justified)."""
    if x == 0.0:
        m, e = 0.0, 0
    else:
        e = int(math.floor(math.log10(abs(x)))) + 1
        m = abs(x) / 10.0 ** e
        if round(m, d) >= 1.0:
            m, e = m / 10.0, e + 1
    t = "%s0.%0*dE%+03d" % ("-" if x < 0 else "", d, int(round(m * 10 ** d)), e)
    return t.rjust(w)


def _fortran_record(f, a):
    b = np.ascontiguousarray(a).tobytes()
    m = np.array([len(b)], dtype="<i4").tobytes()
    f.write(m + b + m)


class STS:
    """Spatio-temporal spectra of one GPE state: sts_gpe
    (src/diagnostics.f90:636-695) snapshots kpsi every sampling interval,
    and the CompSTS post-processing (Run/CompSTS/driver.f90:134-265) turns
    nt snapshots into spectra.  Here the snapshots stay on the device:
    record(time) after each interval, compute() after nt of them returns
    spec2d[w, shell] (shells 0 .. nshells-1, nint(|k|)) and spec1d[w, ky]
    (the kx = 0 modes), frequency index w in FFT order."""

    def __init__(self, gpe, nt):
        h = nat.ctypes.c_void_p()
        nat.check(nat.lib().vk_sts_create(nat.ctypes.byref(h), gpe.h, int(nt)))
        self.h, self.gpe, self.nt, self.N = h, gpe, int(nt), gpe.N
        self.nshells = nat.lib().vk_sts_shells(self.h)
        self.nmodes = nat.lib().vk_sts_modes(self.h)
        self.times = []

    def close(self):
        if getattr(self, "h", None) and nat is not None:
            nat.lib().vk_sts_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def sts_gpe(self, time=0.0):
        """Snapshot the current kpsi (the reference's sts_gpe)."""
        nat.check(nat.lib().vk_sts_record(self.h))
        self.times.append(float(time))

    record = sts_gpe

    def reset(self):
        nat.check(nat.lib().vk_sts_reset(self.h))
        self.times = []

    @property
    def count(self):
        return nat.lib().vk_sts_count(self.h)

    def compute(self):
        s2 = np.empty((self.nt, self.nshells))
        s1 = np.empty((self.nt, self.N))
        nat.check(nat.lib().vk_sts_compute(self.h, nat.ptr(s2), nat.ptr(s1)))
        return s2, s1

    def kernel_times(self, reps=10):
        """Average ms of the compute kernels: (time transforms, shell sums, kx=0 lines)."""
        out = np.zeros(3)
        nat.check(nat.lib().vk_sts_kernel_times(self.h, int(reps), nat.ptr(out)))
        return tuple(out)

    def x0(self):
        """Recorded kpsi(ky, kx=0) columns, (count, N) complex128."""
        out = np.empty((self.count, self.N), dtype=np.complex128)
        nat.check(nat.lib().vk_sts_get_x0(self.h, nat.ptr(out)))
        return out

    def snapshot(self, n):
        """Recorded snapshot n as kpsi_t[kx, ky] (zero off the retained modes)."""
        out = np.empty((self.N, self.N), dtype=np.complex128)
        nat.check(nat.lib().vk_sts_get_kpsi(self.h, int(n), nat.ptr(out)))
        return out

    def write_sts_files(self, directory, nrst, nsplit=1):
        """The files sts_gpe appends during a run, written at once (see
        write_sts_files)."""
        write_sts_files(directory, nrst, self.times, self.x0(),
                        (self.snapshot(n) for n in range(self.count)), self.N, nsplit)

    def write_spectra(self, directory, nrst, spec2d, spec1d):
        """CompSTS's text outputs: sts1d_kx0_nrst<n>.dat (one line of Ny
        values per frequency) and sts2d_nrst<n>.dat (shells 0 .. int(Nx/3))."""
        import os
        np.savetxt(os.path.join(directory, "sts1d_kx0_nrst%d.dat" % nrst), spec1d, fmt="%25.17E")
        np.savetxt(os.path.join(directory, "sts2d_nrst%d.dat" % nrst), spec2d[:, :self.N // 3 + 1], fmt="%25.17E")


def write_sts_files(directory, nrst, times, x0, planes, N, nsplit=1):
    """The files of sts_gpe (src/diagnostics.f90:636-695): sts_time.dat_nrst<n>
    (one E15.6 line per snapshot), sts1d/kpsi_x0_nrst<n>.dat (one unformatted
    record of Ny complex per snapshot) and sts2d/kpsi_nrst<n>p<r>.dat (per
    slab of N/nsplit kx columns, one record of Ny x N/nsplit complex per
    snapshot, ky fastest).  planes yields kpsi_t[kx, ky] per snapshot."""
    import os
    if N % nsplit:
        raise ValueError("nsplit must divide N")
    os.makedirs(os.path.join(directory, "sts1d"), exist_ok=True)
    os.makedirs(os.path.join(directory, "sts2d"), exist_ok=True)
    with open(os.path.join(directory, "sts_time.dat_nrst%d" % nrst), "w") as f:
        for t in times:
            f.write(_fortran_e(t) + "\n")
    with open(os.path.join(directory, "sts1d", "kpsi_x0_nrst%d.dat" % nrst), "wb") as f:
        for row in x0:
            _fortran_record(f, np.asarray(row, dtype=np.complex128))
    L = N // nsplit
    files = [open(os.path.join(directory, "sts2d", "kpsi_nrst%dp%d.dat" % (nrst, r + 1)), "wb")
             for r in range(nsplit)]
    try:
        for k in planes:
            k = np.asarray(k, dtype=np.complex128)
            for r in range(nsplit):
                _fortran_record(files[r], k[r * L:(r + 1) * L])
    finally:
        for f in files:
            f.close()


def write_wf(directory, iopoint, time, dt, psi, nsplit=1):
    """wf checkpoint files of io_fields (src/io_fields.f90:41-50) from a
    host psi[y, x] complex128 array."""
    a = np.ascontiguousarray(psi, dtype=np.complex128)
    nat.check(nat.lib().vk_wf_write(str(directory).encode(), int(iopoint), float(time), float(dt), a.shape[0],
                                    int(nsplit), nat.ptr(a)))


def read_wf(directory, nrestart, N, nsplit_orig=1):
    """(iopoint, time, dt, psi[y, x]) from the nsplit_orig wf files of
    checkpoint nrestart (restart_gpe, src/io_fields.f90:71-224)."""
    psi = np.empty((N, N), dtype=np.complex128)
    icp = nat.ctypes.c_int()
    t, dt = np.zeros(1), np.zeros(1)
    nat.check(nat.lib().vk_wf_read(str(directory).encode(), int(nrestart), int(N), int(nsplit_orig),
                                   nat.ctypes.byref(icp), nat.ptr(t), nat.ptr(dt), nat.ptr(psi)))
    return icp.value, float(t[0]), float(dt[0]), psi


class Ensemble:
    """Independent realizations of unequal size stepped together: every pass
    of an RK4 step is one batched launch, and a step replays as a captured
    HIP graph (no reference counterpart; the reference runs one field per
    MPI job)."""

    def __init__(self, params_list):
        n = len(params_list)
        arr = (nat.Params * n)(*params_list)
        self.params = list(params_list)
        self.sizes = [p.N for p in params_list]
        h = nat.ctypes.c_void_p()
        nat.check(nat.lib().vk_ens_create(nat.ctypes.byref(h), n, arr))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and nat is not None:
            nat.lib().vk_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __len__(self):
        return len(self.params)

    def set_psi(self, r, psi, sync=True):
        a = np.ascontiguousarray(psi, dtype=np.complex128)
        nat.check(nat.lib().vk_ens_set_psi(self.h, int(r), nat.ptr(a), 1 if sync else 0))

    def set_ic_vortices(self, r, xs, ys, qs, nimg=3, sync=True):
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        ys = np.ascontiguousarray(ys, dtype=np.float64)
        qs = np.ascontiguousarray(qs, dtype=np.int32)
        nat.check(nat.lib().vk_ens_set_ic_vortices(self.h, int(r), len(xs), nat.ptr(xs), nat.ptr(ys),
                                                   qs.ctypes.data_as(nat.ip), int(nimg), 1 if sync else 0))

    def psi(self, r):
        N = self.params[r].N
        out = np.empty((N, N), dtype=np.complex128)
        nat.check(nat.lib().vk_ens_get_psi(self.h, int(r), nat.ptr(out), None))
        return out

    def kpsi(self, r):
        N = self.params[r].N
        out = np.empty((N, N), dtype=np.complex128)
        nat.check(nat.lib().vk_ens_get_psi(self.h, int(r), None, nat.ptr(out)))
        return out

    def diagnostics(self, r):
        ns = nshell(self.params[r].N)
        scal, spec, pscal = np.zeros(8), np.zeros(6 * (ns + 1)), np.zeros(5)
        nat.check(nat.lib().vk_ens_diagnostics(self.h, int(r), nat.ptr(scal), nat.ptr(spec), nat.ptr(pscal)))
        return scal, spec.reshape(6, ns + 1)

    def diagnostics_all(self):
        """diagnostics(r) of every member in one call: a list of
        (scalars[8], spectra[6, nshell+1]) per member."""
        ns = [nshell(p.N) for p in self.params]
        scal = np.zeros(8 * len(ns))
        spec = np.zeros(sum(6 * (n + 1) for n in ns))
        nat.check(nat.lib().vk_ens_diagnostics_all(self.h, nat.ptr(scal), nat.ptr(spec)))
        out, o = [], 0
        for r, n in enumerate(ns):
            out.append((scal[8 * r:8 * r + 8], spec[o:o + 6 * (n + 1)].reshape(6, n + 1)))
            o += 6 * (n + 1)
        return out

    def rk4(self, nsteps):
        nat.check(nat.lib().vk_ens_rk4_steps(self.h, int(nsteps)))

    def rk4_async(self, nsteps):
        nat.check(nat.lib().vk_ens_rk4_steps_async(self.h, int(nsteps)))

    @property
    def stream(self):
        return nat.lib().vk_ens_stream(self.h)

    def set_mode(self, mode, graph=True):
        nat.check(nat.lib().vk_ens_set_mode(self.h, int(mode), 1 if graph else 0))

    def mode(self):
        m, g = nat.ctypes.c_int(), nat.ctypes.c_int()
        nat.check(nat.lib().vk_ens_get_mode(self.h, nat.ctypes.byref(m), nat.ctypes.byref(g)))
        return m.value, bool(g.value)

    def autotune(self, nsteps=3):
        ms = np.zeros(8)
        nat.check(nat.lib().vk_ens_autotune(self.h, int(nsteps), nat.ptr(ms)))
        return ms.reshape(4, 2)

    def set_shape(self, N, var):
        nat.check(nat.lib().vk_ens_set_shape(self.h, int(N), int(var)))

    def shape(self, N):
        v = nat.ctypes.c_int()
        nat.check(nat.lib().vk_ens_get_shape(self.h, int(N), nat.ctypes.byref(v)))
        return v.value

    def autotune_shapes(self, nsteps=3):
        """Times both member shapes per size; returns {N: (chosen, ms0, ms1)}."""
        sizes = sorted(set(self.sizes))
        vs = np.zeros(len(sizes), dtype=np.int32)
        ms = np.zeros(2 * len(sizes))
        nat.check(nat.lib().vk_ens_autotune_shapes(self.h, int(nsteps), vs.ctypes.data_as(nat.ip), nat.ptr(ms)))
        return {N: (int(vs[i]), float(ms[2 * i]), float(ms[2 * i + 1])) for i, N in enumerate(sizes)}


def ic_vortices_host(N, xs, ys, qs, xi, nimg=3):
    """Padé-core vortex product on the CPU (small grids; input generation)."""
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    qs = np.ascontiguousarray(qs, dtype=np.int32)
    out = np.empty((N, N), dtype=np.complex128)
    nat.check(nat.lib().vk_ic_vortices_host(int(N), len(xs), nat.ptr(xs), nat.ptr(ys),
                                            qs.ctypes.data_as(nat.ip), float(xi), int(nimg), nat.ptr(out)))
    return out


def vortex_velocity_host(N, xs, ys, qs, xi, alpha, nimg=3):
    """Imposed velocity u = 2 alpha rho0 grad(theta0) of a vortex list on
    the CPU (the field vk_gpe_set_vortex_advection builds on the device)."""
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    qs = np.ascontiguousarray(qs, dtype=np.int32)
    ux = np.empty((N, N))
    uy = np.empty((N, N))
    nat.check(nat.lib().vk_vortex_velocity_host(int(N), len(xs), nat.ptr(xs), nat.ptr(ys), qs.ctypes.data_as(nat.ip),
                                                float(xi), float(alpha), int(nimg), nat.ptr(ux), nat.ptr(uy)))
    return ux, uy


def lattice_vortices(n=4, L=2.0 * math.pi):
    """n x n lattice at ((i+1/2)L/n, (j+1/2)L/n), charge (-1)^(i+j)."""
    xs, ys, qs = [], [], []
    for i in range(n):
        for j in range(n):
            xs.append((i + 0.5) * L / n)
            ys.append((j + 0.5) * L / n)
            qs.append(1 if (i + j) % 2 == 0 else -1)
    return np.array(xs), np.array(ys), np.array(qs, dtype=np.int32)


def random_vortices(nv, seed, L=2.0 * math.pi):
    """nv vortices at seeded uniform positions with zero net charge."""
    rng = np.random.default_rng(seed)
    xs = rng.uniform(0.0, L, nv)
    ys = rng.uniform(0.0, L, nv)
    qs = np.array([1] * (nv // 2) + [-1] * (nv - nv // 2), dtype=np.int32)
    rng.shuffle(qs)
    return xs, ys, qs
