"""Benchmark: RK4 steps/s of the FP64 4096^2 GPE step on one MI355X.

Workload (the 4096^2 configuration the headline metric is quoted on):
4096^2 periodic box L=2pi, psi complex128, g=1, xi=2.5 dx, density 1,
256 random-sign vortices (seed 3, zero net charge, Padé cores), 32 inertial
particles (seed 4) with Gaussian potential V0=10 g, width 2 xi, mass 5x the
displaced fluid; RK4 dt=2.5e-4, 2/3 dealiasing.  Inputs are resident in HBM
before timing starts; the timed region is K full RK4 steps (4 stages, all
passes and the particle update) bracketed by HIP events on the library's
stream.

Also reported: the pure-GPE (no particle) rate on the same field, the
per-kernel roofline of the dominant pass (algorithmic bytes / measured
average duration, against the 8 TB/s HBM peak), the step-level roofline
(compulsory bytes of all passes per step, and the counter-measured bytes),
and the CPU baseline: the plain-C oracle (same algorithm as the reference,
16 unfused 2D FFTs per step plus the per-particle transforms) timed on
bounded samples of the same workload on the host cores, one thread and all; the batched ensemble; and ARGLE time-to-1e-10
at 512^2 with the learned warm start against plain ARGLE from the Pade
ansatz (and against the Thomas-Fermi start); and the spatio-temporal
spectra at the reference's CompSTS size (128^2, 2048 snapshots).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK4 steps/s at 4096² fp64 (HBM roofline %); ARGLE time-to-1e-10 warm vs plain"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(G, N, dt, particles, mass=None):
    """The 4096^2 workload (quantumvortices_amd.workloads): 256 random-sign
    vortices (seed 3) and, with particles, 32 particles (seed 4) of mass 5x
    the fluid one particle displaces in its ARGLE ground state."""
    from quantumvortices_amd import workloads as W
    g, info = W.particle_workload(N, dt, npart=32 if particles else 0, mass=mass)
    return g, info


def time_steps(torch, g, steps, warmup):
    ext = torch.cuda.ExternalStream(g.stream)
    g.rk4_async(warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(ext)
    g.rk4_async(steps)
    e1.record(ext)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(G, N, dt, info, threads, psi_gpu):
    """The oracle (the reference's algorithm: 16 unfused 2D FFTs per RK4
    step, per-particle spectral translation and forces) on the host cores,
    on bounded samples of the workloads:
      * 4096^2, `threads` threads: K = 2 full pure-GPE RK4 steps (measured),
        one translation of the 32-particle potential and one particle force
        (measured); the particle step is the pure step plus, per stage, the
        translation and 32 forces (src/gpe.f90:498-671 calls the translation
        once per right-hand side, src/gpe_particle.f90:429-531 one force per
        particle per stage);
      * 4096^2, one thread: one right-hand side and one inverse FFT,
        extrapolated to a pure step as 4 x (rhs + ifft);
      * 128^2 (the reference's own CPU-runnable case, mpiexec -n 1): 200 full
        RK4 steps on one thread and on `threads` threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    from quantumvortices_amd import workloads as W
    kw = info["kw"]
    npart = kw.get("gp_N_obj", 0)
    op = O.make_params(N, dt, W.XI_IN, np_=npart, V0_in=W.V0_IN, d_in=W.D_IN, mass=kw.get("gp_mass_part", 1.0))
    legs = {}
    O.set_threads(threads)
    # pure GPE: full steps
    pure = O.State(O.make_params(N, dt, W.XI_IN))
    pure.set_psi(psi_gpu, sync=True)
    K = 2
    t0 = time.perf_counter()
    pure.rk4(K)
    t_pure = (time.perf_counter() - t0) / K
    legs["4096_pure_threads"] = {"sec_per_step": round(t_pure, 3), "steps_timed": K, "threads": threads}
    # particle parts
    st = O.State(op)
    st.set_psi(psi_gpu, sync=True)
    st.part(0)[...] = info["pos"]
    st.particle_potential()
    t0 = time.perf_counter()
    st.translate_all()
    t_tr = time.perf_counter() - t0
    st.shortrange()
    t0 = time.perf_counter()
    st.particle_force(0)
    t_force = time.perf_counter() - t0
    t_part = t_pure + 4.0 * (t_tr + npart * t_force)
    legs["4096_particles_threads"] = {"sec_per_step": round(t_part, 2), "threads": threads,
                                      "translate_sec": round(t_tr, 3), "force_sec": round(t_force, 3)}
    del st
    # one thread, pure GPE
    O.set_threads(1)
    t0 = time.perf_counter()
    pure.rhs()
    t_rhs1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.fft2d(pure.kpsi, +1)
    t_ifft1 = time.perf_counter() - t0
    legs["4096_pure_1thread"] = {"sec_per_step": round(4.0 * (t_rhs1 + t_ifft1), 2), "threads": 1,
                                 "rhs_sec": round(t_rhs1, 3), "ifft_sec": round(t_ifft1, 3)}
    del pure
    # the reference's CPU-runnable case: 128^2, 4x4 vortex lattice, dt = 1e-3
    n1 = 128
    p1 = G.params_from_namelist(n1, 1e-3, W.XI_IN)
    xs, ys, qs = G.lattice_vortices(4)
    psi1 = G.ic_vortices_host(n1, xs, ys, qs, p1.xi)
    for thr in (1, threads):
        O.set_threads(thr)
        s1 = O.State(O.make_params(n1, 1e-3, W.XI_IN))
        s1.set_psi(psi1, sync=True)
        t0 = time.perf_counter()
        s1.rk4(200)
        legs["128_%s" % ("1thread" if thr == 1 else "threads")] = {
            "steps_per_s": round(200 / (time.perf_counter() - t0), 1), "steps_timed": 200, "threads": thr}
    sample = ("oracle at %d^2: %d full pure-GPE RK4 steps on %d threads (%.2f s/step), plus one "
              "32-particle potential translation (%.2f s) and one particle force (%.2f s) measured and "
              "extrapolated to 4 x (translation + %d forces) per step -> %.1f s/step"
              % (N, K, threads, t_pure, t_tr, t_force, npart, t_part))
    return {"value": float("%.4g" % (1.0 / t_part)), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": sample, "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)), "pure_gpe_steps_per_s": float("%.4g" % (1.0 / t_pure)),
            "legs": legs}


def ensemble_bench(G, torch, steps, warmup):
    """Batched ensemble: 384 independent realizations, 128 each at 128^2,
    256^2 and 512^2, with 8-64 zero-net-charge random vortices (seed =
    realization id); RK4 dt=1e-3; launch schedule and member shapes picked
    by the autotuners."""
    import numpy as np
    sizes = [128] * 128 + [256] * 128 + [512] * 128
    ps = [G.params_from_namelist(N, 1e-3, 2.5) for N in sizes]
    e = G.Ensemble(ps)
    for r in range(len(sizes)):
        rng = np.random.default_rng(r)
        nv = 2 * int(rng.integers(4, 33))          # 8..64, even
        xs, ys, qs = G.random_vortices(nv, seed=r)
        e.set_ic_vortices(r, xs, ys, qs)
    tune = e.autotune(3)
    shapes = e.autotune_shapes(3) if e.mode()[0] != 2 else {}
    ext = torch.cuda.ExternalStream(e.stream)
    e.rk4_async(warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(ext)
    e.rk4_async(steps)
    e1.record(ext)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    mode, graph = e.mode()
    points = sum(N * N for N in sizes)
    # every member's diagnostics (energies and six spectra): one call for
    # all members, then one call per member (the first call allocates the
    # diagnostics' work fields and is not timed)
    e.diagnostics_all()
    t0 = time.perf_counter()
    e.diagnostics_all()
    diag_all_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    for r in range(len(sizes)):
        e.diagnostics(r)
    diag_loop_ms = (time.perf_counter() - t0) * 1e3
    e.close()
    # compulsory HBM bytes of one step of every member: the four passes of
    # the four stages, per member size (the members run the same pass
    # bodies as a single field of that size)
    per_size = {}
    for N in sorted(set(sizes)):
        g1 = G.GPE(G.params_from_namelist(N, 1e-3, 2.5))
        pb = g1.pass_bytes()
        per_size[N] = 4.0 * sum(pb[k] for k in ("p1_rho_row", "p2_dealias_col", "p3_nonlin_row", "p4_rk4_col"))
        g1.close()
    step_bytes = sum(per_size[N] for N in sizes)
    ach = step_bytes / (ms * 1e-3) / 1e9
    names = ["all sizes per launch", "one launch per size", "one launch per realization",
             "one persistent launch per step"]
    return {"workload": "384 realizations (128 x 128^2, 128 x 256^2, 128 x 512^2), 8-64 vortices, RK4 dt=1e-3",
            "ensemble_steps_per_s": round(1e3 / ms, 2), "ms_per_step": round(ms, 4),
            "realization_steps_per_s": round(len(sizes) * 1e3 / ms, 1),
            "grid_point_steps_per_s": float("%.4g" % (points * 1e3 / ms)),
            "schedule": {"launch": names[mode], "graph": graph},
            "diagnostics_all_members_ms": {"one call": round(diag_all_ms, 2),
                                           "one call per member": round(diag_loop_ms, 2)},
            "shapes": {str(N): {"chosen": ("8 points per thread" if v[0] else "16 points per thread"),
                                "ms_16pt": round(v[1], 4) if v[1] > 0 else None,
                                "ms_8pt": round(v[2], 4) if v[2] > 0 else None} for N, v in shapes.items()},
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4), "compulsory_bytes_per_step": step_bytes},
            "autotune_ms_per_step": {names[m]: {"graph": round(float(tune[m][0]), 4), "direct": round(float(tune[m][1]), 4)}
                                     for m in range(4)}}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "rk4_4096_traffic.txt")
PASS_KERNEL = {"p1_rho_row": "vk::k_p1x<4096, TR=1 >", "p2_dealias_col": "vk::k_p2<4096, 1, 0>",
               "p3_nonlin_row": "vk::k_p3x<4096, TR=1, 6>", "p4_rk4_col": "vk::k_p4<4096, 2, avg>"}


# row-pass variants the traffic summary was taken with (P1 one row per
# workgroup, P3 one row)
PROFILED_ROWS = (1, 1)
ROW_P1_NAMES = ["P1 TR rows per workgroup", "P1 one row per workgroup", "P1 persistent with LDS prefetch"]
ROW_P3_NAMES = ["P3 TR rows per workgroup", "P3 one row per workgroup"]


def tune_rows(g):
    """Row-pass variant autotuning (N >= 2048): ((p1, p3), ms table), or
    (None, None) where the size has one variant."""
    from quantumvortices_amd import _native as nat
    try:
        return g.autotune_rows(3)
    except nat.VKError:
        return None, None


def measured_traffic(pass_name, N, rows=PROFILED_ROWS):
    """HBM bytes per launch of a pass at 4096^2 with particles, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a
    device copy); None for other sizes, for a row-pass variant other than
    the profiled one, or when the summary is missing."""
    if N != 4096 or not os.path.exists(TRAFFIC_FILE):
        return None
    if (pass_name == "p1_rho_row" and rows[0] != PROFILED_ROWS[0]) or \
            (pass_name == "p3_nonlin_row" and rows[1] != PROFILED_ROWS[1]):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith(PASS_KERNEL.get(pass_name, "?") + " "):
            return float(line.split()[-2 if len(line.split()) > 7 else -1]) * 1e6
    return None


def argle_bench(G, torch, n_cfg, plain_cap):
    """ARGLE time-to-1e-10 at 512^2: warm start (network + ARGLE) vs plain
    ARGLE from the Pade product ansatz vs the Thomas-Fermi start, on held-out
    seeded configurations (16-128 vortices, 0-8 obstacles, V0 in [2, 20] g).
    Wall time from the first call (input planes / ansatz) to the step that
    meets the residual; plain runs are capped at plain_cap steps."""
    from quantumvortices_amd import warmstart as W
    fno = W.FNO.load(os.path.join(ROOT, "quantumvortices_amd", "data", "fno_argle_512.npz"), batch=1)
    g = G.GPE(W.solver_params(512, W.DT_MAX))
    # untimed first pass: kernel loading and LDS limits
    W.warm_start(g, fno, W.random_config(899999), torch.empty((512, 512), dtype=torch.complex128, device="cuda"))
    rows = W.compare(fno, g, [W.random_config(900000 + k) for k in range(n_cfg)], plain_cap=plain_cap,
                     plain_check=2000, log=log)
    fno.close()
    g.close()
    conv = [r for r in rows if r["plain_converged"]]
    gm = lambda xs: float(math.exp(sum(math.log(x) for x in xs) / len(xs))) if xs else None
    return {
        "workload": "512^2 ARGLE to residual 1e-10, dt 0.08 (lower for overlapping obstacles), "
                    "held-out seeds 900000+, network trained on %d solver-generated ground states"
                    % fno.info.get("samples", 0),
        "warm_ms": [round(1e3 * r["warm_sec"], 2) for r in rows],
        "tf_ms": [round(1e3 * r["tf_sec"], 2) for r in rows],
        "plain_ms": [round(1e3 * r["plain_sec"], 1) for r in rows],
        "warm_steps": [r["warm_steps"] for r in rows],
        "tf_steps": [r["tf_steps"] for r in rows],
        "plain_steps": [r["plain_steps"] for r in rows],
        "plain_converged": [r["plain_converged"] for r in rows],
        "speedup_vs_plain": round(gm([r["plain_sec"] / r["warm_sec"] for r in rows]), 1),
        "speedup_vs_plain_is_lower_bound": len(conv) < len(rows),
        "speedup_vs_tf": round(gm([r["tf_sec"] / r["warm_sec"] for r in rows]), 3),
        "omega_warm_minus_plain_rel": [float("%.3g" % ((r["warm_omega"] - r["plain_omega"]) / abs(r["plain_omega"])))
                                       for r in rows],
        "omega_warm_minus_tf_rel": [float("%.3g" % ((r["warm_omega"] - r["tf_omega"]) / abs(r["tf_omega"])))
                                    for r in rows],
        "surrogate_raw_rel_err": [float("%.3g" % r["raw_rel_err"]) for r in rows],
        "tf_raw_rel_err": [float("%.3g" % r["tf_raw_rel_err"]) for r in rows],
    }


def diagnostics_bench(G, torch):
    """diagnostics_gpe (energies, norm, momentum and the six shell spectra)
    at 1024^2 (64 seeded vortices, the 1024^2 configuration's field) and on
    the 4096^2 workload field: wall time of one synchronous call (kernels,
    the host copies of the scalars and spectra, and the sync), after one
    untimed call that allocates the diagnostics' work fields."""
    out = {}
    for N, nv, seed in ((1024, 64, 1), (4096, 256, 3)):
        p = G.params_from_namelist(N, 5e-4, 2.5)
        g = G.GPE(p)
        xs, ys, qs = G.random_vortices(nv, seed)
        g.set_ic_vortices(xs, ys, qs, nimg=3)
        g.diagnostics_gpe()
        torch.cuda.synchronize()
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            g.diagnostics_gpe()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        out["%d" % N] = {"ms_per_call": round(ms, 3), "shells": g.nshell}
        g.close()
    return out


def small_field_bench(G, torch):
    """RK4 of one 1024^2 field (the 1024^2 configuration's size, 64 seeded
    vortices, dt = 5e-4, no particles) in both transform shapes: ms per
    step over 200 steps after 20 warm, and the compulsory HBM bytes of a
    step (the four passes of the four stages) over the faster time."""
    N = 1024
    p = G.params_from_namelist(N, 5e-4, 2.5)
    xs, ys, qs = G.random_vortices(64, 1)
    out = {"workload": "GPE RK4 1024^2 FP64, dt=5e-4, 64 random-sign vortices (seed 1), one field"}
    best = None
    for var in (0, 1):
        g = G.GPE(p)
        g.set_ic_vortices(xs, ys, qs, nimg=3)
        g.set_shape(var)
        ms = time_steps(torch, g, 200, 20) / 200
        out["ms_per_step_shape%d" % var] = round(ms, 4)
        if best is None or ms < best[0]:
            pb = g.pass_bytes()
            best = (ms, 4.0 * sum(pb[k] for k in ("p1_rho_row", "p2_dealias_col", "p3_nonlin_row", "p4_rk4_col")))
        g.close()
    ms, step_bytes = best
    out["steps_per_s"] = round(1e3 / ms, 1)
    ach = step_bytes / (ms * 1e-3) / 1e9
    out["step_roofline"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(ach / HBM_PEAK_GBS, 4), "compulsory_bytes_per_step": step_bytes}
    return out


def sts_bench(G, torch):
    """Spatio-temporal spectra at the size of the reference's CompSTS driver
    (128^2, Nt = 2048 snapshots): one snapshot of the retained modes per
    record (a gather on the library's stream), then the windowed time
    transforms of all retained modes, the shell sums and the kx = 0 lines,
    with the copies of both spectra to the host."""
    N, nt = 128, 2048
    p = G.params_from_namelist(N, 1e-3, 2.5)
    g = G.GPE(p)
    xs, ys, qs = G.lattice_vortices(4)
    g.set_psi(G.ic_vortices_host(N, xs + 0.01, ys + 0.02, qs, p.xi), sync=True)
    sts = G.STS(g, nt)
    ext = torch.cuda.ExternalStream(g.stream)
    for _ in range(nt):                       # untimed: first launches
        sts.record()
    sts.compute()
    sts.reset()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record(ext)
    for _ in range(nt):
        sts.record()
    e1.record(ext)
    sts.compute()
    e2.record(ext)
    torch.cuda.synchronize()
    rec_ms, comp_ms = e0.elapsed_time(e1), e1.elapsed_time(e2)
    kt = sts.kernel_times(20)
    nm = sts.nmodes
    # time transforms: read the snapshots (16 B per mode and snapshot), write
    # the power (8 B); shell sums: read the power once
    bytes_time = 24.0 * nm * nt
    bytes_bin = 8.0 * nm * nt
    sts.close()
    g.close()
    ach = bytes_time / (kt[0] * 1e-3) / 1e9
    return {"workload": "%d^2, Nt=%d snapshots, %d retained modes, %d shells" % (N, nt, nm, sts.nshells),
            "record_us_per_snapshot": round(1e3 * rec_ms / nt, 3),
            "compute_ms": round(comp_ms, 4),
            "kernel_ms": {"k_sts_time": round(kt[0], 4), "k_sts_bin": round(kt[1], 4), "k_sts_x0": round(kt[2], 4)},
            "roofline": {"bound": "hbm", "kernel": "k_sts_time", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_launch": bytes_time,
                         "avg_ms": round(kt[0], 4)},
            "shell_sum_GBps": round(bytes_bin / (kt[1] * 1e-3) / 1e9, 1)}


def cpu_baseline_sts(G, threads):
    """CPU leg of the spectra: the numpy oracle (CompSTS restated: Hann
    window, numpy's FFT along time, nint(|k|) shells) on the 128^2 x 2048
    snapshots of the same run, fetched from the device."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import sts_oracle
    N, nt = 128, 2048
    p = G.params_from_namelist(N, 1e-3, 2.5)
    g = G.GPE(p)
    xs, ys, qs = G.lattice_vortices(4)
    g.set_psi(G.ic_vortices_host(N, xs + 0.01, ys + 0.02, qs, p.xi), sync=True)
    sts = G.STS(g, nt)
    for _ in range(nt):
        sts.record()
    snaps = np.array([sts.snapshot(n) for n in range(nt)])
    sts.close()
    g.close()
    t0 = time.perf_counter()
    sts_oracle.sts_spectra(snaps)
    sec = time.perf_counter() - t0
    return {"value": round(sec * 1e3, 1), "unit": "ms per spectra computation", "cores": 1, "kind": "port",
            "sample": "the full 128^2 x 2048 computation (numpy FFT, single thread)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=4096)
    ap.add_argument("--dt", type=float, default=2.5e-4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pure", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-argle", action="store_true")
    ap.add_argument("--argle-configs", type=int, default=2)
    ap.add_argument("--argle-plain-cap", type=int, default=400000)
    ap.add_argument("--no-sts", action="store_true")
    ap.add_argument("--no-diag", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only (the single-field path would shard as "
              "1D slabs over RCCL; that is not implemented); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from quantumvortices_amd import gpe as G

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    torch.cuda.set_device(0)
    N, dt = args.N, args.dt

    log("bench: building %d^2 workload with particles" % N)
    g, wl = make_workload(G, N, dt, particles=True)
    # row-pass variants: timed from the initial state, which is restored
    rows, rows_ms = tune_rows(g)
    log("bench: row variants (p1, p3) =", rows, None if rows_ms is None else rows_ms.tolist())
    ms = time_steps(torch, g, args.steps, args.warmup)
    steps_per_s = args.steps / (ms / 1e3)
    log("bench: particles %d steps in %.1f ms -> %.2f steps/s" % (args.steps, ms, steps_per_s))
    # per-kernel durations (events around every launch, separate run)
    pt = g.pass_times(max(2, min(10, args.steps)))
    pb = g.pass_bytes()
    log("bench: pass ms", pt)
    psi_for_cpu = g.psi if not args.no_cpu else None
    g.close()

    pure = None
    pure_pt = None
    pure_rows = None
    if not args.no_pure:
        gp, _ = make_workload(G, N, dt, particles=False)
        pure_rows, _ = tune_rows(gp)
        pms = time_steps(torch, gp, args.steps, args.warmup)
        pure = args.steps / (pms / 1e3)
        pure_pt = gp.pass_times(max(2, min(10, args.steps)))
        log("bench: pure GPE %.2f steps/s" % pure, pure_pt)
        gp.close()

    # roofline of the dominant pass (largest time per step; each runs 4x per step)
    passes = ["p1_rho_row", "p2_dealias_col", "p3_nonlin_row", "p4_rk4_col"]
    dom = max(passes, key=lambda k: pt[k])
    achieved = pb[dom] / (pt[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, N, rows or PROFILED_ROWS),
            "bytes_per_launch": pb[dom], "avg_ms": pt[dom]}
    per_pass = {k: {"ms": round(pt[k], 4), "GB/s": round(pb[k] / (pt[k] * 1e-3) / 1e9, 1) if pt[k] > 0 else None}
                for k in passes}
    # attainable bandwidth of the passes' access patterns on this GPU (no
    # arithmetic): the ceilings below the 8 TB/s peak
    import numpy as np
    from quantumvortices_amd import _native as nat
    pat = np.zeros(4)
    nat.check(nat.lib().vk_pattern_bandwidth(20, nat.ptr(pat), 4))
    attain = {"contiguous_copy": round(pat[0], 1), "row_pass_panel_gather": round(pat[1], 1),
              "column_pass_panel_copy": round(pat[2], 1), "column_pass_one_column": round(pat[3], 1)}
    ceiling = {"p1_rho_row": pat[1], "p2_dealias_col": pat[3], "p3_nonlin_row": pat[1], "p4_rk4_col": pat[2]}
    for k in passes:
        if per_pass[k]["GB/s"]:
            per_pass[k]["frac_of_pattern_ceiling"] = round(per_pass[k]["GB/s"] / ceiling[k], 3)
    roof["attainable_GBps"] = attain
    roof["frac_of_pattern_ceiling"] = round(achieved / ceiling[dom], 4)
    log("bench: attainable pattern bandwidth", attain)
    # step-level roofline: the compulsory bytes of the four passes of the
    # four stages (the particle kernels move a few MB), and the bytes the
    # counters measured for the same passes
    step_bytes = 4.0 * sum(pb[k] for k in passes)
    meas = [measured_traffic(k, N, rows or PROFILED_ROWS) for k in passes]
    step_meas = 4.0 * sum(meas) if all(m is not None for m in meas) else None
    step_gbs = step_bytes * steps_per_s / 1e9
    step_roof = {"bound": "hbm", "compulsory_bytes_per_step": step_bytes, "achieved": round(step_gbs, 1),
                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4),
                 "measured_bytes_per_step": step_meas,
                 "measured_frac": round(step_meas * steps_per_s / 1e9 / HBM_PEAK_GBS, 4) if step_meas else None}

    ens = None
    if not args.no_ensemble:
        log("bench: ensemble")
        ens = ensemble_bench(G, torch, max(10, args.steps // 4), 3)
        log("bench: ensemble", ens)

    argle = None
    if not args.no_argle:
        log("bench: ARGLE warm start vs plain at 512^2")
        argle = argle_bench(G, torch, args.argle_configs, args.argle_plain_cap)
        log("bench: argle", argle)

    diag = None
    small = None
    if not args.no_diag:
        log("bench: diagnostics at 1024^2 and 4096^2")
        diag = diagnostics_bench(G, torch)
        log("bench: one 1024^2 field, both transform shapes")
        small = small_field_bench(G, torch)
        log("bench: rk4_1024", small)
        log("bench: diagnostics", diag)

    sts = None
    if not args.no_sts:
        log("bench: spatio-temporal spectra 128^2 x 2048")
        sts = sts_bench(G, torch)
        log("bench: sts", sts)

    cpu = None
    if not args.no_cpu:
        # the host cores this process may use, capped by OMP_NUM_THREADS
        # (16 when unset)
        share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
        threads = args.cpu_threads or max(1, min(len(os.sched_getaffinity(0)), share))
        log("bench: CPU baseline (oracle) with %d threads" % threads)
        cpu = cpu_baseline(G, N, dt, wl, threads, psi_for_cpu)
        log("bench: cpu", cpu)
        if sts is not None:
            sts["cpu_baseline"] = cpu_baseline_sts(G, threads)
            log("bench: sts cpu", sts["cpu_baseline"])

    out = {
        "metric": METRIC,
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded vortex gas + particles; no dataset)",
        "config": {"workload": "GPE RK4 4096^2 FP64, dt=2.5e-4, 256 random-sign vortices (seed 3), "
                               "32 inertial particles (seed 4), V0=10g, d=2xi, 2/3 dealias",
                   "N": N, "dt": dt, "vortices": 256, "particles": 32, "particle_mass": wl["mass"],
                   "particle_mass_rule": "5 x integral of (1 - rho) of one particle's ARGLE ground state",
                   "parallelism": "single GPU"},
        "roofline": roof,
        "per_pass": per_pass,
        "row_variants": None if rows is None else {
            "p1": ROW_P1_NAMES[rows[0]], "p3": ROW_P3_NAMES[rows[1]], "TR": {4096: 2, 2048: 4}.get(N),
            "ms_per_step": {"%s / %s" % (ROW_P1_NAMES[a], ROW_P3_NAMES[b]): round(float(rows_ms[a][b]), 4)
                            for a in range(3) for b in range(2) if rows_ms[a][b] > 0},
            "pure_gpe": None if (pure is None or pure_rows is None) else
            {"p1": ROW_P1_NAMES[pure_rows[0]], "p3": ROW_P3_NAMES[pure_rows[1]]}},
        "pure_gpe_steps_per_s": round(pure, 3) if pure else None,
        "step_roofline": step_roof,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "argle_warm_start": argle,
        "spatio_temporal_spectra": sts,
        "diagnostics": diag,
        "rk4_1024": small,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
