"""Benchmark: the SEAM-shaped single FP32 shot on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: nx=4444, nz=3819 (padded 4552 x 3927 with nbc=50, fd_order=8),
dx=dz=12.5 m, free surface, synthetic SEAM-like medium (seed 2), dt from CFL
0.4, 10 Hz Ricker pressure source at the surface centre, 4444 vz receivers
on the surface.  A step is one time step of the whole grid (velocity, HABC,
stress, HABC, free surface, receivers); the default K = 11520 steps is the
full shot.  W warm-up steps run first (clocks, first launches), then the
wavefield and gather are reset, the step graphs are captured (no step runs),
and exactly K steps are timed from step 1.

The step is the fused single-pass form (one velocity+stress kernel over
ping-pong buffers, then velocity HABC, band stress, stress HABC, free
surface + receivers).  Prints one JSON line: Mcell-updates/s over the padded
grid, the fused kernel's HBM roofline (algorithmic 52 B per cell-update: read
5 fields + rho_vx, lam, mu_txx, write 5 fields; over
its mean duration measured with HIP events on the library's stream, at a step
inside a 64-step timing shot; HBM traffic per launch from the committed
rocprofv3 counter summary), and the CPU oracle timed on a bounded sample of
the same workload on the CPUs this process may use.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per padded cell-update (FP32, dx == dz so dt/dx/rho and
# dt/dz/rho are one array each):
BYTES_FUSED = 60     # read 5 old fields + dt/rho_vx, dt/rho_vz, lam, lam+2mu, mu_txz; write 5 fields
# FP32 media whose rho_vz / mu_txz are init_medium's averages: the fused kernel
# reads rho_vx, lam, mu_txx and forms the other coefficients (5 + 3 reads, 5 writes)
BYTES_FUSED_DERIVED = 52
# the same 15 array accesses per cell with 2-byte coefficients and/or fields
BYTES_FUSED_STORAGE = {"f32": 60, "bf16": 30, "f16": 30, "f16c": 50, "f24": 50}
# the derived-coefficient form per storage: FP32 coefficients (3 reads) and
# 4-byte (FP32) or 3-byte (F24: 16-bit plane + byte plane) fields
BYTES_FUSED_DERIVED_STORAGE = {"f32": BYTES_FUSED_DERIVED, "f24": 42}
BYTES_VELOCITY = 36  # two-pass velocity: read txx tzz txz, 2 buoyancies; read+write vx vz
BYTES_STRESS = 44    # two-pass stress: read vx vz lam lam+2mu mu_txz; read+write txx tzz txz
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_fused_seam_traffic.txt")
# the same for the split step's interior launches of k_fused
TRAFFIC_FILE_SPLIT = os.path.join(ROOT, "profiles", "k_fused_interior_seam_traffic.txt")
# and for the double step's two-step interior launches of k_fused2
TRAFFIC_FILE_DOUBLE = os.path.join(ROOT, "profiles", "k_fused2_interior_seam_traffic.txt")
METRIC = "Mcell-updates/s & %HBM roofline, 1 GPU; shots/hr + p99 latency (served)"
# The autotuner's trials and the per-kernel timings run on a timing shot of
# TIMING_NT steps (same medium, source cell and receivers), so that every
# timed launch does a step's full work whatever K the timed loop has: a
# kernel launched at a step k > nt of its shot returns at once.
TIMING_NT = 64
STAGE_STEP = 32


def stage_step(nt: int) -> int:
    """The step the per-kernel launches are timed at: inside [1, nt] of the shot they run on."""
    if nt < 1:
        raise ValueError("a shot has at least one step")
    return max(1, min(STAGE_STEP, nt))


def cpu_share() -> dict:
    """The host CPUs this process may use: its affinity set, capped by the
    cgroup's CPU quota when one is set (cgroup v2 cpu.max, else v1)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        try:
            q = int(open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us").read())
            period = int(open("/sys/fs/cgroup/cpu/cpu.cfs_period_us").read())
            if q > 0:
                quota = q / period
        except (OSError, ValueError):
            pass
    usable = aff if quota is None else max(1, min(aff, int(quota + 0.999)))
    return {"usable": usable, "host_cpus": os.cpu_count(), "affinity": aff,
            "cgroup_quota": None if quota is None else round(quota, 2)}


def split_interior_cells(nx: int, nz: int, nbc: int, M: int) -> int:
    """Cells the split step's interior tiles write: columns [xi0, xi1), rows
    [zi0, zi1), 2M + 2 cells inside the band frame of width nbc + M + 1."""
    mg = nbc + M + 1 + 2 * M + 2
    xi0, xi1, zi0, zi1 = (mg + 1) & ~1, (nx - mg) & ~1, mg, nz - mg
    return max(0, xi1 - xi0) * max(0, zi1 - zi0)


def check_roofline(frac: float, kernel_ms: float, ms_per_step: float, steps_per_launch: int = 1) -> None:
    """A roofline line is only printed for a kernel that did its work: its
    fraction of peak lies in (0, 1] and one launch is no longer than the steps
    it advances (one, or two for the double step's interior kernel)."""
    if not (0.0 < frac <= 1.0):
        raise SystemExit(f"bench.py: roofline fraction {frac:.4f} is outside (0, 1]: "
                         "the timed kernel did not do a step's work")
    if not (0.0 < kernel_ms <= steps_per_launch * ms_per_step):
        raise SystemExit(f"bench.py: one kernel launch ({kernel_ms:.4f} ms) is not within "
                         f"{steps_per_launch} step(s) ({steps_per_launch * ms_per_step:.4f} ms)")


def build_seam(nt: int, nbc=50, fd_order=8, with_weights=False, storage="f32", rec_types=("vz",)):
    from fomo_julia_amd import elastic2d as E
    from fomo_julia_amd import init as I
    from fomo_julia_amd import media
    F32 = np.float32
    dx = F32(12.5)
    vp, vs, rho = media.seam_like()
    dt = media.cfl_dt(vp, dx, 0.4)
    M = E.init_medium(vp, vs, rho, dx, dx, nbc, fd_order, None, free_surface=True, storage=storage)
    H = E.init_habc(M.nx, M.nz, nbc, dt, dx, dx, I.julia_mean_f32(vp), None, weights=with_weights)
    p = E.SimParams(dt, nt, dx, dx, fd_order)
    a = I.get_fd_coefficients(fd_order)
    wav = I.ricker_wavelet(F32(10.0), dt, nt)
    src = E.Source(I.coord_to_index(F32(M.x_max / F32(2)), M.dx, M.pad),
                   I.coord_to_index(F32(0.0), M.dz, M.pad), wav, "pressure")
    nx_in = vp.shape[0]
    # a receiver on every surface column for each type (types in blocks of nx_in)
    xr = np.tile(np.arange(nx_in, dtype=F32) * dx, len(rec_types))
    types = rec_types[0] if len(rec_types) == 1 else [t for t in rec_types for _ in range(nx_in)]
    rec = E.setup_receivers(xr, np.zeros(len(xr), F32), M, type=types)
    return dict(M=M, H=H, p=p, a=a, src=src, rec=rec, vp=vp, velocity=(vp, vs, rho),
                dt=dt, dx=dx, nbc=nbc, fd_order=fd_order, rec_types=tuple(rec_types))


def model_blocks(b, dm, M, H) -> int:
    """Pitched coefficient arrays the device model holds (5 stored; 7 when it
    also holds rho_vx and mu_txx for the derived-coefficient fused kernel)."""
    pitch = (M.nx + 272 + 255) // 256 * 256
    block = 4 * ((M.nz + 36) * pitch + 512)
    return round((int(b.lib.fomo_model_device_bytes(dm.h)) - 12 * (H.nbc + 2)) / block)


def cpu_baseline(case, steps: int, threads: int):
    """The oracle (CPU restatement of the reference's path) on the first
    `steps` steps of the same shot, OpenMP over `threads` threads."""
    import oracle as O
    from fomo_julia_amd import init as I
    M, H, p, src, rec = case["M"], case["H"], case["p"], case["src"], case["rec"]
    w = [I.habc_weight_matrix(M.nx, M.nz, H.nbc, off) for off in (0.0, 0.5, 0.75)]
    r = O.OracleRun(nx=M.nx, nz=M.nz, pad=M.pad, nbc=H.nbc, fd_order=p.fd_order,
                    free_surface=M.is_free_surface, dt=p.dt, dtx=p.dtx, dtz=p.dtz,
                    q=(H.qx, H.qz, H.qt_x, H.qt_z, H.qxt), a=case["a"], rho_vx=M.rho_vx,
                    rho_vz=M.rho_vz, lam=M.lam, mu_txx=M.mu_txx, mu_txz=M.mu_txz,
                    w_vx=w[0], w_vz=w[1], w_tau=w[2], src_i=src.i, src_j=src.j, src_type=0,
                    wavelet=src.wavelet[:steps], rec_i=rec.i, rec_j=rec.j, rec_type=0, nt=steps,
                    flags=O.FMA, threads=threads)
    t0 = time.perf_counter()
    r.run(1, steps)
    el = time.perf_counter() - t0
    cells = M.nx * M.nz * steps
    return cells / el / 1e6, el


def _oracle_for(case, steps: int, dtype, threads: int, storage: str = "f32", fscale=None):
    """The oracle on `steps` steps of a build_seam case."""
    import oracle as O
    from fomo_julia_amd import init as I
    M, H, p, src, rec = case["M"], case["H"], case["p"], case["src"], case["rec"]
    w = [I.habc_weight_matrix(M.nx, M.nz, H.nbc, off) for off in (0.0, 0.5, 0.75)]
    return O.OracleRun(nx=M.nx, nz=M.nz, pad=M.pad, nbc=H.nbc, fd_order=p.fd_order,
                       free_surface=M.is_free_surface, dt=p.dt, dtx=p.dtx, dtz=p.dtz,
                       q=(H.qx, H.qz, H.qt_x, H.qt_z, H.qxt), a=case["a"], rho_vx=M.rho_vx,
                       rho_vz=M.rho_vz, lam=M.lam, mu_txx=M.mu_txx, mu_txz=M.mu_txz,
                       w_vx=w[0], w_vz=w[1], w_tau=w[2], src_i=src.i, src_j=src.j, src_type=0,
                       wavelet=src.wavelet[:steps], rec_i=rec.i, rec_j=rec.j,
                       rec_type=rec.type_codes(), nt=steps, dtype=dtype, flags=O.FMA,
                       threads=threads, storage=storage, fscale=fscale)


def _component_errors(g, g64, types, n_col) -> dict:
    """Per receiver type: relative L2 and peak errors of gather g against g64."""
    out = {}
    for k, t in enumerate(types):
        a = np.asarray(g[:, k * n_col:(k + 1) * n_col], np.float64)
        r = np.asarray(g64[:, k * n_col:(k + 1) * n_col], np.float64)
        peak = float(np.abs(r).max())
        out[t] = {"rel_l2": float(np.linalg.norm(a - r) / np.linalg.norm(r)),
                  "max_abs_err_over_peak": float(np.abs(a - r).max() / peak),
                  "peak_amplitude_err": float((np.abs(a).max() - peak) / peak)}
    return out


def bench_precision(args) -> dict:
    """Each storage's error against the Float64 run of the same scheme on the
    SEAM-shaped shot (4444 x 3819, fd_order 8, nbc 50, free surface, 10 Hz
    Ricker pressure source at the surface centre): the first K steps (default
    400), receivers on every surface column recording vx, vz and pressure.
    GPU runs for FP32, F24 (fields rounded to a 16-bit significand, FP32
    coefficients), FP16 coefficients (FP32 fields), FP16 and BF16; the
    oracle's emulation of the block floating-point field format (no kernel)
    beside them.  Float64 is the oracle on the host's CPUs."""
    from fomo_julia_amd import elastic2d as E
    steps = args.steps if args.steps != 11520 else 400
    types = ("vx", "vz", "p")
    case = build_seam(steps, rec_types=types)
    n_col = case["vp"].shape[0]
    threads = cpu_share()["usable"]
    t0 = time.perf_counter()
    r64 = _oracle_for(case, steps, np.float64, threads)
    r64.run(1, steps)
    g64 = r64.gather_nt_by_nrec().copy()
    f64_s = time.perf_counter() - t0
    del r64
    b = E.backend("mi355x")
    b.set_tiling(2, 0, 16)
    vp, vs, rho = case["velocity"]
    out = {"metric": "gather error vs Float64 per component, SEAM-shaped shot, 1 GPU",
           "steps": steps, "components": list(types), "receivers_per_component": n_col,
           "float64_oracle_seconds": round(f64_s, 1), "oracle_threads": threads,
           "data": "synthetic (seeded SEAM-like medium)", "gpu": {}, "emulated": {}}
    for storage in ("f32", "f24", "f16c", "f16", "bf16"):
        Ms = E.init_medium(vp, vs, rho, case["dx"], case["dx"], case["nbc"], case["fd_order"], None,
                           free_surface=True, storage=storage)
        W = E.Wavefield(Ms.nx, Ms.nz, b)
        res = E.run_shot(b, W, Ms, case["H"], case["a"], case["src"], case["rec"], case["p"])
        out["gpu"][storage] = _component_errors(res.gather, g64, types, n_col)
        print(json.dumps({storage: out["gpu"][storage]}), file=sys.stderr, flush=True)
        W.close()
        b._models.clear()
    for storage in ("bfp16", "bfp16c"):
        r = _oracle_for(case, steps, np.float32, threads, storage=storage)
        r.run(1, steps)
        out["emulated"][storage] = _component_errors(r.gather_nt_by_nrec(), g64, types, n_col)
        print(json.dumps({storage: out["emulated"][storage]}), file=sys.stderr, flush=True)
        del r
    return out


def bench_batch(args) -> dict:
    """Batched shots: 64 shots on the Marmousi-shaped 1700 x 350 grid
    (dx = dz = 10 m, fd_order 12, nbc 40, free surface, nt 4000, 25 Hz Ricker
    pressure sources evenly spaced on the surface, 1700 surface vz receivers),
    all in one persistent launch per kernel with captured steps and an
    autotuned tile shape.  Timed: the batch's nt steps (tuning excluded)."""
    import torch
    from fomo_julia_amd import elastic2d as E
    from fomo_julia_amd import init as I
    from fomo_julia_amd import media
    from fomo_julia_amd import tune as T
    F32 = np.float32
    n_shots, nt = args.shots, args.steps if args.steps != 11520 else 4000
    dx = F32(10.0)
    vp, vs, rho = media.marmousi_like()
    dt = media.cfl_dt(vp, dx, 0.4)
    M = E.init_medium(vp, vs, rho, dx, dx, 40, 12, None, free_surface=True, storage=args.storage)
    H = E.init_habc(M.nx, M.nz, 40, dt, dx, dx, I.julia_mean_f32(vp), None, weights=False)
    p = E.SimParams(dt, nt, dx, dx, 12)
    a = I.get_fd_coefficients(12)
    wav = I.ricker_wavelet(F32(25.0), dt, nt)
    b = E.backend("mi355x")
    split = args.split not in (None, "", "off")
    double = split and args.double and args.storage == "f32"
    if split:
        er, bpc = (int(v) for v in args.split.split(","))
        b.set_step_mode(3 if double else 2)
        b.set_split(er, bpc)
    W = E.Wavefield(M.nx, M.nz, b)
    dm = E.bind(b, W, M, H, a, p)
    nx_in = vp.shape[0]
    rec = E.setup_receivers(np.arange(nx_in, dtype=F32) * dx, np.zeros(nx_in, F32), M, type="vz")
    xs = ((np.arange(n_shots) + 0.5) * (float(M.x_max) / n_shots)).astype(F32)
    wfs, dss = [], []
    for x in xs:
        src = E.Source(I.coord_to_index(F32(x), M.dx, M.pad), I.coord_to_index(F32(0.0), M.dz, M.pad),
                       wav, "pressure")
        w = E.Wavefield(M.nx, M.nz, b)
        w._bind(dm)
        wfs.append(w)
        dss.append(E._DeviceShot(dm, src, rec, nt))
    tuning = T.autotune_batch(b, [dm] * n_shots, wfs, dss)
    lib = b.lib
    arr = lambda hs: (ctypes.c_void_p * n_shots)(*[h.value for h in hs])
    bh = ctypes.c_void_p()
    b.check(lib.fomo_batch_create(b.ctx, n_shots, arr([dm.h] * n_shots), arr([w._h for w in wfs]),
                                  arr([d.h for d in dss]), ctypes.byref(bh)), "batch_create")
    stream = torch.cuda.ExternalStream(b.stream_handle())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    b.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    b.check(lib.fomo_batch_run(bh, 1, nt), "batch_run")
    e1.record(stream)
    b.synchronize()
    gpu_s = e0.elapsed_time(e1) / 1e3
    gathers = [d.gather() for d in dss]
    wall = time.perf_counter() - t0
    lib.fomo_batch_destroy(bh)
    cells = M.nx * M.nz * nt * n_shots
    serial = None
    if args.serial_shots > 0:
        del wfs, dss
        serial = bench_batch_serial(b, (vp, vs, rho), dx, dt, xs, nt, args.serial_shots)
    return {"metric": "shots/hour and Mcell-updates/s, batched shots, 1 GPU",
            "value": round(n_shots / gpu_s * 3600.0, 1), "unit": "shots/hour", "n_gpus": 1,
            "steps": nt, "higher_is_better": True, "dtype": "f32", "storage": args.storage,
            "data": "synthetic (seeded Marmousi-like medium)",
            "config": {"workload": f"{n_shots} batched shots, Marmousi-shaped 1700x350, dx=dz=10 m, "
                                   f"fd_order=12, nbc=40, nt={nt}, 25 Hz Ricker, 1700 receivers, "
                                   "one persistent launch per kernel, captured steps",
                       "padded_cells_per_shot": M.nx * M.nz,
                       "step_mode": ("double, split edges " + args.split) if double else
                                    ("split fused " + args.split) if split else "fused"},
            "mcell_updates_per_s": round(cells / gpu_s / 1e6, 1),
            "ms_per_step": round(gpu_s * 1e3 / nt, 4), "gpu_seconds": round(gpu_s, 3),
            "wall_seconds_with_gathers": round(wall, 3),
            "tiling": {"cols_per_lane": tuning["tiling"][0], "rows_per_tile": tuning["tiling"][1],
                       "trials_ms_per_step": tuning["trials"]},
            "gathers_finite": bool(all(np.isfinite(g).all() for g in gathers)),
            "serial_f32": serial,
            "speedup_vs_serial_f32": (round(n_shots / gpu_s * 3600.0 / serial["shots_per_hour"], 3)
                                      if serial else None)}


def bench_batch_serial(b, velocity, dx, dt, xs, nt: int, n_timed: int) -> dict:
    """The plain FP32 GPU path on the same shots: run_shots one shot at a time
    (the reference's sequential loop, shot_manager.jl:157-176) on one
    wavefield, FP32 storage, tile shape autotuned for the single shot.  One
    untimed shot first (graph capture); then `n_timed` shots, synchronized
    wall time, which gives seconds per shot and shots/hour."""
    import torch  # noqa: F401  (device initialisation as in the other workloads)
    from fomo_julia_amd import elastic2d as E
    from fomo_julia_amd import init as I
    from fomo_julia_amd import tune as T
    F32 = np.float32
    vp, vs, rho = velocity
    M = E.init_medium(vp, vs, rho, dx, dx, 40, 12, None, free_surface=True, storage="f32")
    H = E.init_habc(M.nx, M.nz, 40, dt, dx, dx, I.julia_mean_f32(vp), None, weights=False)
    p = E.SimParams(dt, nt, dx, dx, 12)
    a = I.get_fd_coefficients(12)
    wav = I.ricker_wavelet(F32(25.0), dt, nt)
    nx_in = vp.shape[0]
    rec = E.setup_receivers(np.arange(nx_in, dtype=F32) * dx, np.zeros(nx_in, F32), M, type="vz")
    W = E.Wavefield(M.nx, M.nz, b)
    dm = E.bind(b, W, M, H, a, p)
    src_t = E.Source(I.coord_to_index(F32(xs[0]), M.dx, M.pad), I.coord_to_index(F32(0.0), M.dz, M.pad),
                     I.ricker_wavelet(F32(25.0), dt, TIMING_NT), "pressure")
    tuning = T.autotune_batch(b, [dm], [W], [E._DeviceShot(dm, src_t, rec, TIMING_NT)])
    n = min(n_timed + 1, len(xs))
    cfg = E.MultiShotConfig(xs[:n], np.zeros(n, F32), wav, source_type="pressure")
    times = []
    E.run_shots(b, W, M, H, a, rec, E.MultiShotConfig(xs[:1], np.zeros(1, F32), wav), p, batch=1)
    b.synchronize()
    t_last = [time.perf_counter()]

    def done(res):
        b.synchronize()
        t = time.perf_counter()
        times.append(t - t_last[0])
        t_last[0] = t

    E.run_shots(b, W, M, H, a, rec, E.MultiShotConfig(xs[1:n], np.zeros(n - 1, F32), wav), p,
                batch=1, on_shot_complete=done)
    per_shot = float(np.mean(times))
    return {"shots_timed": len(times), "seconds_per_shot": round(per_shot, 4),
            "shots_per_hour": round(3600.0 / per_shot, 1),
            "ms_per_step": round(per_shot * 1e3 / nt, 4),
            "tiling": {"cols_per_lane": tuning["tiling"][0], "rows_per_tile": tuning["tiling"][1]},
            "note": "one shot at a time through run_shots(batch=1): run_shot on one wavefield, "
                    "gather downloaded after each shot; synchronized wall time per shot; "
                    "the batch's step mode"}


def bench_edges(args) -> dict:
    """Two edges of the path on the SEAM-shaped FP32 shot (4444 x 3819, nbc 50,
    fd_order 8):
    - medium setup, wall time to a ready device model (synchronized, best of
      3): init_medium on the host plus the upload of its padded coefficient
      arrays, against init_medium_device (upload of the unpadded vp, vs, rho;
      k_setup pads them and forms the coefficients on the device);
    - on-step frames: the replayed loop writing a 4x-decimated pressure frame
      every 10 steps into a device ring (fomo_run_time_loop_frames), against
      the plain loop; HIP-event time per step on the library's stream, both
      from step 1 of a reset shot, 2 columns per lane, with the whole-grid
      step and with the double step (whose frames are taken after the
      second step of each pair).  Draining the ring to the host is not
      timed."""
    import torch
    from fomo_julia_amd import elastic2d as E
    from fomo_julia_amd import init as I
    from fomo_julia_amd import media
    from fomo_julia_amd import video as V
    F32 = np.float32
    steps = args.steps if args.steps != 11520 else 2000
    skip, dsamp = 10, 4
    b = E.backend("mi355x")
    lib = b.lib
    vp, vs, rho = media.seam_like()
    dx, nbc, order = F32(12.5), 50, 8
    dt = media.cfl_dt(vp, dx, 0.4)
    v_ref = I.julia_mean_f32(vp)
    p = E.SimParams(dt, steps, dx, dx, order)
    a = I.get_fd_coefficients(order)
    setup = {}
    for name, init in (("host", E.init_medium), ("device", E.init_medium_device)):
        best = None
        for _ in range(3):
            b._models.clear()
            b.synchronize()
            t0 = time.perf_counter()
            M = init(vp, vs, rho, dx, dx, nbc, order, None, free_surface=True)
            H = E.init_habc(M.nx, M.nz, nbc, dt, dx, dx, v_ref, None, weights=False)
            E._model(b, M, H, a, p)
            b.synchronize()
            t = time.perf_counter() - t0
            best = t if best is None else min(best, t)
        setup[name] = round(best, 3)
    # the device-built model (the last one created) runs the shot, at the
    # tiling the SEAM autotune picks (2 columns per lane, automatic rows),
    # with the whole-grid step and with the double step (split edges 16,1)
    wav = I.ricker_wavelet(F32(10.0), dt, steps)
    rec = E.setup_receivers(np.arange(vp.shape[0], dtype=F32) * dx, np.zeros(vp.shape[0], F32), M,
                            type="vz")
    stream = torch.cuda.ExternalStream(b.stream_handle())

    def frames_cost(step_mode: int) -> dict:
        b.set_tiling(2, 0, 16)
        b.set_step_mode(step_mode)
        b.set_split(16, 1 if step_mode >= 2 else 0)
        try:
            W = E.Wavefield(M.nx, M.nz, b)
            dm = E.bind(b, W, M, H, a, p)
            src = E.Source(I.coord_to_index(F32(M.x_max / F32(2)), M.dx, M.pad),
                           I.coord_to_index(F32(0.0), M.dz, M.pad), wav, "pressure")
            ds = E._shot_for(dm, src, rec, steps)
            store = V._FrameStore(b, M.nx, M.nz, [V.FIELDS["p"]], skip, dsamp, steps // skip)

            def run(frames: bool, k0: int, n: int):
                if frames:
                    b.check(lib.fomo_run_time_loop_frames(dm.h, W._h, ds.h, k0, n, store.h), "loop with frames")
                else:
                    b.check(lib.fomo_run_time_loop(dm.h, W._h, ds.h, k0, n), "loop")

            per_step = {}
            for frames in (False, True):
                run(frames, 1, 64)  # warm-up: graph capture
                E.reset(b, W)
                b.check(lib.fomo_shot_reset(ds.h), "shot reset")
                b.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                run(frames, 1, steps)
                e1.record(stream)
                b.synchronize()
                per_step["frames" if frames else "plain"] = e0.elapsed_time(e1) / steps
            frame = store.download(b, steps // skip - 1, 1)
            return {"field": "p", "skip": skip, "downsample": dsamp, "steps": steps,
                    "step_mode": {1: "whole-grid", 3: "double, split edges 16,1"}[step_mode],
                    "frame_shape": list(store.shape[1:]),
                    "ms_per_step_plain": round(per_step["plain"], 4),
                    "ms_per_step_with_frames": round(per_step["frames"], 4),
                    "overhead": round(per_step["frames"] / per_step["plain"] - 1.0, 4),
                    "last_frame_finite_nonzero": bool(np.isfinite(frame).all() and np.abs(frame).max() > 0)}
        finally:
            b.set_step_mode(1)
            b.set_split(16, 0)

    return {"metric": "edges of the path, SEAM-shaped FP32 shot, 1 GPU",
            "setup_seconds": setup,
            "setup_speedup": round(setup["host"] / setup["device"], 2),
            "frames": frames_cost(1),
            "frames_double": frames_cost(3),
            "data": "synthetic (seeded SEAM-like medium)"}


def bench_serve(n_req: int, seed: int = 5, load: float = 0.8, storages=("f32", "bf16")) -> dict:
    """Served stream (requests with grids nx 300-2000, nz 300-1200, nt
    1000-6000, fd_order 4/8/12, storage alternating over `storages`: FP32 and
    BF16 by default): serial plain FP32 execution, then the continuous-batching
    server with every request queued at once (capacity), then with Poisson
    arrivals at `load` times the serial service rate (latency)."""
    from fomo_julia_amd import elastic2d as E
    from fomo_julia_amd import serve as S
    b = E.backend("mi355x")
    reqs = S.c5_stream(n_req, seed=seed, storages=tuple(storages))
    every = max(1, n_req // 20)

    def progress(phase):
        count = [0]
        t_start = time.perf_counter()

        def cb(c):
            count[0] += 1
            if count[0] % every == 0 or count[0] == n_req:
                print(f"served sample, {phase}: {count[0]}/{n_req} shots done, "
                      f"{time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)
        return cb

    t = time.perf_counter()
    prep_times = []
    ser = S.serial(reqs, b, storage="f32", prep_times=prep_times, on_complete=progress("serial FP32"))
    w_ser = time.perf_counter() - t
    w_ser_gpu = w_ser - sum(prep_times)  # the serial loop without its host preparation
    srv = S.Server(b, chunk=64, capacity=64)
    t = time.perf_counter()
    sat = srv.serve(reqs, realtime=False, on_complete=progress("capacity"))
    w_sat = time.perf_counter() - t
    rate = load / (w_ser / n_req)
    srv2 = S.Server(b, chunk=64, capacity=64)
    t = time.perf_counter()
    poi = srv2.serve(S.with_arrivals(reqs, rate), realtime=True, on_complete=progress("Poisson"))
    w_poi = time.perf_counter() - t
    s_ser, s_sat, s_poi = S.summarize(ser, w_ser), S.summarize(sat, w_sat), S.summarize(poi, w_poi)
    return {"requests": n_req, "storages": list(storages),
            "serial_f32": {**s_ser, "prepare_s": round(sum(prep_times), 3)},
            "served_capacity": s_sat,
            "served_poisson": {**s_poi, "arrival_rate_per_s": round(rate, 3), "load": load},
            "speedup_shots_per_hour": round(s_sat["shots_per_hour"] / s_ser["shots_per_hour"], 3),
            # against the serial loop with its host preparation taken out (the
            # server prepares on worker threads, overlapped with the GPU)
            "speedup_vs_serial_without_prepare": round(w_ser_gpu / w_sat, 3),
            "server_stats": {k: (round(v, 3) if isinstance(v, float) else v)
                             for k, v in srv.stats.items()},
            "failed_requests": [f.__dict__ for f in srv.failed + srv2.failed]}


def pattern_probe(reps: int = 20) -> dict:
    """Bandwidth of the fused step's access pattern with no stencil arithmetic
    (csrc/pattern_probe.hip) and of plain streaming, on a 4608 x 3920 grid of
    pitched rows (the SEAM shot's padded 4552 x 3927 rounded to whole tiles).
    nread = 8 is the fused step's FP32 stream set (5 fields + rho_vx, lam,
    mu_txx), 12 the set without derived coefficients; 5 fields are written.
    An LDS reservation of 48 KiB per 4-wave block caps residency at 3 blocks
    per CU = 3 waves per SIMD, the fused step's own occupancy (150 VGPRs)."""
    import torch
    from fomo_julia_amd import build as B
    if not os.path.exists(B.PROBE_PATH):
        B.build_probe()
    lib = ctypes.CDLL(B.PROBE_PATH)
    lib.fomo_probe_pattern.argtypes = [ctypes.c_int] * 9 + [ctypes.POINTER(ctypes.c_double)]
    lib.fomo_probe_stream.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]
    lib.fomo_probe_runtime.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double)]
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    us = ctypes.c_double()
    out = {"cus": cus, "stream": [], "runtime": [], "pattern": []}
    for nbytes in (1 << 30, 4 << 30):
        for mode, name, moved in ((0, "copy", 2), (1, "read", 1), (2, "write", 1)):
            for grid in (cus * 8, cus * 32):
                rc = lib.fomo_probe_stream(mode, nbytes, grid, reps, ctypes.byref(us))
                if rc:
                    raise RuntimeError(f"fomo_probe_stream({name}, grid {grid}) returned {rc}")
                gbs = moved * nbytes / (us.value * 1e3)
                out["stream"].append({"kind": name, "bytes": nbytes, "grid": grid,
                                      "us": round(us.value, 1), "GB/s": round(gbs, 1),
                                      "frac": round(gbs / HBM_PEAK_GBS, 3)})
                print(json.dumps(out["stream"][-1]), file=sys.stderr, flush=True)
    # the runtime's own copy and fill over 4 GiB: points that do not depend
    # on the probe kernels above
    nbytes = 4 << 30
    for mode, name, moved in ((0, "hipMemcpyAsync D2D", 2), (1, "hipMemsetAsync", 1),
                              (2, "hipMemsetD32Async", 1)):
        rc = lib.fomo_probe_runtime(mode, nbytes, reps, ctypes.byref(us))
        if rc:
            raise RuntimeError(f"fomo_probe_runtime({name}) returned {rc}")
        gbs = moved * nbytes / (us.value * 1e3)
        out["runtime"].append({"kind": name, "bytes": nbytes, "us": round(us.value, 1),
                               "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 3)})
        print(json.dumps(out["runtime"][-1]), file=sys.stderr, flush=True)
    nx, nz, tz = 4608, 3920, 40
    for nread in (8, 12):
        for cpt in (1, 2, 4):
            for lds, grid in ((0, cus * 8), (0, cus * 16), (49152, cus * 3), (49152, cus * 12)):
                rc = lib.fomo_probe_pattern(nread, 5, cpt, nx, nz, tz, grid, lds, reps, ctypes.byref(us))
                if rc:
                    raise RuntimeError(f"fomo_probe_pattern({nread}, 5, {cpt}, grid {grid}, "
                                       f"lds {lds}) returned {rc}")
                gbs = (nread + 5) * 4 * nx * nz / (us.value * 1e3)
                out["pattern"].append({"nread": nread, "nwrite": 5, "cols_per_lane": cpt, "grid": grid,
                                       "lds": lds, "us": round(us.value, 1), "GB/s": round(gbs, 1),
                                       "frac": round(gbs / HBM_PEAK_GBS, 3)})
                print(json.dumps(out["pattern"][-1]), file=sys.stderr, flush=True)
    for key in ("stream", "runtime", "pattern"):
        best = max(out[key], key=lambda r: r["GB/s"])
        out[f"best_{key}"] = best
    capped = [r for r in out["pattern"] if r["nread"] == 8 and r["lds"]]
    out["best_pattern_8r5w_3_waves_per_simd"] = max(capped, key=lambda r: r["GB/s"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=11520)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=80, help="steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stage-reps", type=int, default=50)
    ap.add_argument("--tiling", default="", help="cols_per_lane,rows_per_tile,steps_per_graph")
    ap.add_argument("--two-pass", action="store_true", help="time the two-pass in-place step instead")
    ap.add_argument("--split", default=None, metavar="EDGE_ROWS,BLOCKS_PER_CU",
                    help="split fused step: the interior tiles beside the edge tiles and the strip "
                         "kernels (rows per edge tile, interior blocks per CU, 0 = no cap); 'off' "
                         "runs the whole-grid fused step.  Default: 16,1 for the single shot "
                         "(16,2 with F24 storage), off for the batch (whose edge tiles and strip "
                         "kernels cover a third of its cells: split, its step is slower)")
    ap.add_argument("--double", default=True, action=argparse.BooleanOptionalAction,
                    help="double step (step mode 3, the default with FP32 storage): the interior "
                         "advanced two time steps per pass, the split step's edge chain for each "
                         "step beside it; --no-double runs the split step (or --split off, the "
                         "whole-grid step)")
    ap.add_argument("--storage", default="f32", choices=sorted(BYTES_FUSED_STORAGE),
                    help="device storage of coefficients/fields (arithmetic stays FP32)")
    ap.add_argument("--workload", default="seam",
                    choices=("seam", "batch", "serve", "edges", "precision"),
                    help="seam: the single SEAM-shaped shot (default); batch: 64 batched "
                         "Marmousi-shaped shots; serve: the served request stream; edges: "
                         "device-side setup and on-step frames on the SEAM-shaped shot; "
                         "precision: each storage's gather error vs Float64 on the SEAM shot")
    ap.add_argument("--shots", type=int, default=64, help="shots of the batch workload")
    ap.add_argument("--serial-shots", type=int, default=4,
                    help="batch workload: shots timed one at a time on the plain FP32 path")
    ap.add_argument("--requests", type=int, default=40,
                    help="requests of the served sample (default run) or stream (serve)")
    ap.add_argument("--no-serve", action="store_true", help="skip the served sample")
    ap.add_argument("--serve-storages", default="f32,bf16",
                    help="storages the served requests alternate over (f32,bf16: the stream as "
                         "specified; f32,f24: its reduced half in the F24 form within 1e-3)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    ap.add_argument("--pattern-probe", action="store_true",
                    help="only measure the bandwidth ceiling of the fused step's access "
                         "pattern and of plain streaming (libfomo_probe.so)")
    args = ap.parse_args()
    if args.pattern_probe:
        print(json.dumps({"probe": "hbm access-pattern ceiling", **pattern_probe()}))
        return
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 18  # float32 elements
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.workload == "edges":
        print(json.dumps(bench_edges(args)))
        return
    if args.workload == "precision":
        print(json.dumps(bench_precision(args)))
        return
    if args.workload == "batch":
        print(json.dumps(bench_batch(args)))
        return
    if args.workload == "serve":
        print(json.dumps({"metric": "shots/hour + p50/p99 latency (served), 1 GPU",
                          **bench_serve(args.requests, storages=args.serve_storages.split(","))}))
        return

    import torch
    from fomo_julia_amd import elastic2d as E

    nt = max(args.steps, args.warmup, 1)
    t_setup = time.perf_counter()
    if args.two_pass and args.storage != "f32":
        sys.exit("bench.py: the two-pass step exists for FP32 storage only")
    case = build_seam(nt, storage=args.storage)
    M, H, p, src, rec = case["M"], case["H"], case["p"], case["src"], case["rec"]
    t_setup = time.perf_counter() - t_setup
    cells = M.nx * M.nz

    b = E.backend("mi355x")
    fused = not args.two_pass
    if args.split is None:
        # F24's interior kernel gains more from a second block per CU than the
        # strip kernels beside it lose (0.2665 against 0.277 ms on SEAM)
        args.split = "16,2" if args.storage == "f24" else "16,1"
    split = fused and args.split not in ("", "off")
    double = split and args.double and args.storage == "f32"
    b.set_step_mode(3 if double else 2 if split else 1 if fused else 0)
    if split:
        er, bpc = (int(v) for v in args.split.split(","))
        b.set_split(er, bpc)
    W = E.Wavefield(M.nx, M.nz, b)
    dm = E.bind(b, W, M, H, case["a"], p)
    ds = E._shot_for(dm, src, rec, p.nt)
    # the timing shot: TIMING_NT steps of the same source and receivers
    from fomo_julia_amd import init as I
    src_t = E.Source(src.i, src.j, I.ricker_wavelet(np.float32(10.0), p.dt, TIMING_NT), src.type)
    ds_t = E._DeviceShot(dm, src_t, rec, TIMING_NT)
    lib = b.lib
    stream = torch.cuda.ExternalStream(b.stream_handle())
    tuning = None
    if args.tiling:
        cpt, tz, gs = (int(v) for v in args.tiling.split(","))
        b.set_tiling(cpt, tz, gs)
    elif fused:
        from fomo_julia_amd import tune as T
        tuning = T.autotune_batch(b, [dm], [W], [ds_t])

    # warm-up (clocks, first launches), then the step graphs of the timed
    # run are captured before timing starts
    if args.warmup:
        b.check(lib.fomo_run_time_loop(dm.h, W._h, ds.h, 1, args.warmup), "warmup")
    b.synchronize()
    E.reset(b, W)
    b.check(lib.fomo_shot_reset(ds.h), "shot_reset")
    b.check(lib.fomo_run_time_loop_prepare(dm.h, W._h, ds.h), "graph capture")
    b.synchronize()
    torch.cuda.synchronize()

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    b.check(lib.fomo_run_time_loop(dm.h, W._h, ds.h, 1, args.steps), "timed loop")
    ev1.record(stream)
    b.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / max(args.steps, 1)
    value = cells * args.steps / (gpu_ms / 1e3) / 1e6
    t_g = time.perf_counter()
    gather = ds.gather()  # device -> host, [nt x n_rec]
    gather_s = time.perf_counter() - t_g
    finite = bool(np.isfinite(gather).all())
    # the boundary's host hand-over on either side of the timed region: the
    # medium as three unpadded arrays (device-side setup), the gather back
    t_u = time.perf_counter()
    dm_up = E._DeviceModel(b, E.init_medium_device(*case["velocity"], M.dx, M.dz, H.nbc,
                                                   p.fd_order, None, storage=args.storage),
                           H, case["a"], p)
    b.synchronize()
    upload_s = time.perf_counter() - t_u
    del dm_up

    # per-kernel mean durations (HIP events on the library's stream), on the
    # timing shot at a step inside its nt
    k_stage = stage_step(TIMING_NT)

    def stage_ms(which: int, reps: int) -> float:
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        b.check(lib.fomo_launch_stage(dm.h, W._h, ds_t.h, which, k_stage, 2), "stage warm")
        b.synchronize()
        e0.record(stream)
        b.check(lib.fomo_launch_stage(dm.h, W._h, ds_t.h, which, k_stage, reps), "stage")
        e1.record(stream)
        b.synchronize()
        return e0.elapsed_time(e1) / reps

    reps = args.stage_reps
    main_cells = cells
    if fused:
        main_ms = stage_ms(5, reps)
        kernels = {"fused": main_ms, "band_stress": stage_ms(6, reps),
                   "habc_velocity": stage_ms(7, reps), "habc_stress": stage_ms(8, reps),
                   "fs_record": stage_ms(4, reps)}
        main_bytes = BYTES_FUSED_STORAGE[args.storage]
        form = "stored coefficients"
        if args.storage in BYTES_FUSED_DERIVED_STORAGE and model_blocks(b, dm, M, H) == 7:
            main_bytes = BYTES_FUSED_DERIVED_STORAGE[args.storage]
            form = "coefficients formed from rho_vx, lam, mu_txx"
        main_name = f"k_fused, {form} ({main_bytes} B/cell-update x padded cells per launch)"
        if split:
            # the split step runs the fused pass as edge tiles, then interior tiles
            kernels["fused_edge"] = stage_ms(9, reps)
            kernels["fused_interior"] = main_ms = stage_ms(10, reps)
            main_cells = split_interior_cells(M.nx, M.nz, H.nbc, p.fd_order // 2)
            main_name = (f"k_fused over the split step's interior tiles, {form} ({main_bytes} "
                         f"B/cell-update x {main_cells} interior cells per launch)")
        if double:
            # two steps per launch over the interior, with the five stored
            # coefficient arrays (the library's default form for k_fused2):
            # one read of 5 fields + 5 coefficients and one write of the 5
            # fields per cell and launch
            kernels["dbl_interior_2steps"] = main_ms = stage_ms(11, reps)
            kernels["dbl_edge_step_k"] = stage_ms(12, reps)
            kernels["dbl_edge_step_k1"] = stage_ms(13, reps)
            main_bytes = BYTES_FUSED
            main_name = (f"k_fused2 over the double step's interior tiles (two time steps per "
                         f"launch), stored coefficients ({main_bytes} B per cell and launch x "
                         f"{main_cells} interior cells)")
    else:
        main_ms = stage_ms(1, reps)
        kernels = {"velocity": stage_ms(0, reps), "stress": main_ms,
                   "habc_velocity": stage_ms(2, reps), "habc_stress": stage_ms(3, reps),
                   "fs_record": stage_ms(4, reps)}
        main_bytes, main_name = BYTES_STRESS, "k_stress (44 B/cell-update x padded cells per launch)"
    main_gbs = main_bytes * main_cells / (main_ms / 1e3) / 1e9
    # the step as a whole on the single-step model of the medium's form
    step_bytes = (BYTES_FUSED_DERIVED if double else main_bytes) if fused else BYTES_VELOCITY + BYTES_STRESS
    step_gbs = step_bytes * cells / (ms_per_step / 1e3) / 1e9
    check_roofline(main_gbs / HBM_PEAK_GBS, main_ms, ms_per_step, 2 if double else 1)
    traffic = None
    tfile = TRAFFIC_FILE_DOUBLE if double else TRAFFIC_FILE_SPLIT if split else TRAFFIC_FILE
    if fused and args.storage == "f32" and os.path.exists(tfile) and \
            (not split or args.split == "16,1"):
        for line in open(tfile):
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = int(line.split(":")[1])

    # the served sample runs with the SEAM shot's device objects released and
    # before the CPU baseline (whose OpenMP threads would share the host cores
    # with the server's worker threads)
    del W, dm, ds, ds_t
    import gc
    gc.collect()
    b.synchronize()
    served = None
    if not args.no_serve and args.requests > 0:
        try:
            served = bench_serve(args.requests)
        except Exception as e:  # reported, never fatal
            served = {"error": str(e)}

    cpu = None
    if not args.no_cpu:
        share = cpu_share()
        threads = share["usable"]
        try:
            v, el = cpu_baseline(case, args.cpu_steps, threads)
            cpu = {"value": round(v, 2), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
                   "host_cpus": share,
                   "sample": f"first {args.cpu_steps} steps of the same SEAM-shaped shot "
                             f"({M.nx}x{M.nz} padded cells/step), C++ oracle, OpenMP stencils, "
                             f"sequential HABC, {el:.1f} s; full shot extrapolated "
                             f"{el / args.cpu_steps * 11520 / 60:.1f} min; OpenMP threads = "
                             "the CPUs this process may use (affinity, cgroup quota)"}
        except Exception as e:  # reported, never fatal
            cpu = {"value": None, "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
                   "host_cpus": share, "sample": f"failed: {e}"}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "storage": args.storage,
        "data": "synthetic (seeded SEAM-like medium)",
        "config": {"workload": f"SEAM-shaped single shot, {args.storage.upper()} storage, FP32 "
                               "arithmetic: nx=4444 nz=3819 dx=dz=12.5 m, "
                               "fd_order=8, nbc=50, free surface, padded 4552x3927, "
                               f"nt={args.steps}, 10 Hz Ricker pressure source at the surface "
                               "centre, 4444 vz receivers",
                   "padded_cells": cells, "physical_cells": 4444 * 3819,
                   "shot_seconds": round(gpu_ms / 1e3 * 11520 / max(args.steps, 1), 3),
                   "wall_seconds_timed": round(wall, 3), "setup_seconds": round(t_setup, 2),
                   "gather_finite": finite,
                   "pcie_inclusive": {
                       "model_upload_and_device_setup_s": round(upload_s, 4),
                       "gather_download_s": round(gather_s, 4),
                       "shot_seconds": round(gpu_ms / 1e3 * 11520 / max(args.steps, 1)
                                             + upload_s + gather_s, 3)}},
        "roofline": {"bound": "hbm", "achieved": round(main_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(main_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": (f"profiles/{os.path.basename(tfile)}: rocprofv3 --pmc "
                                        "FETCH_SIZE and WRITE_SIZE per launch, scaled by a plain "
                                        "copy's counts (not measured in this run)") if traffic else None,
                     "kernel": main_name, "kernel_ms": round(main_ms, 5),
                     "kernel_timing": f"HIP events around {reps} back-to-back launches on the "
                                      f"library's stream, step {k_stage} of a {TIMING_NT}-step shot",
                     "algorithmic_bytes_per_launch": main_bytes * main_cells},
        "kernels_ms": {k: round(v, 5) for k, v in kernels.items()},
        "tiling": ({"cols_per_lane": tuning["tiling"][0], "rows_per_tile": tuning["tiling"][1],
                    "steps_per_graph": b.tiling[2],
                    "autotuned": True, "trials_ms_per_step": tuning["trials"]} if tuning else
                   {"fixed": args.tiling or "library default"}),
        "step_mode": ("double, split edges " + args.split) if double else
                     ("split fused " + args.split) if split else "fused" if fused else "two-pass",
        "step_gbs": round(step_gbs, 1),
        "step_frac_hbm": round(step_gbs / HBM_PEAK_GBS, 4),
        "step_bytes_model": f"{step_bytes} B per cell-update (one step's compulsory traffic) x "
                            "padded cells / time per step",
        "cpu_baseline": cpu,
        "served": served,
        "reference_published": {"cpu_wall_min": 35, "rtx3060_wall_min_upper": 3,
                                "derived_cpu_mcells": 98.1, "derived_rtx3060_mcells_lower": 1144,
                                "note": "README wall times for one SEAM shot, other hardware"},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
