"""Parity at the benchmark's full size: the SEAM-shaped grid (4444 x 3819,
nbc 50, order 8, padded 4552 x 3927), the benchmark's kernel and tile shape
(two columns per lane, coefficients formed in the kernel for FP32, paired
stores), replayed HIP graphs.  Every field starts from a random state, so
every HABC piece, corner, band item and free-surface row is active from the
first step; after 20 steps the five fields and the gather equal the oracle's
bit for bit (the oracle runs the same 20 steps on the host, about a second;
for reduced storage it emulates the storage's rounding at every field write
and its dithered coefficients).  Each storage of the precision form runs:
FP32, BF16, FP16 (fields at the shot's power-of-two scales) and FP16
coefficients with FP32 fields."""
import numpy as np
import pytest

from fomo_julia_amd import elastic2d as E
from fomo_julia_amd import init as I
from fomo_julia_amd import media
from tests.helpers import REC_CODES, SRC_CODES

pytestmark = pytest.mark.gpu
F32 = np.float32
NT = 20


def _bf16_rne(v):
    """v rounded to BF16 (nearest, ties to even), as FP32 values."""
    u = v.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(F32)


def _f16_rne(v, scale):
    """v stored as FP16 at `scale` (x * scale rounded to nearest even), as FP32 values."""
    return (np.float16(v.astype(np.float64) * scale).astype(np.float64) / scale).astype(F32)


def _f24_rne(v):
    """v stored as F24 (a 16-bit significand, nearest, ties to even), as FP32 values."""
    u = v.view(np.uint32).astype(np.uint64)
    return ((u + 0x7F + ((u >> 8) & 1)) & 0xFFFFFF00).astype(np.uint32).view(F32)


@pytest.mark.parametrize("storage,mode", [("f32", 1), ("bf16", 1), ("f16", 1), ("f16c", 1),
                                          ("f24", 1), ("f32", 2), ("bf16", 2), ("f24", 2),
                                          ("f32", 3)])
def test_seam_random_state_bit_exact(storage, mode):
    import oracle as O
    dx = F32(12.5)
    vp, vs, rho = media.seam_like()
    dt = media.cfl_dt(vp, dx, 0.4)
    nbc, order = 50, 8
    M = E.init_medium(vp, vs, rho, dx, dx, nbc, order, None, free_surface=True, storage=storage)
    H = E.init_habc(M.nx, M.nz, nbc, dt, dx, dx, I.julia_mean_f32(vp), None)
    p = E.SimParams(dt, NT, dx, dx, order)
    a = I.get_fd_coefficients(order)
    src = E.Source(I.coord_to_index(F32(M.x_max / F32(2)), M.dx, M.pad),
                   I.coord_to_index(F32(0.0), M.dz, M.pad), I.ricker_wavelet(F32(10.0), dt, NT),
                   "pressure")
    xr = np.arange(0, vp.shape[0], 7, dtype=F32) * dx
    types = [("vz", "vx", "p")[k % 3] for k in range(len(xr))]
    rec = E.setup_receivers(xr, np.full(len(xr), 2 * dx, F32), M, type=types)
    print(f"case SEAM random state {storage} mode {mode} {M.nx}x{M.nz}, {NT} steps, "
          f"{len(xr)} receivers", flush=True)

    b = E.backend("mi355x")
    b.set_tiling(2, 0, 16)
    b.set_step_mode(mode)
    if mode == 3:  # the benchmark's double step: one interior block per CU
        b.set_split(16, 1)
    W = E.Wavefield(M.nx, M.nz, b)
    dm = E.bind(b, W, M, H, a, p)
    scales = np.ones(5, F32)
    if storage == "f16":
        # a step from the zero state makes the wavefield adopt the shot's FP16
        # field scales; the random state below is then stored at those scales
        ds = E._shot_for(dm, src, rec, NT)
        b.check(b.lib.fomo_time_step(dm.h, W._h, ds.h, 1), "scale adoption step")
        b.synchronize()
        scales = ds.field_scales()
        assert all(float(x) == 2.0 ** np.round(np.log2(float(x))) for x in scales)
    r = O.OracleRun(nx=M.nx, nz=M.nz, pad=M.pad, nbc=nbc, fd_order=order, free_surface=True,
                    dt=p.dt, dtx=p.dtx, dtz=p.dtz, q=(H.qx, H.qz, H.qt_x, H.qt_z, H.qxt), a=a,
                    rho_vx=M.rho_vx, rho_vz=M.rho_vz, lam=M.lam, mu_txx=M.mu_txx,
                    mu_txz=M.mu_txz, w_vx=H.w_vx, w_vz=H.w_vz, w_tau=H.w_tau, src_i=src.i,
                    src_j=src.j, src_type=SRC_CODES[src.type], wavelet=src.wavelet, rec_i=rec.i,
                    rec_j=rec.j, rec_type=rec.type_codes(), nt=NT, dtype=F32, flags=1,
                    storage=storage, fscale=[float(x) for x in scales])
    rng = np.random.default_rng(2026)
    for f in range(5):
        v = rng.standard_normal((M.nx, M.nz), dtype=F32)
        if storage == "f16":
            # about 100 in stored units: the 20 steps' stencil growth (the
            # largest modulus times dt/dx over the smallest impedance, about
            # 70x) stays inside FP16's range
            v = (v * F32(100.0 / float(scales[f]))).astype(F32)
        W.upload(f, v)
        if storage == "bf16":
            r.field(f)[:] = _bf16_rne(v)
        elif storage == "f24":
            r.field(f)[:] = _f24_rne(v)
        elif storage == "f16":
            r.field(f)[:] = _f16_rne(v, float(scales[f]))
        else:
            r.field(f)[:] = v
        del v
    try:
        E.run_time_loop(b, W, M, H, a, src, rec, p)
        b.synchronize()
    finally:
        b.set_tiling(1, 0, 16)
        b.set_step_mode(1)
        b.set_split(16, 0)
    if mode == 3:  # the double step ran: its third state buffer exists
        fresh = E.Wavefield(M.nx, M.nz, b)
        fresh._bind(dm)
        assert b.lib.fomo_wavefield_device_bytes(W._h) > b.lib.fomo_wavefield_device_bytes(fresh._h)
        fresh.close()
    r.run(1, NT)
    assert np.array_equal(rec.data, r.gather_nt_by_nrec())
    for f in range(5):
        got, want = W.download(f), r.field(f)
        if not np.array_equal(got, want):
            bad = np.argwhere(got != want)
            i, j = bad[0]
            raise AssertionError(f"field {f}: {len(bad)} cells differ, first (i={i}, j={j}) "
                                 f"got {got[i, j]!r} want {want[i, j]!r}")


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_marmousi_batch_of_64_random_states_bit_exact(mode):
    """The batched workload at its full size: 64 Marmousi-shaped shots (1700 x
    350, nbc 40, order 12, padded 1792 x 442) in one batch, two columns per
    lane (the batch's autotuned shape), each from its own random state and
    source position; after 20 steps every shot's fields and gather equal its
    own oracle run.  Mode 1 is the whole-grid fused step, mode 2 the split
    step (interior tiles of all 64 shots beside their edge tiles and strip
    kernels, 16 edge rows, one interior block per CU); mode 3 asks for the
    double step, which order 12 does not take (the split step runs)."""
    import ctypes
    import oracle as O
    dx = F32(10.0)
    vp, vs, rho = media.marmousi_like()
    dt = media.cfl_dt(vp, dx, 0.4)
    nbc, order, n_shots = 40, 12, 64
    M = E.init_medium(vp, vs, rho, dx, dx, nbc, order, None, free_surface=True)
    H = E.init_habc(M.nx, M.nz, nbc, dt, dx, dx, I.julia_mean_f32(vp), None)
    p = E.SimParams(dt, NT, dx, dx, order)
    a = I.get_fd_coefficients(order)
    wav = I.ricker_wavelet(F32(25.0), dt, NT)
    nx_in = vp.shape[0]
    rec = E.setup_receivers(np.arange(nx_in, dtype=F32) * dx, np.zeros(nx_in, F32), M, type="vz")
    print(f"case Marmousi batch {n_shots} x {M.nx}x{M.nz}, order {order}, {NT} steps, "
          f"step mode {mode}", flush=True)

    b = E.backend("mi355x")
    b.set_tiling(2, 0, 16)
    b.set_step_mode(mode)
    b.set_split(16, 1)
    W0 = E.Wavefield(M.nx, M.nz, b)
    dm = E.bind(b, W0, M, H, a, p)
    xs = ((np.arange(n_shots) + 0.5) * (float(M.x_max) / n_shots)).astype(F32)
    srcs, wfs, dss, states = [], [], [], []
    rng = np.random.default_rng(7)
    for x in xs:
        src = E.Source(I.coord_to_index(F32(x), M.dx, M.pad), I.coord_to_index(F32(0.0), M.dz, M.pad),
                       wav, "pressure")
        w = E.Wavefield(M.nx, M.nz, b)
        w._bind(dm)
        st = [rng.standard_normal((M.nx, M.nz), dtype=F32) for _ in range(5)]
        for f in range(5):
            w.upload(f, st[f])
        srcs.append(src)
        wfs.append(w)
        states.append(st)
        dss.append(E._DeviceShot(dm, src, rec, NT))
    arr = lambda hs: (ctypes.c_void_p * n_shots)(*[h.value for h in hs])
    bh = ctypes.c_void_p()
    b.check(b.lib.fomo_batch_create(b.ctx, n_shots, arr([dm.h] * n_shots), arr([w._h for w in wfs]),
                                    arr([d.h for d in dss]), ctypes.byref(bh)), "batch_create")
    try:
        b.check(b.lib.fomo_batch_run(bh, 1, NT), "batch_run")
        b.synchronize()
    finally:
        b.lib.fomo_batch_destroy(bh)
        b.set_tiling(1, 0, 16)
        b.set_step_mode(1)
        b.set_split(16, 0)
    for s in range(n_shots):
        r = O.OracleRun(nx=M.nx, nz=M.nz, pad=M.pad, nbc=nbc, fd_order=order, free_surface=True,
                        dt=p.dt, dtx=p.dtx, dtz=p.dtz, q=(H.qx, H.qz, H.qt_x, H.qt_z, H.qxt), a=a,
                        rho_vx=M.rho_vx, rho_vz=M.rho_vz, lam=M.lam, mu_txx=M.mu_txx,
                        mu_txz=M.mu_txz, w_vx=H.w_vx, w_vz=H.w_vz, w_tau=H.w_tau, src_i=srcs[s].i,
                        src_j=srcs[s].j, src_type=SRC_CODES["pressure"], wavelet=wav,
                        rec_i=rec.i, rec_j=rec.j, rec_type=rec.type_codes(), nt=NT, dtype=F32,
                        flags=1)
        for f in range(5):
            r.field(f)[:] = states[s][f]
        r.run(1, NT)
        assert np.array_equal(dss[s].gather(), r.gather_nt_by_nrec()), f"shot {s}: gather"
        for f in range(5):
            got, want = wfs[s].download(f), r.field(f)
            if not np.array_equal(got, want):
                bad = np.argwhere(got != want)
                i, j = bad[0]
                raise AssertionError(f"shot {s} field {f}: {len(bad)} cells differ, first (i={i}, "
                                     f"j={j}) got {got[i, j]!r} want {want[i, j]!r}")
        states[s] = None


def test_grid_beyond_32_bit_buffer_offsets_bit_exact():
    """A grid whose coefficient blocks exceed the fused kernel's 2 GiB buffer
    range (8600 x 8600 padded, 74 M cells; 7 blocks of 310 MB): the fused step
    is requested, the library takes the two-pass step with 64-bit addressing,
    and 4 steps from a random state equal the oracle's bit for bit."""
    import oracle as O
    nbc, order = 20, 8
    pad = nbc + order // 2
    nx_in = nz_in = 8600 - 2 * pad
    rng = np.random.default_rng(5)
    vp = np.full((nx_in, nz_in), 3000.0, F32)
    vp[:, nz_in // 2:] = 3600.0
    vp += rng.uniform(-50, 50, vp.shape).astype(F32)
    vs = (vp / F32(1.73)).astype(F32)
    rho = np.full((nx_in, nz_in), 2200.0, F32)
    dx = F32(10.0)
    dt = media.cfl_dt(vp, dx, 0.4)
    M = E.init_medium(vp, vs, rho, dx, dx, nbc, order, None, free_surface=True)
    del vs, rho
    H = E.init_habc(M.nx, M.nz, nbc, dt, dx, dx, I.julia_mean_f32(vp), None)
    del vp
    p = E.SimParams(dt, 4, dx, dx, order)
    a = I.get_fd_coefficients(order)
    src = E.Source(I.coord_to_index(F32(M.x_max / F32(2)), M.dx, M.pad),
                   I.coord_to_index(F32(0.0), M.dz, M.pad), I.ricker_wavelet(F32(15.0), dt, 4),
                   "pressure")
    xr = np.arange(0, nx_in, 13, dtype=F32) * dx
    rec = E.setup_receivers(xr, np.full(len(xr), 2 * dx, F32), M, type="p")
    print(f"case large grid {M.nx}x{M.nz} ({M.nx * M.nz / 1e6:.0f} M cells), 4 steps", flush=True)

    b = E.backend("mi355x")
    b.set_step_mode(1)
    b.set_tiling(2, 0, 16)
    W = E.Wavefield(M.nx, M.nz, b)
    E.bind(b, W, M, H, a, p)
    r = O.OracleRun(nx=M.nx, nz=M.nz, pad=M.pad, nbc=nbc, fd_order=order, free_surface=True,
                    dt=p.dt, dtx=p.dtx, dtz=p.dtz, q=(H.qx, H.qz, H.qt_x, H.qt_z, H.qxt), a=a,
                    rho_vx=M.rho_vx, rho_vz=M.rho_vz, lam=M.lam, mu_txx=M.mu_txx,
                    mu_txz=M.mu_txz, w_vx=H.w_vx, w_vz=H.w_vz, w_tau=H.w_tau, src_i=src.i,
                    src_j=src.j, src_type=SRC_CODES[src.type], wavelet=src.wavelet, rec_i=rec.i,
                    rec_j=rec.j, rec_type=rec.type_codes(), nt=4, dtype=F32, flags=1)
    for f in range(5):
        v = rng.standard_normal((M.nx, M.nz), dtype=F32)
        W.upload(f, v)
        r.field(f)[:] = v
        del v
    try:
        E.run_time_loop(b, W, M, H, a, src, rec, p)
    finally:
        b.set_tiling(1, 0, 16)
    r.run(1, 4)
    assert np.array_equal(rec.data, r.gather_nt_by_nrec())
    for f in range(5):
        got, want = W.download(f), r.field(f)
        if not np.array_equal(got, want):
            bad = np.argwhere(got != want)
            i, j = bad[0]
            raise AssertionError(f"field {f}: {len(bad)} cells differ, first (i={i}, j={j}) "
                                 f"got {got[i, j]!r} want {want[i, j]!r}")
        del got
