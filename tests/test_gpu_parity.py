"""MI355X kernels against the CPU oracle, through the C ABI.

FP32 results are compared bit for bit with the oracle's FMA mode (the kernels
and the oracle share one arithmetic contract); the oracle's unfused mode and
its Float64 run are compared by relative L2 (tolerance 1e-5 for
FP32 gathers)."""
import os

import numpy as np
import pytest

import oracle as O
from fomo_julia_amd import elastic2d as E
from tests.helpers import Case, rel_l2, run_jl_case

pytestmark = pytest.mark.gpu
F32 = np.float32
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "shot_1.bin")


@pytest.fixture(scope="module")
def b():
    bk = E.backend("mi355x")
    bk.set_tiling(1, 0, 16)
    return bk


def gpu_shot(b, c, cpt=1, tz=0, gsteps=16, mode=1):
    b.set_tiling(cpt, tz, gsteps)
    b.set_step_mode(mode)
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    res = E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p)
    return res, W


def assert_same_state(W, r, what=""):
    for f in range(10):
        got = W.download(f)
        want = r.field(f)
        if not np.array_equal(got, want):
            bad = np.argwhere(got != want)
            i, j = bad[0]
            raise AssertionError(f"{what} field {f}: {len(bad)} cells differ, first (i={i}, j={j}) "
                                 f"got {got[i, j]!r} want {want[i, j]!r}")


@pytest.mark.parametrize("mode", [1, 0, 2])
@pytest.mark.parametrize("fs", [True, False])
@pytest.mark.parametrize("src_type", ["pressure", "force_z", "force_x"])
def test_full_shot_bit_exact(b, fs, src_type, mode):
    c = Case(fs=fs, src_type=src_type, nt=150, water=fs)
    print(f"case fs={fs} src={src_type} mode={mode} grid={c.M.nx}x{c.M.nz}", flush=True)
    res, W = gpu_shot(b, c, mode=mode)
    b.set_step_mode(1)
    r = c.oracle()
    r.run(1, c.nt)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    assert_same_state(W, r)


@pytest.mark.parametrize("mode", [1, 0, 2])
@pytest.mark.parametrize("fd_order", [2, 4, 6, 10, 12, 16])
def test_orders_bit_exact(b, fd_order, mode):
    c = Case(fd_order=fd_order, nt=80, nbc=10)
    print(f"case order={fd_order} mode={mode}", flush=True)
    res, W = gpu_shot(b, c, mode=mode)
    b.set_step_mode(1)
    r = c.oracle()
    r.run(1, c.nt)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    assert_same_state(W, r)


@pytest.mark.parametrize("mode", [1, 0, 2])
@pytest.mark.parametrize("cpt,tz,gsteps", [(1, 8, 2), (2, 17, 6), (4, 64, 32), (4, 5, 16), (1, 0, 4)])
def test_tilings_bit_exact(b, cpt, tz, gsteps, mode):
    c = Case(nx_in=150, nz_in=90, nbc=20, nt=71)
    print(f"case cpt={cpt} tz={tz} gsteps={gsteps} mode={mode}", flush=True)
    res, W = gpu_shot(b, c, cpt, tz, gsteps, mode)
    b.set_step_mode(1)
    r = c.oracle()
    r.run(1, c.nt)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    assert_same_state(W, r)
    b.set_tiling(1, 0, 16)


@pytest.mark.parametrize("edge_rows,bpc", [(1, 0), (7, 1), (16, 2), (64, 0)])
@pytest.mark.parametrize("rec_where", ["frame", "interior"])
def test_split_step_shapes_bit_exact(b, edge_rows, bpc, rec_where):
    """Step mode 2 (edge tiles, then interior tiles beside the strip kernels):
    edge tile heights, caps on the interior's blocks per CU, receivers only
    in the frame (recorded beside the strip kernels) or also inside the
    interior (recorded after the join), two columns per lane, odd nx."""
    c = Case(nx_in=151, nz_in=97, nbc=14, nt=75, rec_types=("vz", "vx", "p"))
    if rec_where == "interior":
        xr = np.linspace(0.0, float(c.M.x_max), 41).astype(F32)
        zr = np.linspace(0.0, 90.0 * float(c.M.dz), 41).astype(F32)
        c.rec = E.setup_receivers(xr, zr, c.M, type=[("vz", "vx", "p")[k % 3] for k in range(41)])
    print(f"case split edge_rows={edge_rows} bpc={bpc} receivers={rec_where}", flush=True)
    b.set_split(edge_rows, bpc)
    try:
        res, W = gpu_shot(b, c, 2, 0, 8, 2)
    finally:
        b.set_split(16, 0)
        b.set_step_mode(1)
        b.set_tiling(1, 0, 16)
    r = c.oracle()
    r.run(1, c.nt)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    assert_same_state(W, r)


STAGES = [("backup", E.backup_boundary), ("velocity", E.update_velocity),
          ("habc_v", E.apply_habc_velocity), ("stress", E.update_stress),
          ("habc_s", E.apply_habc_stress), ("free_surface", E.apply_free_surface)]


@pytest.mark.parametrize("fs", [True, False])
@pytest.mark.parametrize("name,fn", STAGES)
def test_each_stage_bit_exact_on_random_state(b, fs, name, fn):
    c = Case(fs=fs, nt=10)
    print(f"case stage={name} fs={fs}", flush=True)
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    E.bind(b, W, c.M, c.H, c.a, c.p)
    r = c.oracle()
    rng = np.random.default_rng(11)
    for f in range(10):
        v = rng.standard_normal((c.M.nx, c.M.nz)).astype(F32)
        W.upload(f, v)
        r.field(f)[:] = v
    fn(b, W)
    b.synchronize()
    r.run(1, 1, O.STAGES[name])
    assert_same_state(W, r, name)


@pytest.mark.parametrize("nx_in", [150, 151])
@pytest.mark.parametrize("fd_order", [4, 8, 12])
@pytest.mark.parametrize("nbc", [20, 21])
def test_column_pair_stores_bit_exact(b, nx_in, fd_order, nbc):
    """Two columns per lane: even grids store each lane's column pair with one
    instruction and widen the fused kernel's stress mask to even bounds (the
    band kernel rewrites the extra column); odd grids keep per-column stores.
    Both grid parities, both band-width parities (nbc + M + 1), three orders."""
    c = Case(nx_in=nx_in, nz_in=60, nbc=nbc, fd_order=fd_order, nt=90)
    print(f"case pairs nx={c.M.nx} band={nbc + fd_order // 2 + 1} order={fd_order}", flush=True)
    res, W = gpu_shot(b, c, cpt=2, tz=0, gsteps=8)
    b.set_tiling(1, 0, 16)
    r = c.oracle()
    r.run(1, c.nt)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    assert_same_state(W, r)


@pytest.mark.parametrize("cpt", [2, 4])
@pytest.mark.parametrize("fd_order", [10, 12, 16])
def test_high_order_wide_tiles_bit_exact(b, cpt, fd_order):
    """Orders 10-16 with 2 and 4 columns per lane: the fused kernel's z-march
    is fully unrolled there too (its ring windows stay in registers)."""
    c = Case(nx_in=150, nz_in=70, nbc=20, fd_order=fd_order, nt=70)
    print(f"case wide tiles order={fd_order} cpt={cpt} nx={c.M.nx}", flush=True)
    res, W = gpu_shot(b, c, cpt=cpt, tz=0, gsteps=8)
    b.set_tiling(1, 0, 16)
    r = c.oracle()
    r.run(1, c.nt)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    assert_same_state(W, r)


@pytest.mark.parametrize("nbc", [1, 2, 33, 53, 57, 63])
@pytest.mark.parametrize("fs", [True, False])
@pytest.mark.parametrize("name", ["habc_v", "habc_s"])
def test_habc_strip_widths_bit_exact_on_random_state(b, nbc, fs, name):
    """Every HABC piece at strip widths from 1 to the widest one wavefront
    spans (63), on a random state.  The bottom-right corner's sweep reads one
    group of steps past its block, which for nbc > 52 runs into the LDS the
    launch adds for it; 33 and 57 are odd widths (a partial last group)."""
    c = Case(nx_in=30, nz_in=26, nbc=nbc, fs=fs, nt=4)
    print(f"case stage={name} nbc={nbc} fs={fs} grid={c.M.nx}x{c.M.nz}", flush=True)
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    E.bind(b, W, c.M, c.H, c.a, c.p)
    r = c.oracle()
    rng = np.random.default_rng(nbc)
    for f in range(10):
        v = rng.standard_normal((c.M.nx, c.M.nz)).astype(F32)
        W.upload(f, v)
        r.field(f)[:] = v
    dict(STAGES)[name](b, W)
    b.synchronize()
    r.run(1, 1, O.STAGES[name])
    assert_same_state(W, r, name)


@pytest.mark.parametrize("nbc", [53, 63])
def test_wide_strip_full_shot_bit_exact(b, nbc):
    c = Case(nx_in=40, nz_in=36, nbc=nbc, nt=140)
    print(f"case full shot nbc={nbc} grid={c.M.nx}x{c.M.nz}", flush=True)
    res, W = gpu_shot(b, c)
    r = c.oracle()
    r.run(1, c.nt)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    assert_same_state(W, r)


def test_inject_and_record_stages(b):
    c = Case(nt=12, src_type="pressure")
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    E.bind(b, W, c.M, c.H, c.a, c.p)
    r = c.oracle()
    rng = np.random.default_rng(5)
    for f in range(5):
        v = rng.standard_normal((c.M.nx, c.M.nz)).astype(F32)
        W.upload(f, v)
        r.field(f)[:] = v
    E.inject_source(b, W, c.src, 7, rec=c.rec)
    b.synchronize()
    r.run(7, 1, O.STAGES["inject"])
    assert_same_state(W, r, "inject")
    E.record_receivers(b, W, c.rec, 7, src=c.src, nt=c.nt)
    r.run(7, 1, O.STAGES["record"])
    assert np.array_equal(c.rec.data, r.gather_nt_by_nrec())


@pytest.mark.parametrize("mode", [1, 0])
def test_time_step_api_matches_run_loop(b, mode):
    b.set_step_mode(mode)
    c = Case(nt=40)
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    E.bind(b, W, c.M, c.H, c.a, c.p)
    for k in range(1, 41):
        E.time_step(b, W, c.M, c.H, c.a, c.src, c.rec, k, c.p)
    b.synchronize()
    r = c.oracle()
    r.run(1, 40)
    assert_same_state(W, r, "time_step")
    # per-stage calls continue from the fused state (strip backups included)
    E.backup_boundary(b, W)
    E.update_velocity(b, W)
    E.apply_habc_velocity(b, W)
    b.synchronize()
    r.run(41, 1, O.STAGES["backup"] | O.STAGES["velocity"] | O.STAGES["habc_v"])
    assert_same_state(W, r, "stages after time_step")
    b.set_step_mode(1)


@pytest.mark.parametrize("mode", [1, 2])
def test_batched_uneven_grids_bit_exact(b, mode):
    cases = [Case(nx_in=72, nz_in=56, nt=90, seed=1),
             Case(nx_in=130, nz_in=64, nt=90, seed=2, fs=False),
             Case(nx_in=60, nz_in=90, nt=90, seed=3, src_type="force_z"),
             Case(nx_in=180, nz_in=120, nt=90, seed=4, src_type="force_x")]
    import ctypes
    from fomo_julia_amd import _lib
    lib = _lib.load()
    print(f"case uneven batch of {len(cases)}, step mode {mode}", flush=True)
    b.set_tiling(2, 16, 8)
    b.set_step_mode(mode)
    b.set_split(7, 1)
    hs = []
    for c in cases:
        dm = E._model(b, c.M, c.H, c.a, c.p)
        W = E.Wavefield(c.M.nx, c.M.nz, b)
        W._bind(dm)
        ds = E._DeviceShot(dm, c.src, c.rec, c.nt)
        hs.append((dm, W, ds))
    n = len(hs)
    arr = lambda xs: (ctypes.c_void_p * n)(*[x.value for x in xs])
    bh = ctypes.c_void_p()
    b.check(lib.fomo_batch_create(b.ctx, n, arr([h[0].h for h in hs]), arr([h[1]._h for h in hs]),
                                  arr([h[2].h for h in hs]), ctypes.byref(bh)), "batch_create")
    b.check(lib.fomo_batch_run(bh, 1, 50), "batch_run")
    b.check(lib.fomo_batch_run(bh, 51, 40), "batch_run")
    b.synchronize()
    lib.fomo_batch_destroy(bh)
    for c, (dm, W, ds) in zip(cases, hs):
        r = c.oracle()
        r.run(1, c.nt)
        assert np.array_equal(ds.gather(), r.gather_nt_by_nrec())
        assert_same_state(W, r, "batch")
    b.set_tiling(1, 0, 16)
    b.set_step_mode(1)
    b.set_split(16, 0)


@pytest.mark.parametrize("nts", [(60, 35), (40, 40, 70)])
def test_two_stream_batch_members_that_end_apart(b, nts):
    """A whole-grid batch of two or more members steps as two halves on two
    streams; a member whose shot ends first idles in its half while the
    other half runs on, and every member still equals its own oracle run,
    its state read from the buffer of its own last step (runs that go on
    past a member's end leave the members at different buffer parities)."""
    import ctypes
    from fomo_julia_amd import _lib
    lib = _lib.load()
    cases = [Case(nx_in=70 + 20 * q, nz_in=48 + 6 * q, nt=nt, seed=q + 5) for q, nt in enumerate(nts)]
    print(f"case two-stream batch, nt {nts}", flush=True)
    b.set_tiling(2, 0, 8)
    b.set_step_mode(1)
    hs = []
    for c in cases:
        dm = E._model(b, c.M, c.H, c.a, c.p)
        W = E.Wavefield(c.M.nx, c.M.nz, b)
        W._bind(dm)
        hs.append((dm, W, E._DeviceShot(dm, c.src, c.rec, c.nt)))
    n = len(hs)
    arr = lambda xs: (ctypes.c_void_p * n)(*[x.value for x in xs])
    bh = ctypes.c_void_p()
    b.check(lib.fomo_batch_create(b.ctx, n, arr([h[0].h for h in hs]), arr([h[1]._h for h in hs]),
                                  arr([h[2].h for h in hs]), ctypes.byref(bh)), "batch_create")
    try:
        # two chunks of 8 and 5 steps; then runs in which members end at
        # different buffer parities, and one in which only some still step
        b.check(lib.fomo_batch_run(bh, 1, 21), "batch_run")
        b.check(lib.fomo_batch_run(bh, 22, 20), "batch_run")
        b.check(lib.fomo_batch_run(bh, 42, max(nts) - 41), "batch_run")
        b.synchronize()
    finally:
        lib.fomo_batch_destroy(bh)
        b.set_tiling(1, 0, 16)
    for c, (dm, W, ds) in zip(cases, hs):
        r = c.oracle()
        r.run(1, c.nt)
        assert np.array_equal(ds.gather(), r.gather_nt_by_nrec())
        assert_same_state(W, r, f"two-stream batch member with nt {c.nt}")


def test_run_shots_batched_equals_sequential(b):
    c = Case(nx_in=90, nz_in=50, nt=60)
    xs = [F32(100.0), F32(300.0), F32(550.0)]
    cfg = E.MultiShotConfig(xs, [F32(30.0)] * 3, c.src.wavelet)
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    seq = E.run_shots(b, W, c.M, c.H, c.a, c.rec, cfg, c.p, batch=1)
    seen = []
    bat = E.run_shots(b, W, c.M, c.H, c.a, c.rec, cfg, c.p, batch=3,
                      on_shot_complete=lambda r: seen.append(r.shot_id))
    assert seen == [1, 2, 3]
    for s, t in zip(seq, bat):
        assert (s.src_i, s.src_j, s.shot_id) == (t.src_i, t.src_j, t.shot_id)
        assert np.array_equal(s.gather, t.gather)


def test_reference_golden_gather_and_fp64_error(b):
    d = run_jl_case()
    W = E.Wavefield(d["M"].nx, d["M"].nz, b)
    res = E.run_shot(b, W, d["M"], d["H"], d["a"], d["src"], d["rec"], d["p"])
    gold = np.fromfile(GOLDEN, dtype="<f4").reshape(100, 800).T
    err = rel_l2(res.gather, gold)
    print(f"run.jl gather vs reference shot_1.bin: rel L2 {err:.3e}")
    assert err < 1e-5


def test_fp32_gather_within_tolerance_of_unfused_and_fp64_oracle(b):
    c = Case(nx_in=120, nz_in=80, nbc=20, nt=300)
    res, W = gpu_shot(b, c)
    sep = c.oracle(flags=0)
    sep.run(1, c.nt)
    r64 = c.oracle(dtype=np.float64)
    r64.run(1, c.nt)
    e_sep = rel_l2(res.gather, sep.gather_nt_by_nrec())
    e_64 = rel_l2(res.gather, r64.gather_nt_by_nrec())
    print(f"gather rel L2: vs unfused FP32 oracle {e_sep:.2e}, vs FP64 oracle {e_64:.2e}")
    assert e_sep <= 1e-5 and e_64 <= 1e-5


def test_closed_form_ramps_equal_weight_matrices(b):
    c = Case(nt=60)
    res_w, _ = gpu_shot(b, c)
    c.H = E.HABCConfig(c.H.nbc, c.H.qx, c.H.qz, c.H.qt_x, c.H.qt_z, c.H.qxt)  # no matrices
    c.rec = E.Receivers(c.rec.i, c.rec.j, np.zeros((1, len(c.rec.i)), F32), c.rec.type)
    res_r, _ = gpu_shot(b, c)
    assert np.array_equal(res_w.gather, res_r.gather)


def test_non_ramp_weights_rejected(b):
    c = Case(nt=10)
    c.H.w_tau = c.H.w_tau.copy()
    c.H.w_tau[5, 5] = 0.123
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    with pytest.raises(Exception, match="edge distance"):
        E.bind(b, W, c.M, c.H, c.a, c.p)


def _perturbed_case(**kw):
    """A case whose mu_txz is not init_medium's harmonic average in one
    interior cell: the model keeps the five stored coefficient arrays."""
    c = Case(**kw)
    i, j = c.M.nx // 2, c.M.nz // 2
    c.M.mu_txz = np.array(c.M.mu_txz, copy=True)
    c.M.mu_txz[i, j] = np.nextafter(c.M.mu_txz[i, j], np.float32(np.inf))
    return c


@pytest.mark.parametrize("fs,src_type", [(True, "pressure"), (False, "force_z")])
def test_stored_and_derived_coefficients_bit_exact(b, fs, src_type):
    """Media with init_medium's averages run the fused kernel that forms the
    coefficients from rho_vx, lam and mu_txx; any other medium keeps the five
    stored arrays.  Both are bit-exact against the oracle on their own arrays."""
    for perturbed in (False, True):
        c = (_perturbed_case if perturbed else Case)(fs=fs, src_type=src_type, nt=120, water=fs)
        print(f"case coefficients perturbed={perturbed} fs={fs} src={src_type}", flush=True)
        res, W = gpu_shot(b, c, cpt=2)
        b.set_tiling(1, 0, 16)
        dm = E._model(b, c.M, c.H, c.a, c.p)
        pitch = (c.M.nx + 272 + 255) // 256 * 256
        block = 4 * ((c.M.nz + 36) * pitch + 512)  # one pitched FP32 array with guards
        ramps = 4 * 3 * (c.H.nbc + 2)
        blocks = (int(b.lib.fomo_model_device_bytes(dm.h)) - ramps) / block
        # 5 + rho_vx, mu_txx for the derived form, 5 otherwise
        assert blocks == (5 if perturbed else 7), blocks
        r = c.oracle()
        r.run(1, c.nt)
        assert np.array_equal(res.gather, r.gather_nt_by_nrec())
        assert_same_state(W, r, f"perturbed={perturbed}")


def test_pool_mixing_stored_and_derived_coefficients(b):
    """A pool that starts with derived-coefficient members and then admits a
    member with stored arrays and a wider HABC strip (more LDS rows)
    recaptures its step and stays bit-exact."""
    import ctypes
    cases = [Case(nx_in=70, nz_in=50, nt=48, seed=1),
             _perturbed_case(nx_in=64, nz_in=58, nt=48, seed=2, nbc=20)]
    print("case pool mixed coefficient forms", flush=True)
    b.set_tiling(1, 0, 16)
    b.set_step_mode(1)
    solo = []
    for c in cases:
        W = E.Wavefield(c.M.nx, c.M.nz, b)
        solo.append(E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p).gather)
    lib = b.lib
    hp = ctypes.c_void_p()
    b.check(lib.fomo_pool_create(b.ctx, 8, 0, 4, ctypes.byref(hp)), "pool_create")
    keep = []
    try:
        for q, c in enumerate(cases):
            dm = E._model(b, c.M, c.H, c.a, c.p)
            W = E.Wavefield(c.M.nx, c.M.nz, b)
            W._bind(dm)
            E.reset(b, W)
            ds = E._DeviceShot(dm, c.src, c.rec, c.nt)
            keep.append((W, ds))
            b.check(lib.fomo_pool_add(hp, dm.h, W._h, ds.h), "pool_add")
            if q == 0:
                b.check(lib.fomo_pool_run(hp, 16), "pool_run")  # captured with the derived form
        # member 0 is 16 steps ahead: run until both are done
        b.check(lib.fomo_pool_run(hp, 48), "pool_run")
        b.check(lib.fomo_pool_synchronize(hp), "pool_sync")
        assert np.array_equal(keep[0][1].gather(), solo[0])
        assert np.array_equal(keep[1][1].gather(), solo[1])
    finally:
        lib.fomo_pool_destroy(hp)
