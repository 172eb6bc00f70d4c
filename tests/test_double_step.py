"""Step mode 3 (double step) against the CPU oracle, bit for bit, through the
C ABI: two time steps per pass over the interior (k_fused2), the split step's
edge tiles and strip kernels for step k (outside the interior shrunk by 2M,
into the third buffer) and step k + 1 (outside the interior) beside it.

Every case also checks that the double step really ran (the third buffer was
allocated) or, where it cannot run (a receiver inside the shrunk interior),
that the split step took over."""
import ctypes

import numpy as np
import pytest

from fomo_julia_amd import _lib
from fomo_julia_amd import elastic2d as E
from tests.helpers import Case

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def b():
    bk = E.backend("mi355x")
    yield bk
    bk.set_step_mode(1)
    bk.set_split(16, 0)
    bk.set_tiling(1, 0, 16)


def _bytes(W):
    return int(_lib.load().fomo_wavefield_device_bytes(W._h))


def _state_equal(W, r, what):
    for f in range(10):
        got, want = W.download(f), r.field(f)
        if not np.array_equal(got, want):
            bad = np.argwhere(got != want)
            i, j = bad[0]
            raise AssertionError(f"{what} field {f}: {len(bad)} cells differ, first (i={i}, j={j}) "
                                 f"got {got[i, j]!r} want {want[i, j]!r}")


def _run(b, c, cpt=2, gsteps=16, edge_rows=16, bpc=1):
    b.set_tiling(cpt, 0, gsteps)
    b.set_step_mode(3)
    b.set_split(edge_rows, bpc)
    try:
        W = E.Wavefield(c.M.nx, c.M.nz, b)
        res = E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p)
        b.synchronize()
        fresh = E.Wavefield(c.M.nx, c.M.nz, b)  # the same model's wavefield, never run
        fresh._bind(W._dm)
        grew = _bytes(W) > _bytes(fresh)
        fresh.close()
    finally:
        b.set_step_mode(1)
        b.set_split(16, 0)
        b.set_tiling(1, 0, 16)
    return res, W, grew


def _xz(c, i, j):
    """Physical source coordinates of padded cell (i, j) (0-based)."""
    return (F32((i - c.M.pad) * float(c.M.dx)), F32((j - c.M.pad) * float(c.M.dz)))


# Case(nx_in=151, nz_in=97, nbc=14, fd_order=8): padded 187 x 133, band 19,
# interior [30, 158) x [29, 104), shrunk interior [38, 150) x [37, 96).
SRC_CELLS = {"frame": None, "deep": (94, 66), "interior_edge": (32, 60), "ring": (27, 60),
             "shrunk_edge": (38, 37)}


@pytest.mark.parametrize("where", list(SRC_CELLS))
@pytest.mark.parametrize("src_type", ["pressure", "force_z", "force_x"])
def test_double_step_sources_bit_exact(b, src_type, where):
    base = Case(nx_in=151, nz_in=97, nbc=14, nt=75)
    cell = SRC_CELLS[where]
    c = Case(nx_in=151, nz_in=97, nbc=14, nt=75, src_type=src_type,
             src_xz=None if cell is None else _xz(base, *cell))
    if cell is not None:  # Source indices are 1-based, as in the reference
        assert (int(c.src.i) - 1, int(c.src.j) - 1) == cell
    print(f"case double src={src_type} at {where} ({int(c.src.i)},{int(c.src.j)}) nt={c.nt}", flush=True)
    res, W, grew = _run(b, c)
    assert grew, "the double step did not run"
    r = c.oracle()
    r.run(1, c.nt)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    _state_equal(W, r, "double")


@pytest.mark.parametrize("fd_order", [4, 8, 12])
@pytest.mark.parametrize("cpt,gsteps,edge_rows,bpc", [(1, 6, 7, 1), (2, 16, 16, 1), (2, 4, 16, 0),
                                                      (4, 0, 3, 2)])
def test_double_step_orders_and_shapes_bit_exact(b, fd_order, cpt, gsteps, edge_rows, bpc):
    """Orders with k_fused2 instances (4 and 8; order 12 takes the split
    step), edge-tile shapes, graph chunks of an odd number of double steps
    (6 steps), no graphs, even and odd nt."""
    nt = 64 if gsteps == 16 else 71
    c = Case(nx_in=160, nz_in=110, nbc=12, fd_order=fd_order, nt=nt, fs=fd_order != 12,
             src_xz=None if fd_order == 4 else (F32(700.0), F32(500.0)))
    print(f"case double order={fd_order} cpt={cpt} gsteps={gsteps} edge_rows={edge_rows} "
          f"bpc={bpc} nt={nt}", flush=True)
    res, W, grew = _run(b, c, cpt, gsteps, edge_rows, bpc)
    assert grew == (fd_order != 12), "the double step ran at the wrong orders"
    r = c.oracle()
    r.run(1, c.nt)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    _state_equal(W, r, "double")


@pytest.mark.parametrize("rec_where", ["interior_ring", "shrunk_interior"])
def test_double_step_receivers(b, rec_where):
    """Receivers inside the interior but outside the shrunk interior are
    recorded after the join (step k + 1) and from the third buffer (step k);
    one inside the shrunk interior makes the run take split steps."""
    c = Case(nx_in=151, nz_in=97, nbc=14, nt=70, rec_types=("vz", "vx", "p"))
    j = 33 if rec_where == "interior_ring" else 60  # padded rows; shrunk interior rows [37, 96)
    xr = np.linspace(0.0, float(c.M.x_max), 37).astype(F32)
    zr = np.full(37, (j - c.M.pad) * float(c.M.dz), F32)
    c.rec = E.setup_receivers(xr, zr, c.M, type=[("vz", "vx", "p")[k % 3] for k in range(37)])
    assert int(c.rec.j[0]) - 1 == j  # 1-based, as in the reference
    print(f"case double receivers {rec_where} (row {j})", flush=True)
    res, W, grew = _run(b, c)
    assert grew == (rec_where == "interior_ring")
    r = c.oracle()
    r.run(1, c.nt)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    _state_equal(W, r, "double receivers")


def test_double_step_batch_in_pieces_bit_exact(b):
    """A batch of uneven members with interiors, and one with a member too
    small for a shrunk interior (the whole batch then takes split steps),
    each run in pieces that end on a single and on a double step."""
    lib = _lib.load()
    groups = {
        "interiors": [Case(nx_in=150, nz_in=100, nt=90, seed=1, nbc=10),
                      Case(nx_in=200, nz_in=120, nt=90, seed=2, fs=False, nbc=10,
                           src_type="force_z", src_xz=(F32(900.0), F32(600.0))),
                      Case(nx_in=170, nz_in=140, nt=90, seed=4, nbc=10, src_type="force_x")],
        "one_small": [Case(nx_in=150, nz_in=100, nt=90, seed=1, nbc=10),
                      Case(nx_in=40, nz_in=40, nt=90, seed=3, nbc=10)],
    }
    for name, cases in groups.items():
        print(f"case double batch {name}: {len(cases)} members, runs 51 + 39", flush=True)
        b.set_tiling(2, 0, 8)
        b.set_step_mode(3)
        b.set_split(16, 1)
        hs = []
        try:
            for c in cases:
                dm = E._model(b, c.M, c.H, c.a, c.p)
                W = E.Wavefield(c.M.nx, c.M.nz, b)
                W._bind(dm)
                ds = E._DeviceShot(dm, c.src, c.rec, c.nt)
                hs.append((dm, W, ds, _bytes(W)))
            n = len(hs)
            arr = lambda xs: (ctypes.c_void_p * n)(*[x.value for x in xs])
            bh = ctypes.c_void_p()
            b.check(lib.fomo_batch_create(b.ctx, n, arr([h[0].h for h in hs]), arr([h[1]._h for h in hs]),
                                          arr([h[2].h for h in hs]), ctypes.byref(bh)), "batch_create")
            b.check(lib.fomo_batch_run(bh, 1, 51), "batch_run")
            b.check(lib.fomo_batch_run(bh, 52, 39), "batch_run")
            b.synchronize()
            lib.fomo_batch_destroy(bh)
        finally:
            b.set_step_mode(1)
            b.set_split(16, 0)
            b.set_tiling(1, 0, 16)
        grew = [_bytes(h[1]) > h[3] for h in hs]
        assert all(grew) if name == "interiors" else not any(grew)
        for c, (dm, W, ds, _) in zip(cases, hs):
            r = c.oracle()
            r.run(1, c.nt)
            assert np.array_equal(ds.gather(), r.gather_nt_by_nrec())
            _state_equal(W, r, f"double batch {name}")


def test_double_step_then_stages_bit_exact(b):
    """After a run that ends on a double step, the strip backup the per-stage
    API reads comes from the third buffer: a single time_step and then the
    per-stage kernels continue the shot bit for bit."""
    c = Case(nx_in=151, nz_in=97, nbc=14, nt=44, src_type="force_x", src_xz=(F32(900.0), F32(500.0)))
    print("case double then time_step and stages", flush=True)
    b.set_tiling(2, 0, 16)
    b.set_step_mode(3)
    b.set_split(16, 1)
    try:
        W = E.Wavefield(c.M.nx, c.M.nz, b)
        dm = E.bind(b, W, c.M, c.H, c.a, c.p)
        ds = E._shot_for(dm, c.src, c.rec, c.nt)
        b.check(b.lib.fomo_run_time_loop(dm.h, W._h, ds.h, 1, 40), "run 40")
        r = c.oracle()
        r.run(1, 40)
        _state_equal(W, r, "after 40 double-step steps")  # old fields from the third buffer
        E.time_step(b, W, c.M, c.H, c.a, c.src, c.rec, 41, c.p)
        E.backup_boundary(b, W)
        E.update_velocity(b, W)
        E.apply_habc_velocity(b, W)
        b.synchronize()
    finally:
        b.set_step_mode(1)
        b.set_split(16, 0)
        b.set_tiling(1, 0, 16)
    import oracle as O
    r.run(41, 1)
    r.run(42, 1, O.STAGES["backup"] | O.STAGES["velocity"] | O.STAGES["habc_v"])
    _state_equal(W, r, "time_step and stages after double steps")


@pytest.mark.parametrize("skip,slots", [(4, 2), (6, 3), (3, 2)])
def test_double_step_frames_ring_wraps(b, monkeypatch, skip, slots):
    """A frame ring of a few slots splits the run into segments of
    slots * skip steps; with an even skip every segment starts on an odd
    step and takes double steps, its last segment ending on an even or odd
    step; with an odd skip the segments take split steps.  Frames, gather
    and final state equal the oracle's."""
    from fomo_julia_amd import video as V
    from tests.test_frames import _check, _oracle_frames
    names = ["vx", "p"]
    c = Case(nx_in=151, nz_in=97, nbc=14, nt=41)
    rec = V.MultiFieldRecorder(c.M.nx, c.M.nz, c.dt, V.VideoConfig(fields=names, skip=skip, downsample=3))
    per = len(names) * rec.nx_d * rec.nz_d * 4
    monkeypatch.setattr(V, "RING_BYTES", slots * per)
    print(f"case double frames ring skip={skip} slots={slots}", flush=True)
    b.set_tiling(2, 0, 8)
    b.set_step_mode(3)
    b.set_split(16, 1)
    try:
        W = E.Wavefield(c.M.nx, c.M.nz, b)
        res = E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p, on_step=rec)
    finally:
        b.set_step_mode(1)
        b.set_split(16, 0)
        b.set_tiling(1, 0, 16)
    want, r = _oracle_frames(c, names, skip, 3, c.M.nx, c.M.nz)
    done = (c.nt // skip) * skip
    if done < c.nt:
        r.run(done + 1, c.nt - done)
    assert rec.device_frames == c.nt // skip
    _check(rec, want, names)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    _state_equal(W, r, f"frames ring skip {skip}")


@pytest.mark.parametrize("skip", [6, 5])
def test_double_step_frames_and_batched_run_shots(b, skip):
    """On-step frames in step mode 3: with an even skip every frame step is the
    second step of a pair and the frames are taken inside the double steps;
    with an odd skip the run takes split steps.  Both equal the oracle's
    frames.  run_shots batched in mode 3 equals run_shots one shot at a time
    in mode 1."""
    from fomo_julia_amd import video as V
    from tests.test_frames import _check, _oracle_frames
    names = ["p", "vz"]
    c = Case(nx_in=151, nz_in=97, nbc=14, nt=42)
    print(f"case double frames (skip {skip}) + batched run_shots", flush=True)
    xs = [F32(500.0), F32(750.0), F32(1000.0)]
    cfg = E.MultiShotConfig(xs, [F32(40.0)] * 3, c.src.wavelet)
    b.set_tiling(2, 0, 16)
    b.set_step_mode(3)
    b.set_split(16, 1)
    try:
        rec = V.MultiFieldRecorder(c.M.nx, c.M.nz, c.dt, V.VideoConfig(fields=names, skip=skip, downsample=2))
        W = E.Wavefield(c.M.nx, c.M.nz, b)
        res = E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p, on_step=rec)
        Wb = E.Wavefield(c.M.nx, c.M.nz, b)
        bat = E.run_shots(b, Wb, c.M, c.H, c.a, c.rec, cfg, c.p, batch=3)
    finally:
        b.set_step_mode(1)
        b.set_split(16, 0)
        b.set_tiling(1, 0, 16)
    want, r = _oracle_frames(c, names, skip, 2, c.M.nx, c.M.nz)
    done = (c.nt // skip) * skip  # the oracle ran up to the last frame step
    if done < c.nt:
        r.run(done + 1, c.nt - done)
    assert rec.device_frames == c.nt // skip
    _check(rec, want, names)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    _state_equal(W, r, f"frames skip {skip}")
    Ws = E.Wavefield(c.M.nx, c.M.nz, b)
    seq = E.run_shots(b, Ws, c.M, c.H, c.a, c.rec, cfg, c.p, batch=1)
    for s, t in zip(seq, bat):
        assert (s.src_i, s.src_j) == (t.src_i, t.src_j)
        assert np.array_equal(s.gather, t.gather)


def test_double_step_stored_coefficient_model(b):
    """A medium whose mu_txz is not init_medium's harmonic average in one
    interior cell keeps only the five stored coefficient arrays; the double
    step reads those (as it does for every FP32 medium) and stays bit-exact,
    the perturbed cell included."""
    from tests.test_gpu_parity import _perturbed_case
    c = _perturbed_case(nx_in=151, nz_in=97, nbc=14, nt=50, src_type="force_z",
                        src_xz=(F32(760.0), F32(480.0)))
    print("case double step, stored-coefficient medium", flush=True)
    res, W, grew = _run(b, c)
    assert grew, "the double step did not run"
    r = c.oracle()
    r.run(1, c.nt)
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    _state_equal(W, r, "double, stored coefficients")
