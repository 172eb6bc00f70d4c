"""Wavefield frames: the on_step video callback (video_recorder.jl) on the GPU.

CPU: the oracle's frame extraction pinned against element-by-element loops
written in the reference's 1-based form, closed forms (ds = 1 identity,
3-4-5 magnitude, p of equal stresses) and edge clamping; the recorder's
configuration, titles and frame rate.

GPU: frames taken inside the graph-replayed loop (fomo_run_time_loop_frames),
through a small device ring that wraps many times, and through the plain
per-step callback (fomo_wavefield_frame) are compared bit for bit with the
oracle's state at the same steps; the gather of a recorded run equals that of
an unrecorded one."""
import numpy as np
import pytest

from fomo_julia_amd import elastic2d as E
from fomo_julia_amd import video as V
from oracle import frames as OF
from tests.helpers import Case

F32 = np.float32
gpu = pytest.mark.gpu


def _loop_frame(raw, ds, nx_frame, nz_frame):
    """_extract_frame! as written: 1-based loops, min((i-1)*ds + 1, nx)."""
    nx, nz = raw.shape
    nx_d, nz_d = -(-nx_frame // ds), -(-nz_frame // ds)
    out = np.zeros((nx_d, nz_d), F32)
    for i in range(1, nx_d + 1):
        for j in range(1, nz_d + 1):
            i_src = min((i - 1) * ds + 1, nx)
            j_src = min((j - 1) * ds + 1, nz)
            out[i - 1, j - 1] = raw[i_src - 1, j_src - 1]
    return out


def test_extract_frame_matches_reference_loops():
    rng = np.random.default_rng(4)
    raw = rng.standard_normal((11, 7)).astype(F32)
    for ds in (1, 2, 3, 4, 5, 12):
        for nxf, nzf in ((11, 7), (14, 9), (8, 5), (1, 1)):
            got = OF.extract_frame(raw, ds, nxf, nzf)
            assert got.shape == (-(-nxf // ds), -(-nzf // ds))
            assert np.array_equal(got, _loop_frame(raw, ds, nxf, nzf)), (ds, nxf, nzf)
    assert np.array_equal(OF.extract_frame(raw, 1, 11, 7), raw)
    # a frame wider than the grid repeats the last column
    wide = OF.extract_frame(raw, 2, 16, 7)
    assert np.array_equal(wide[-1], raw[-1, ::2]) and np.array_equal(wide[-3], raw[-1, ::2])


def test_field_data_closed_forms():
    vx = np.full((3, 2), 3.0, F32)
    vz = np.full((3, 2), -4.0, F32)
    assert np.all(OF.field_data(vx, vz, vx, vx, "vel") == 5.0)
    x = np.array([[1.5, -2.25], [1e-3, 7.0]], F32)
    assert np.array_equal(OF.field_data(x, x, x, x, "p"), x)
    assert np.array_equal(OF.field_data(x, -x, x, -x, "p"), np.zeros_like(x))
    assert np.array_equal(OF.field_data(x, vx[:2], x, x, "vx"), x)
    tiny = np.array([[1e-30]], F32)  # squares underflow in Float32 and sqrt(0) = 0
    assert OF.field_data(tiny, tiny, tiny, tiny, "vel")[0, 0] == 0.0
    with pytest.raises(ValueError):
        OF.field_data(x, x, x, x, "txz")


def test_video_config_and_titles():
    c = V.VideoConfig()
    assert c.fields == ["p"] and c.skip == 1 and c.downsample == 1 and c.clim is None
    assert c.time_scale == F32(0.01) and c.prefix == "wavefield" and c.colormap == "balance"
    assert V.VideoConfig(fields=[":p", ":vel"]).fields == ["p", "vel"]
    with pytest.raises(ValueError):
        V.VideoConfig(fields=["txz"])
    with pytest.raises(ValueError):
        V.VideoConfig(skip=0)
    assert V.field_title("p", F32(0.0125), 10, 800) == \
        "Pressure (Txx+Tzz)/2 | Time: 0.0125 s | Step: 10/800"
    assert V.field_title("vel", F32(1.0), 800, 800) == \
        "Velocity Magnitude | Time: 1.0000 s | Step: 800/800"
    assert V.field_label("vel") == "|V|"
    # run.jl's config: dt 0.00125, skip 10, time_scale 0.01 -> 8000 fps, clamped
    assert V.video_fps(F32(0.00125), 10, F32(0.01)) == 120
    assert V.video_fps(F32(1e-3), 50, F32(0.5)) == 40
    assert V.video_fps(F32(0.1), 10, F32(10.0)) == 1
    rec = V.MultiFieldRecorder(308, 201, F32(1e-3), V.VideoConfig(fields=["p", "vz"], downsample=2))
    assert (rec.nx_d, rec.nz_d) == (154, 101)


# ---------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def b():
    bk = E.backend("mi355x")
    bk.set_tiling(1, 0, 16)
    bk.set_step_mode(1)
    return bk


def _oracle_frames(c, names, skip, ds, nxf, nzf, storage="f32", fscale=None):
    r = c.oracle(storage=storage, fscale=fscale)
    out = {n: [] for n in names}
    k = 1
    while k + skip - 1 <= c.nt:
        r.run(k, skip)
        k += skip
        fr = OF.frames_of_state([r.field(q) for q in range(5)], names, ds, nxf, nzf)
        for n in names:
            out[n].append(fr[n])
    return out, r


def _check(rec, want, names):
    for n in names:
        got = rec.frames(n)
        assert got.shape[0] == len(want[n]), (n, got.shape, len(want[n]))
        for q, w in enumerate(want[n]):
            if not np.array_equal(got[q], w):
                bad = np.argwhere(got[q] != w)
                i, j = bad[0]
                raise AssertionError(f"{n} frame {q}: {len(bad)} cells differ, first ({i},{j}) "
                                     f"got {got[q][i, j]!r} want {w[i, j]!r}")


@gpu
@pytest.mark.parametrize("storage", ["f32", "bf16", "f16"])
def test_graph_frames_match_oracle(b, storage):
    names = ["p", "vel", "vx", "vz"]
    c = Case(nx_in=70, nz_in=52, nt=64, fd_order=8, storage=storage, water=True)
    print(f"case frames graph storage={storage}", flush=True)
    skip, ds = 7, 3
    rec = V.MultiFieldRecorder(c.M.nx + 2, c.M.nz, c.dt, V.VideoConfig(fields=names, skip=skip,
                                                                       downsample=ds))
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    res = E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p, on_step=rec)
    assert rec.device_frames == c.nt // skip
    assert rec.recorders["p"].steps == [skip * (q + 1) for q in range(c.nt // skip)]
    W2 = E.Wavefield(c.M.nx, c.M.nz, b)
    plain = E.run_shot(b, W2, c.M, c.H, c.a, c.src, c.rec, c.p)
    assert np.array_equal(res.gather, plain.gather)
    fscale = None
    if storage != "f32":
        dm = E._model(b, c.M, c.H, c.a, c.p)
        fscale = [float(s) for s in E._shot_for(dm, c.src, c.rec, c.nt).field_scales()]
    want, r = _oracle_frames(c, names, skip, ds, c.M.nx + 2, c.M.nz, storage, fscale)
    assert np.abs(want["p"][-1]).max() > 0
    _check(rec, want, names)


@gpu
def test_frame_ring_wraps(b, monkeypatch):
    names = ["vel", "p"]
    c = Case(nx_in=64, nz_in=48, nt=75, fd_order=4, fs=False)
    print("case frames ring", flush=True)
    skip, ds = 4, 2
    rec = V.MultiFieldRecorder(c.M.nx, c.M.nz, c.dt, V.VideoConfig(fields=names, skip=skip,
                                                                   downsample=ds))
    per = len(names) * rec.nx_d * rec.nz_d * 4
    monkeypatch.setattr(V, "RING_BYTES", 3 * per)  # 3 slots for 18 frames
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p, on_step=rec)
    want, _ = _oracle_frames(c, names, skip, ds, c.M.nx, c.M.nz)
    _check(rec, want, names)


@gpu
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_callback_frames_equal_graph_frames(b, storage):
    names = ["p", "vz"]
    c = Case(nx_in=60, nz_in=40, nt=30, fd_order=8, storage=storage)
    print(f"case frames callback storage={storage}", flush=True)
    cfg = V.VideoConfig(fields=names, skip=5, downsample=1)
    graph = V.MultiFieldRecorder(c.M.nx, c.M.nz, c.dt, cfg)
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p, on_step=graph)
    cb = V.MultiFieldRecorder(c.M.nx, c.M.nz, c.dt, cfg)
    seen = []

    def on_step(Wk, info):  # a plain callback: per-step segments, the recorder called on W
        seen.append(info.k)
        return cb(Wk, info)

    W2 = E.Wavefield(c.M.nx, c.M.nz, b)
    E.run_shot(b, W2, c.M, c.H, c.a, c.src, c.rec, c.p, on_step=on_step)
    assert seen == list(range(1, c.nt + 1))
    assert cb.device_frames == 0
    for n in names:
        assert np.array_equal(cb.frames(n), graph.frames(n)), n
    assert cb.recorders["p"].titles == graph.recorders["p"].titles
    fscale = None
    if storage != "f32":  # the callback decodes with the state's own FP16 scales
        dm = E._model(b, c.M, c.H, c.a, c.p)
        fscale = [float(v) for v in E._shot_for(dm, c.src, c.rec, c.nt).field_scales()]
    want, _ = _oracle_frames(c, names, 5, 1, c.M.nx, c.M.nz, storage, fscale)
    _check(cb, want, names)


@gpu
def test_recorder_close_writes_frames(b, tmp_path):
    c = Case(nx_in=50, nz_in=40, nt=20, fd_order=4)
    rec = V.MultiFieldRecorder(c.M.nx, c.M.nz, c.dt,
                               V.VideoConfig(fields=["p"], skip=4, downsample=2,
                                             output_dir=str(tmp_path), prefix="shot"))
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p, on_step=rec)
    rec.close()
    data = np.fromfile(tmp_path / "shot_p.f32", dtype="<f4").reshape(5, rec.nz_d, rec.nx_d)
    assert np.array_equal(data.transpose(0, 2, 1), rec.frames("p"))
    import json
    meta = json.loads((tmp_path / "shot_p.json").read_text())
    assert meta["steps"] == [4, 8, 12, 16, 20] and meta["frames"] == 5


@gpu
def test_run_shots_with_recorder_keeps_every_shot(b):
    """run_shots(...; on_step=recorder) as run.jl calls it: one video over all
    shots (frames of shot 1, then shot 2), gathers equal to plain runs."""
    c = Case(nx_in=60, nz_in=44, nt=24, fd_order=8)
    print("case frames run_shots", flush=True)
    cfg = E.MultiShotConfig(np.array([100.0, 400.0], F32), np.array([20.0, 30.0], F32),
                            c.src.wavelet)
    rec = V.MultiFieldRecorder(c.M.nx, c.M.nz, c.dt, V.VideoConfig(fields=["p"], skip=6,
                                                                   downsample=2))
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    seen = []
    res = E.run_shots(b, W, c.M, c.H, c.a, c.rec, cfg, c.p, on_step=rec,
                      on_shot_complete=lambda r: seen.append(r.shot_id))
    assert seen == [1, 2]
    assert rec.recorders["p"].steps == [6, 12, 18, 24] * 2
    plain = E.run_shots(b, E.Wavefield(c.M.nx, c.M.nz, b), c.M, c.H, c.a, c.rec, cfg, c.p, batch=1)
    for r0, r1 in zip(res, plain):
        assert np.array_equal(r0.gather, r1.gather)
    # each shot's frames equal a recorder run on that shot alone
    for q, shot in enumerate(cfg.shots):
        src = E._create_source(c.M, shot, cfg.wavelet, cfg.source_type)
        solo = V.MultiFieldRecorder(c.M.nx, c.M.nz, c.dt, V.VideoConfig(fields=["p"], skip=6,
                                                                        downsample=2))
        E.run_shot(b, E.Wavefield(c.M.nx, c.M.nz, b), c.M, c.H, c.a, src, c.rec, c.p, on_step=solo)
        assert np.array_equal(rec.frames("p")[4 * q:4 * q + 4], solo.frames("p")), q


@gpu
def test_frame_download_range_is_checked(b):
    """Ranges outside the ring are refused before any copy, including one whose
    end overflows 32 bits."""
    import ctypes
    print("case frames download range", flush=True)
    st = V._FrameStore(b, 32, 24, [V.FIELDS["p"]], 1, 1, 3)
    buf = np.zeros(1, F32)
    ptr = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    for first, count in [(-1, 1), (0, -1), (3, 1), (2, 2), (2**31 - 1, 1), (1, 2**31 - 1)]:
        assert b.lib.fomo_frames_download(st.h, first, count, ptr) != 0, (first, count)
    assert st.download(b, 1, 2).shape == (2, 1, 24, 32)
