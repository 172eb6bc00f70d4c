"""Gather files and survey geometry (src/io/output.jl, src/io/geometry_io.jl)."""
import json

import numpy as np
import pytest

from fomo_julia_amd import elastic2d as E
from fomo_julia_amd import geometry as G
from fomo_julia_amd import output as O

F32 = np.float32


def _result(nt=7, n_rec=5, shot_id=3):
    g = (np.arange(nt * n_rec, dtype=F32).reshape((n_rec, nt)).T * F32(0.5)).copy(order="F")
    return E.ShotResult(g, shot_id, 155, 57, list(range(55, 55 + n_rec)), [56] * n_rec)


def test_save_gather_is_julia_column_major(tmp_path):
    r = _result()
    p = tmp_path / "shot.bin"
    O.save_gather(r, str(p))
    raw = np.fromfile(p, dtype="<f4")
    # Julia's write(io, ::Matrix{Float32}) emits element (k, r) at k + r*nt
    nt, n_rec = r.gather.shape
    for rr in range(n_rec):
        for k in range(nt):
            assert raw[k + rr * nt] == r.gather[k, rr]
    back = O.load_gather(str(p), nt, n_rec)
    assert back.shape == (nt, n_rec) and np.array_equal(back, r.gather)
    with pytest.raises(ValueError):
        O.save_gather(r, str(p), format="png")
    with pytest.raises(EOFError):
        O.load_gather(str(p), nt + 1, n_rec)


def test_reference_golden_gather_loads():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "shot_1.bin")
    g = O.load_gather(path, 800, 100)
    assert g.shape == (800, 100) and np.isfinite(g).all() and np.abs(g).max() > 0
    # the file is exactly the 80 000 values in column-major order
    assert np.array_equal(g.ravel(order="F"), np.fromfile(path, dtype="<f4"))


def test_create_geometry_follows_reference_arithmetic():
    class Med:  # the Medium fields create_geometry reads
        pad, dx, dz, nx, nz = 54, F32(10.0), F32(10.0), 308, 308

    class Par:
        dt, nt = F32(1e-3), 1000

    r = _result()
    g = G.create_geometry(r, Med, Par)
    assert (g.src_x, g.src_z) == (F32(1000.0), F32(20.0))  # (155-54-1)*10, (57-54-1)*10
    assert (g.src_i, g.src_j) == (101, 3)  # index - pad, labelled 0-based by the reference
    assert g.rec_x.tolist() == [F32(k * 10.0) for k in range(5)] and g.rec_z.tolist() == [F32(10.0)] * 5
    assert g.rec_i_idx == [1, 2, 3, 4, 5] and g.rec_j_idx == [2] * 5
    assert (g.nx, g.nz, g.nt) == (200, 200, 1000)
    assert g.t_max == F32(F32(1e-3) * F32(1000))
    mg = G.create_geometry([r, _result(shot_id=4)], Med, Par)
    assert mg.n_shots == 2 and [s.shot_id for s in mg.shots] == [3, 4] and mg.nx == 200


def test_geometry_json_and_text_layouts(tmp_path):
    class Med:
        pad, dx, dz, nx, nz = 54, F32(12.5), F32(12.5), 308, 308

    class Par:
        dt, nt = F32(0.00125), 800

    r = _result()
    g = G.create_geometry(r, Med, Par)
    p = tmp_path / "g.json"
    G.save_geometry(str(p), g)
    text = p.read_text()
    d = json.loads(text)
    assert d["source"] == {"x": 1250.0, "z": 25.0, "i": 101, "j": 3}
    assert '"time": {"dt": 0.00125, "nt": 800, "t_max": 1.0}' in text
    assert '"x": [0.0, 12.5, 25.0, 37.5, 50.0]' in text
    back = G.load_geometry(str(p))
    assert back.rec_x.tolist() == g.rec_x.tolist() and back.src_i == 101 and back.t_max == F32(1.0)
    mg = G.create_geometry([r, _result(shot_id=9)], Med, Par)
    G.save_geometry(str(tmp_path / "m.json"), mg)
    assert G.load_geometry(str(tmp_path / "m.json")).shots[1].shot_id == 9
    G.save_geometry(str(tmp_path / "g.txt"), g)
    lines = (tmp_path / "g.txt").read_text().splitlines()
    assert "src_x        1250.0000    # meters" in lines
    assert "dt           0.001250  # seconds" in lines
    assert "     1        0.0000       12.5000       1       2" in lines
    G.save_geometry(str(tmp_path / "m.txt"), mg)
    assert "shot_9_receivers  5" in (tmp_path / "m.txt").read_text()
    with pytest.raises(ValueError):
        G.save_geometry(str(tmp_path / "g.jld2"), g)
