"""Survey geometry of computed shots (src/io/geometry_io.jl).

``create_geometry`` turns ShotResults into the source and receiver positions
actually used on the grid (geometry_io.jl:85-130), with the reference's
arithmetic: ``(index - pad - 1) * dx`` in Float32, physical indices
``index - pad`` (which the reference labels 0-based although the first
physical cell is 1), grid size without padding, ``t_max = dt * nt`` in
Float32.  ``save_geometry`` writes the reference's JSON and text layouts
(geometry_io.jl:236-372); its JLD2 layout needs JLD2/HDF5, which this
package does not use, so ``load_geometry`` reads the JSON layout instead.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32
_NOTE = "All coordinates are ACTUAL discretized grid positions"


@dataclass
class SurveyGeometry:
    """``SurveyGeometry`` (geometry_io.jl:26-52)."""
    shot_id: int
    src_x: np.float32
    src_z: np.float32
    src_i: int
    src_j: int
    n_rec: int
    rec_x: np.ndarray
    rec_z: np.ndarray
    rec_i_idx: list
    rec_j_idx: list
    dt: np.float32
    nt: int
    t_max: np.float32
    dx: np.float32
    dz: np.float32
    nx: int
    nz: int


@dataclass
class MultiShotGeometry:
    """``MultiShotGeometry`` (geometry_io.jl:59-70)."""
    n_shots: int
    shots: list = field(default_factory=list)
    dt: np.float32 = F32(0)
    nt: int = 0
    dx: np.float32 = F32(0)
    dz: np.float32 = F32(0)
    nx: int = 0
    nz: int = 0


def _one(result, medium, params) -> SurveyGeometry:
    pad = int(medium.pad)
    dx, dz = F32(medium.dx), F32(medium.dz)
    si, sj = int(result.src_i), int(result.src_j)
    ri = [int(v) for v in result.rec_i]
    rj = [int(v) for v in result.rec_j]
    # Int * Float32 is Float32 in Julia
    rec_x = np.array([F32(i - pad - 1) * dx for i in ri], dtype=F32)
    rec_z = np.array([F32(j - pad - 1) * dz for j in rj], dtype=F32)
    dt, nt = F32(params.dt), int(params.nt)
    return SurveyGeometry(int(result.shot_id), F32(F32(si - pad - 1) * dx), F32(F32(sj - pad - 1) * dz),
                          si - pad, sj - pad, len(ri), rec_x, rec_z, [i - pad for i in ri],
                          [j - pad for j in rj], dt, nt, F32(dt * F32(nt)), dx, dz,
                          int(medium.nx) - 2 * pad, int(medium.nz) - 2 * pad)


def create_geometry(result_or_results, medium, params):
    """``create_geometry(result, medium, params)`` (geometry_io.jl:85-112) or,
    for a list of results, the ``MultiShotGeometry`` (geometry_io.jl:119-130)."""
    if isinstance(result_or_results, (list, tuple)):
        shots = [_one(r, medium, params) for r in result_or_results]
        pad = int(medium.pad)
        return MultiShotGeometry(len(shots), shots, F32(params.dt), int(params.nt), F32(medium.dx),
                                 F32(medium.dz), int(medium.nx) - 2 * pad, int(medium.nz) - 2 * pad)
    return _one(result_or_results, medium, params)


def _f(x) -> str:
    """A Float32 as Julia prints it (shortest round-trip digits, '.0' for
    integers, 'e-5' style exponents)."""
    r = str(F32(x))
    # numpy prints the shortest Float32 digits; normalise to Julia's forms
    if "e" in r or "E" in r:
        m, e = r.lower().split("e")
        if "." not in m:
            m += ".0"
        return f"{m}e{int(e)}"
    return r if "." in r else r + ".0"


def _join(v) -> str:
    return ", ".join(_f(x) if isinstance(x, (float, np.floating)) else str(int(x)) for x in v)


def save_geometry(path: str, geom) -> None:
    """``save_geometry(path, geom)`` (geometry_io.jl:154-181): .json or
    .txt/.dat/.geo."""
    ext = os.path.splitext(path)[1].lower()
    multi = isinstance(geom, MultiShotGeometry)
    if ext == ".json":
        text = _multi_json(geom) if multi else _single_json(geom)
    elif ext in (".txt", ".dat", ".geo"):
        text = _multi_txt(geom) if multi else _single_txt(geom)
    elif ext == ".jld2":
        raise ValueError("JLD2 output needs JLD2/HDF5, which this package does not use; "
                         "use .json or .txt")
    else:
        raise ValueError(f"Unsupported format: {ext}. Use .json or .txt")
    with open(path, "w") as f:
        f.write(text)


def _single_json(g: SurveyGeometry) -> str:
    L = ["{", f'  "_note": "{_NOTE}",', f'  "shot_id": {g.shot_id},', '  "source": {',
         f'    "x": {_f(g.src_x)}, "z": {_f(g.src_z)},', f'    "i": {g.src_i}, "j": {g.src_j}', "  },",
         '  "receivers": {', f'    "n": {g.n_rec},', f'    "x": [{_join(g.rec_x)}],',
         f'    "z": [{_join(g.rec_z)}],', f'    "i": [{_join(g.rec_i_idx)}],',
         f'    "j": [{_join(g.rec_j_idx)}]', "  },",
         f'  "time": {{"dt": {_f(g.dt)}, "nt": {g.nt}, "t_max": {_f(g.t_max)}}},',
         f'  "grid": {{"dx": {_f(g.dx)}, "dz": {_f(g.dz)}, "nx": {g.nx}, "nz": {g.nz}}}', "}"]
    return "\n".join(L) + "\n"


def _multi_json(mg: MultiShotGeometry) -> str:
    L = ["{", f'  "_note": "{_NOTE}",', f'  "n_shots": {mg.n_shots},', '  "common": {',
         f'    "dt": {_f(mg.dt)}, "nt": {mg.nt},', f'    "dx": {_f(mg.dx)}, "dz": {_f(mg.dz)},',
         f'    "nx": {mg.nx}, "nz": {mg.nz}', "  },", '  "shots": [']
    for k, s in enumerate(mg.shots):
        comma = "," if k < mg.n_shots - 1 else ""
        L += ["    {", f'      "id": {s.shot_id},', f'      "src_x": {_f(s.src_x)}, "src_z": {_f(s.src_z)},',
              f'      "src_i": {s.src_i}, "src_j": {s.src_j},', f'      "n_rec": {s.n_rec},',
              f'      "rec_x": [{_join(s.rec_x)}],', f'      "rec_z": [{_join(s.rec_z)}],',
              f'      "rec_i": [{_join(s.rec_i_idx)}],', f'      "rec_j": [{_join(s.rec_j_idx)}]',
              f"    }}{comma}"]
    L += ["  ]", "}"]
    return "\n".join(L) + "\n"


def _single_txt(g: SurveyGeometry) -> str:
    L = ["# Survey Geometry - Single Shot", "# Generated by Elastic2D",
         "# NOTE: All coordinates are ACTUAL DISCRETIZED grid positions!", "#" + "=" * 60, "",
         "# Shot Info", f"shot_id      {g.shot_id}", "", "# Source (actual discretized position)",
         f"src_x        {float(g.src_x):.4f}    # meters", f"src_z        {float(g.src_z):.4f}    # meters",
         f"src_i        {g.src_i}        # grid index (0-based)",
         f"src_j        {g.src_j}        # grid index (0-based)", "", "# Time Info",
         f"dt           {float(g.dt):.6f}  # seconds", f"nt           {g.nt}",
         f"t_max        {float(g.t_max):.4f}    # seconds", "", "# Grid Info (physical domain)",
         f"dx           {float(g.dx):.4f}    # meters", f"dz           {float(g.dz):.4f}    # meters",
         f"nx           {g.nx}", f"nz           {g.nz}", "",
         f"# Receivers: {g.n_rec} total (actual discretized positions)",
         "# Format: rec_id  x(m)  z(m)  i_idx  j_idx", f"n_rec        {g.n_rec}", "receivers"]
    for r in range(g.n_rec):
        L.append(f"{r + 1:6d}  {float(g.rec_x[r]):12.4f}  {float(g.rec_z[r]):12.4f}  "
                 f"{g.rec_i_idx[r]:6d}  {g.rec_j_idx[r]:6d}")
    return "\n".join(L) + "\n"


def _multi_txt(mg: MultiShotGeometry) -> str:
    L = ["# Survey Geometry - Multi Shot", "# Generated by Elastic2D",
         "# NOTE: All coordinates are ACTUAL DISCRETIZED grid positions!", "#" + "=" * 60, "",
         "# Common Parameters", f"n_shots      {mg.n_shots}", f"dt           {float(mg.dt):.6f}  # seconds",
         f"nt           {mg.nt}", f"dx           {float(mg.dx):.4f}    # meters",
         f"dz           {float(mg.dz):.4f}    # meters", f"nx           {mg.nx}", f"nz           {mg.nz}", "",
         "# Shot List (actual discretized positions)",
         "# Format: shot_id  src_x(m)  src_z(m)  src_i  src_j  n_rec", "shots"]
    for s in mg.shots:
        L.append(f"{s.shot_id:6d}  {float(s.src_x):12.4f}  {float(s.src_z):12.4f}  "
                 f"{s.src_i:6d}  {s.src_j:6d}  {s.n_rec:6d}")
    L += ["", "# Receiver Coordinates per Shot (actual discretized positions)",
          "# Format: rec_id  x(m)  z(m)  i_idx  j_idx"]
    for s in mg.shots:
        L += ["", f"# Shot {s.shot_id} receivers", f"shot_{s.shot_id}_receivers  {s.n_rec}"]
        for r in range(s.n_rec):
            L.append(f"{r + 1:6d}  {float(s.rec_x[r]):12.4f}  {float(s.rec_z[r]):12.4f}  "
                     f"{s.rec_i_idx[r]:6d}  {s.rec_j_idx[r]:6d}")
    return "\n".join(L) + "\n"


def load_geometry(path: str):
    """Reads a geometry written by ``save_geometry`` in the JSON layout."""
    if os.path.splitext(path)[1].lower() != ".json":
        raise ValueError("load_geometry reads the .json layout")
    with open(path) as f:
        d = json.load(f)
    if "n_shots" in d:
        c = d["common"]
        shots = []
        for s in d["shots"]:
            nt = int(c["nt"])
            shots.append(SurveyGeometry(
                int(s["id"]), F32(s["src_x"]), F32(s["src_z"]), int(s["src_i"]), int(s["src_j"]),
                int(s["n_rec"]), np.array(s["rec_x"], F32), np.array(s["rec_z"], F32),
                [int(v) for v in s["rec_i"]], [int(v) for v in s["rec_j"]], F32(c["dt"]), nt,
                F32(F32(c["dt"]) * F32(nt)), F32(c["dx"]), F32(c["dz"]), int(c["nx"]), int(c["nz"])))
        return MultiShotGeometry(int(d["n_shots"]), shots, F32(c["dt"]), int(c["nt"]), F32(c["dx"]),
                                 F32(c["dz"]), int(c["nx"]), int(c["nz"]))
    s, r, t, g = d["source"], d["receivers"], d["time"], d["grid"]
    return SurveyGeometry(int(d["shot_id"]), F32(s["x"]), F32(s["z"]), int(s["i"]), int(s["j"]),
                          int(r["n"]), np.array(r["x"], F32), np.array(r["z"], F32),
                          [int(v) for v in r["i"]], [int(v) for v in r["j"]], F32(t["dt"]),
                          int(t["nt"]), F32(t["t_max"]), F32(g["dx"]), F32(g["dz"]), int(g["nx"]),
                          int(g["nz"]))
