"""Gather files: ``save_gather`` / ``load_gather`` (src/io/output.jl:18-33,60-70;
src/io/geometry_io.jl:469-475).

The binary format is the reference's: the [nt x n_rec] Float32 matrix written
as Julia writes it, column-major (all nt samples of receiver 1, then receiver
2, ...), little-endian, no header.  PNG output (CairoMakie in the reference)
is plotting and not part of this package.
"""
from __future__ import annotations

import numpy as np

F32 = np.dtype("<f4")


def save_gather(result, path: str, format: str = "bin") -> None:
    """``save_gather(result, path; format=:bin)``: `result` is a ShotResult
    (or anything with a ``gather`` [nt x n_rec] array) or the array itself."""
    fmt = str(format).lstrip(":")
    if fmt != "bin":
        if fmt == "png":
            raise ValueError("format=:png is plotting, which this package does not provide")
        raise ValueError(f"Unknown format: {format}")
    g = np.asarray(getattr(result, "gather", result), dtype=np.float32)
    if g.ndim != 2:
        raise ValueError("a gather is an [nt x n_rec] matrix")
    g.astype(F32).ravel(order="F").tofile(path)


def load_gather(path: str, nt_or_geometry, n_rec: int | None = None) -> np.ndarray:
    """``load_gather(path, nt, n_rec)`` or ``load_gather(path, geom)``:
    the [nt x n_rec] Float32 matrix (Fortran order, as Julia holds it)."""
    if n_rec is None:
        geom = nt_or_geometry
        nt, n_rec = int(geom.nt), int(geom.n_rec)
    else:
        nt = int(nt_or_geometry)
    data = np.fromfile(path, dtype=F32, count=nt * n_rec)
    if data.size != nt * n_rec:
        raise EOFError(f"{path}: expected {nt * n_rec} Float32 values, found {data.size}")
    return data.astype(np.float32).reshape((nt, n_rec), order="F")
