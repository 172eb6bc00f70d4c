"""ORACLE -- test infrastructure only (tests/, smoke(), bench.py cpu_baseline).

Frame extraction of the reference's video callback, restated in numpy
Float32 (src/visualization/video_recorder.jl):

  _get_field_data (:270-283)   p = (txx + tzz) * 0.5f0, vel = sqrt(vx^2 + vz^2),
                                vx, vz; elementwise Float32 broadcasts (no
                                contraction: `@.` does not fuse x^2 + y^2)
  _extract_frame! (:260-268)   buffer[i, j] = raw[min((i-1)*ds + 1, nx),
                                min((j-1)*ds + 1, nz)] (1-based) over a buffer
                                of cld(nx_frame, ds) x cld(nz_frame, ds)
  _SingleRecorder call (:207-228) records when k % skip == 0

Pinned in tests/test_frames.py by element-by-element loops over tiny inputs,
the identity at ds = 1, the 3-4-5 magnitude and the edge clamping of indices.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
FIELDS = {"p": 0, "vel": 1, "vx": 2, "vz": 3}


def field_data(vx, vz, txx, tzz, field: str) -> np.ndarray:
    """_get_field_data (video_recorder.jl:270-283) on (nx, nz) Float32 arrays."""
    if field == "p":
        return ((txx.astype(F32) + tzz.astype(F32)) * F32(0.5)).astype(F32)
    if field == "vx":
        return vx.astype(F32)
    if field == "vz":
        return vz.astype(F32)
    if field == "vel":
        x, z = vx.astype(F32), vz.astype(F32)
        return np.sqrt((x * x + z * z).astype(F32)).astype(F32)
    raise ValueError(f"Unknown field: {field}")


def frame_shape(nx_frame: int, nz_frame: int, ds: int) -> tuple:
    """cld(n, downsample) (video_recorder.jl:175-176)."""
    return (-(-nx_frame // ds), -(-nz_frame // ds))


def extract_frame(raw: np.ndarray, ds: int, nx_frame: int, nz_frame: int) -> np.ndarray:
    """_extract_frame! (video_recorder.jl:260-268): the (nx_d, nz_d) buffer."""
    nx_d, nz_d = frame_shape(nx_frame, nz_frame, ds)
    ii = np.minimum(np.arange(nx_d) * ds, raw.shape[0] - 1)
    jj = np.minimum(np.arange(nz_d) * ds, raw.shape[1] - 1)
    return raw[np.ix_(ii, jj)].astype(F32)


def frames_of_state(fields5, names, ds, nx_frame, nz_frame) -> dict:
    """{name: frame} of one state given as the five (nx, nz) fields
    (vx, vz, txx, tzz, txz)."""
    vx, vz, txx, tzz = fields5[0], fields5[1], fields5[2], fields5[3]
    return {n: extract_frame(field_data(vx, vz, txx, tzz, n), ds, nx_frame, nz_frame)
            for n in names}
