"""Velocity-model ingestion (src/io/model_loader.jl; scripts/convert_model.jl).

``load_model(path; ...)`` dispatches on the extension as the reference does
(model_loader.jl:88-106) and returns a ``VelocityModel`` (:29-53) of Float32
(nx, nz) arrays; ``load_model_files`` (:465-557) reads vp / vs / rho from
separate files of mixed formats.  Missing vs and rho are estimated as the
reference does: vs = vp / 1.73f0, rho = 310f0 * vp^0.25f0 (Gardner).

Formats:
  .bin          raw little-endian samples, column-major (x fastest) or
                ``order="row_major"`` (z fastest, transposed on read)
  .segy / .sgy  SEG-Y rev 0/1, read here (the reference uses SegyIO.jl): the
                data matrix is (samples, traces), as ``segy_read(...).data``;
                sample formats 1 (IBM float), 2 (int32), 3 (int16), 5 (IEEE
                float) and 8 (int8), big-endian
  .npy          numpy arrays (no pickles)
  .mat          MATLAB v5 files through scipy.io.loadmat (v7.3 files are HDF5)
  .npz          this package's saved models (``save_model``)
  .jld2 / .h5   need an HDF5 library, which the image does not have: they
                raise with the reason

``save_model`` writes ``.npz`` (the reference writes JLD2, an HDF5 format).
Host code only: the arrays go to the device through ``init_medium`` or
``init_medium_device``.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

F32 = np.float32


@dataclass(eq=False)
class VelocityModel:
    """``VelocityModel`` (model_loader.jl:29-40)."""
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    dx: np.float32
    dz: np.float32
    nx: int
    nz: int
    x_origin: np.float32 = F32(0)
    z_origin: np.float32 = F32(0)
    name: str = "unnamed"


def velocity_model(vp, vs, rho, dx, dz, x_origin=0.0, z_origin=0.0, name="unnamed") -> VelocityModel:
    """The keyword constructor (model_loader.jl:43-53): sizes from vp, all Float32."""
    vp = np.asarray(vp)
    nx, nz = vp.shape
    if np.shape(vs) != (nx, nz):
        raise ValueError("vs size mismatch")
    if np.shape(rho) != (nx, nz):
        raise ValueError("rho size mismatch")
    f = lambda a: np.asfortranarray(np.asarray(a, dtype=F32))
    return VelocityModel(f(vp), f(vs), f(rho), F32(dx), F32(dz), int(nx), int(nz), F32(x_origin),
                         F32(z_origin), str(name))


def estimate_vs(vp) -> np.ndarray:
    """vp ./ 1.73f0 (model_loader.jl:130, 178)."""
    return (np.asarray(vp, dtype=F32) / F32(1.73)).astype(F32)


def gardner_rho(vp) -> np.ndarray:
    """310.0f0 .* (vp .^ 0.25f0) (model_loader.jl:134, 187).  The power is
    taken in Float64 and rounded to Float32 before the Float32 product."""
    v = np.asarray(vp, dtype=np.float64)
    if np.any(v < 0):  # Julia: DomainError for a negative base and a fractional power
        raise ValueError("Gardner relation needs vp >= 0")
    p = np.power(v, 0.25).astype(F32)
    return (F32(310.0) * p).astype(F32)


# ---------------------------------------------------------------- readers
def read_bin(path, nx, nz, dtype=np.float32, order="column_major") -> np.ndarray:
    """``read_bin`` (model_loader.jl:155-165): nx*nz samples, column-major
    (nx, nz), or row-major data transposed (permutedims(reshape(data, nz, nx)))."""
    if nx is None or nz is None:
        raise ValueError("Binary files require nx and nz parameters")
    n = int(nx) * int(nz)
    dt = np.dtype(dtype).newbyteorder("<")
    data = np.fromfile(path, dtype=dt, count=n)
    if data.size != n:
        raise EOFError(f"{path}: expected {n} samples, found {data.size}")
    if str(order).lstrip(":") == "row_major":
        return np.asfortranarray(data.reshape((int(nz), int(nx)), order="F").T)
    return data.reshape((int(nx), int(nz)), order="F")


def ibm_to_float32(words: np.ndarray) -> np.ndarray:
    """IBM System/360 single precision (sign, base-16 exponent excess 64,
    24-bit fraction) to Float32: exact for every value in Float32's range."""
    w = np.asarray(words, dtype=np.uint32)
    sign = np.where(w >> 31, -1.0, 1.0)
    expo = ((w >> 24) & 0x7F).astype(np.int64) - 64
    frac = (w & 0x00FFFFFF).astype(np.float64) / float(1 << 24)
    return (sign * frac * np.power(16.0, expo)).astype(F32)


_SEGY_FORMATS = {1: (4, None), 2: (4, ">i4"), 3: (2, ">i2"), 5: (4, ">f4"), 8: (1, "i1")}


def read_segy(path) -> np.ndarray:
    """The (samples, traces) Float32 data matrix of a SEG-Y file (what
    ``SegyIO.segy_read(path).data`` holds): 3200-byte textual header, 400-byte
    binary header (samples per trace at bytes 3221-3222, format code at
    3225-3226, extended textual headers at 3505-3506), then traces of a
    240-byte header and the samples."""
    with open(path, "rb") as f:
        head = f.read(3600)
        if len(head) < 3600:
            raise EOFError(f"{path}: shorter than the SEG-Y file headers")
        ns = struct.unpack(">H", head[3220:3222])[0]
        fmt = struct.unpack(">H", head[3224:3226])[0]
        n_ext = struct.unpack(">h", head[3504:3506])[0]
        if fmt not in _SEGY_FORMATS:
            raise ValueError(f"{path}: SEG-Y sample format {fmt} is not supported")
        if ns == 0:
            raise ValueError(f"{path}: zero samples per trace in the binary header")
        size, dt = _SEGY_FORMATS[fmt]
        f.seek(3600 + 3200 * max(n_ext, 0))
        body = f.read()
    rec = 240 + ns * size
    ntr = len(body) // rec
    if ntr == 0 or len(body) % rec:
        raise ValueError(f"{path}: {len(body)} trace bytes are not whole traces of {ns} samples")
    raw = np.frombuffer(body, dtype=np.uint8).reshape(ntr, rec)[:, 240:]
    if fmt == 1:
        vals = ibm_to_float32(raw.copy().view(">u4").astype(np.uint32))
    else:
        vals = raw.copy().view(dt).astype(F32)
    return np.asfortranarray(vals.reshape(ntr, ns).T)  # column k = trace k


def _first_matrix(d: dict, keys=None):
    for k in keys or []:
        if k in d:
            return d[k]
    for k, v in d.items():
        if not k.startswith("__") and isinstance(v, np.ndarray) and v.ndim == 2:
            return v
    return None


def _load_mat_dict(path) -> dict:
    import scipy.io
    try:
        return scipy.io.loadmat(path)
    except NotImplementedError as e:  # v7.3 files are HDF5
        raise ValueError(f"{path}: MATLAB v7.3 (HDF5) files need an HDF5 library") from e


def _no_hdf5(path, what):
    raise ValueError(f"{path}: {what} files are HDF5 containers and need an HDF5 library, "
                     "which this installation does not have; convert to .npz, .npy or .bin")


# ---------------------------------------------------------------- load_model
def load_model(path: str, **kw):
    """``load_model(path; kwargs...)`` (model_loader.jl:88-106)."""
    ext = os.path.splitext(path)[1].lower()
    if ext == ".bin":
        return _load_binary(path, **kw)
    if ext in (".segy", ".sgy"):
        return _load_segy(path, **kw)
    if ext == ".npy":
        return _load_npy(path, **kw)
    if ext == ".mat":
        return _load_mat(path, **kw)
    if ext == ".npz":
        return _load_npz(path, **kw)
    if ext == ".jld2":
        _no_hdf5(path, "JLD2")
    if ext in (".h5", ".hdf5"):
        _no_hdf5(path, "HDF5")
    raise ValueError(f"Unsupported format: {ext}\n"
                     "Supported: .jld2, .segy, .sgy, .bin, .mat, .npy, .h5, .hdf5")


def _resolve(x, reader, vp, fallback):
    if isinstance(x, str):
        return reader(x)
    if x is None:
        return fallback(vp)
    return np.asarray(x, dtype=F32)


def _load_binary(path, nx=None, nz=None, dx=1.0, dz=None, dtype=np.float32, order="column_major",
                 vs=None, rho=None, single=False, **_):
    """``_load_binary`` (model_loader.jl:144-192)."""
    dz = dx if dz is None else dz
    rd = lambda p: read_bin(p, nx, nz, dtype, order)
    vp = rd(path)
    if single:
        return vp
    vs_d = _resolve(vs, rd, vp, estimate_vs)
    rho_d = _resolve(rho, rd, vp, gardner_rho)
    return velocity_model(vp, vs_d, rho_d, dx, dz, name=os.path.basename(path))


def _load_segy(path, dx=None, dz=None, vs=None, rho=None, **_):
    """``_load_segy`` (model_loader.jl:199-248): vs / rho as further SEG-Y files."""
    vp = read_segy(path)
    dx = 1.0 if dx is None else dx
    dz = dx if dz is None else dz
    seg = lambda p: read_segy(p) if p.lower().endswith((".segy", ".sgy")) else None
    vs_d = _resolve(vs, seg, vp, estimate_vs)
    rho_d = _resolve(rho, seg, vp, gardner_rho)
    if vs_d is None or rho_d is None:
        raise ValueError("vs / rho paths of a SEG-Y model must be SEG-Y files")
    return velocity_model(vp, vs_d, rho_d, dx, dz, name=os.path.basename(path))


def _load_npy(path, dx=1.0, dz=None, vs=None, rho=None, **_):
    """``_load_npy`` (model_loader.jl:288-320)."""
    rd = lambda p: np.load(p, allow_pickle=False).astype(F32)
    vp = rd(path)
    dz = dx if dz is None else dz
    return velocity_model(vp, _resolve(vs, rd, vp, estimate_vs), _resolve(rho, rd, vp, gardner_rho),
                          dx, dz, name=os.path.basename(path))


def _load_mat(path, vp_key="vp", vs_key="vs", rho_key="rho", dx=1.0, dz=None, **_):
    """``_load_mat`` (model_loader.jl:254-282)."""
    d = _load_mat_dict(path)
    vp = d.get(vp_key, d.get("Vp", d.get("VP")))
    if vp is None:
        raise ValueError(f"Cannot find vp field. Available keys: {sorted(k for k in d if not k.startswith('__'))}")
    vp = np.asarray(vp, dtype=F32)
    vs = d.get(vs_key, d.get("Vs"))
    rho = d.get(rho_key, d.get("Rho", d.get("density")))
    dz = dx if dz is None else dz
    vs_d = estimate_vs(vp) if vs is None else np.asarray(vs, dtype=F32)
    rho_d = gardner_rho(vp) if rho is None else np.asarray(rho, dtype=F32)
    return velocity_model(vp, vs_d, rho_d, dx, dz, name=os.path.basename(path))


def _load_npz(path, **_):
    with np.load(path, allow_pickle=False) as d:
        return velocity_model(d["vp"], d["vs"], d["rho"], d["dx"], d["dz"],
                              name=str(d["name"]) if "name" in d else os.path.basename(path))


def save_model(path: str, model_or_vp, vs=None, rho=None, dx=None, dz=None, name="model") -> None:
    """``save_model`` (model_loader.jl:359-378), as .npz."""
    m = model_or_vp if isinstance(model_or_vp, VelocityModel) else \
        velocity_model(model_or_vp, vs, rho, dx, dz, name=name)
    if not path.lower().endswith(".npz"):
        raise ValueError("save_model writes .npz (JLD2 needs an HDF5 library)")
    np.savez(path, vp=m.vp, vs=m.vs, rho=m.rho, dx=m.dx, dz=m.dz, nx=m.nx, nz=m.nz, name=m.name)


def convert_model(input: str, output: str, **kw) -> VelocityModel:
    """``convert_model`` (model_loader.jl:397-402)."""
    m = load_model(input, **kw)
    save_model(output, m)
    return m


def model_info(model) -> str:
    """``model_info`` (model_loader.jl:413-427): the summary text (also printed)."""
    m = load_model(model) if isinstance(model, str) else model
    bar, thin = "═" * 50, "─" * 50
    f = lambda x: _jfloat(F32(x))
    lines = [bar, f"Model: {m.name}", bar, f"  Grid:     {m.nx} × {m.nz}",
             f"  Spacing:  dx={f(m.dx)}m, dz={f(m.dz)}m",
             f"  Size:     {f(F32(m.nx) * m.dx)}m × {f(F32(m.nz) * m.dz)}m", thin,
             f"  Vp:   {f(m.vp.min())} - {f(m.vp.max())} m/s",
             f"  Vs:   {f(m.vs.min())} - {f(m.vs.max())} m/s",
             f"  Rho:  {f(m.rho.min())} - {f(m.rho.max())} kg/m³", bar]
    text = "\n".join(lines)
    print(text)
    return text


def _jfloat(x) -> str:
    """A Float32 as Julia prints it (shortest digits, '.0' for integral values)."""
    r = str(F32(x))
    if "e" in r:
        m, e = r.split("e")
        if "." not in m:
            m += ".0"
        return f"{m}e{int(e)}"
    return r if "." in r else r + ".0"


def load_model_files(vp: str, vs=None, rho=None, dx=None, dz=None, nx=None, nz=None,
                     order="column_major", name="model") -> VelocityModel:
    """``load_model_files`` (model_loader.jl:465-557): mixed formats per property."""
    if dx is None:
        raise ValueError("dx is required")
    dz = dx if dz is None else dz

    def read_file(path):
        ext = os.path.splitext(path)[1].lower()
        if ext in (".segy", ".sgy"):
            return read_segy(path)
        if ext == ".bin":
            return read_bin(path, nx, nz, np.float32, order)
        if ext == ".npy":
            return np.load(path, allow_pickle=False).astype(F32)
        if ext == ".mat":
            v = _first_matrix(_load_mat_dict(path))
            if v is None:
                raise ValueError(f"No matrix in MAT file: {path}")
            return np.asarray(v, dtype=F32)
        if ext == ".jld2":
            _no_hdf5(path, "JLD2")
        raise ValueError(f"Unsupported format: {ext}")

    vp_d = read_file(vp)
    vs_d = read_file(vs) if vs is not None else estimate_vs(vp_d)
    rho_d = read_file(rho) if rho is not None else gardner_rho(vp_d)
    return velocity_model(vp_d, vs_d, rho_d, F32(dx), F32(dz), name=name)
