"""Host-side mirror of the reference's Elastic2D API for the MI355X backend.

Same names, argument meaning and error behaviour as the Julia module
(src/Elastic2D.jl:41-80), with ``!`` dropped from function names:

    b = backend("mi355x")                      # backends/backend.jl:39-53
    M = init_medium(vp, vs, rho, dx, dz, nbc, fd_order, b; free_surface=True)
    H = init_habc(M.nx, M.nz, nbc, dt, dx, dz, v_ref, b)
    p = SimParams(dt, nt, dx, dz, fd_order)
    W = Wavefield(M.nx, M.nz, b)
    rec = setup_receivers(x, z, M, type="vz")
    cfg = MultiShotConfig(xs, zs, wavelet, source_type="pressure")
    results = run_shots(b, W, M, H, get_fd_coefficients(fd_order), rec, cfg, p)

Host arrays are numpy (nx, nz) arrays in Julia index order.  All time
stepping runs in libfomo_mi355x.so; this module only prepares inputs (the
Float32 setup of src/utils/init.jl, restated in ``init.py``), owns the
library handles and copies results back.  There is no CPU compute path.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

from . import _lib
from .init import (F32, F64, _compute_staggered_params, _jl, _pad_array, coord_to_index,
                   get_fd_coefficients, habc_coefficients, habc_weight_matrix,
                   receiver_indices, ricker_wavelet)

__all__ = [
    "MI355XBackend", "backend", "to_device", "synchronize", "Medium", "HABCConfig",
    "SimParams", "Source", "Receivers", "Wavefield", "ShotConfig", "MultiShotConfig",
    "ShotResult", "TimeStepInfo", "init_medium", "init_habc", "setup_receivers",
    "get_fd_coefficients", "ricker_wavelet", "time_step", "run_time_loop", "run_shot",
    "run_shots", "update_velocity", "update_stress", "backup_boundary", "apply_habc_velocity",
    "apply_habc_stress", "apply_free_surface", "inject_source", "record_receivers", "reset",
]


def _f32ptr(a: np.ndarray):
    return a.ctypes.data_as(_lib.P_f32)


def _i32ptr(a: np.ndarray):
    return a.ctypes.data_as(_lib.P_i32)


def _host_f32(a) -> np.ndarray:
    """(nx, nz) array -> contiguous Float32 in Julia memory order (i fastest)."""
    return np.asfortranarray(np.asarray(a, dtype=F32))


# ---------------------------------------------------------------- backend
class MI355XBackend:
    """``MI355XBackend <: AbstractBackend``: one GPU, one HIP stream (fomo_ctx)."""

    name = "mi355x"

    def __init__(self, device: int = 0):
        self.device = device
        self._ctx = None
        self._models: dict = {}

    @property
    def lib(self):
        return _lib.load()

    @property
    def ctx(self):
        if self._ctx is None:
            h = ctypes.c_void_p()
            rc = self.lib.fomo_ctx_create(self.device, ctypes.byref(h))
            _lib.check(rc, None, "fomo_ctx_create")
            self._ctx = h
        return self._ctx

    def check(self, rc: int, what: str) -> None:
        _lib.check(rc, self._ctx, what)

    def synchronize(self) -> None:
        self.check(self.lib.fomo_ctx_synchronize(self.ctx), "synchronize")

    def set_tiling(self, cols_per_lane: int = 1, rows_per_tile: int = 0,
                   steps_per_graph: int = 16) -> None:
        """Stencil tile shape (rows_per_tile 0 = automatic) and steps per captured graph (even).
        Batches and pools take the shape current when they are built."""
        self.check(self.lib.fomo_ctx_set_tiling(self.ctx, cols_per_lane, rows_per_tile,
                                                steps_per_graph), "set_tiling")
        self.tiling = (int(cols_per_lane), int(rows_per_tile), int(steps_per_graph))

    tiling = (1, 0, 16)  # the library's default shape

    def set_step_mode(self, mode: int = 1) -> None:
        """1 = fused single-pass step on ping-pong buffers (default), 0 = two-pass in
        place, 2 = fused with the interior tiles beside the strip kernels, 3 = pairs
        of steps with the interior advanced two steps per pass (FP32; runs that
        cannot take it use mode 2).  All give identical results."""
        self.check(self.lib.fomo_ctx_set_step_mode(self.ctx, int(mode)), "set_step_mode")

    def set_split(self, edge_rows: int = 16, interior_blocks_per_cu: int = 0) -> None:
        """Split-step shape (step mode 2): rows per edge tile, interior blocks per CU cap."""
        self.check(self.lib.fomo_ctx_set_split(self.ctx, int(edge_rows), int(interior_blocks_per_cu)),
                   "set_split")

    def stream_handle(self) -> int:
        return int(self.lib.fomo_ctx_stream(self.ctx) or 0)

    def __repr__(self):
        return f"MI355XBackend(device={self.device})"


_BACKENDS: dict = {}


def backend(name="mi355x", device: int = 0) -> MI355XBackend:
    """``backend(:mi355x)`` (backends/backend.jl:39-53 plus the new branch).

    ``:cpu`` and ``:cuda`` are the reference's own backends and are not
    provided by this package."""
    key = str(name).lower().lstrip(":")
    if key in ("mi355x", "amd", "hip", "rocm", "gpu"):
        if device not in _BACKENDS:
            _BACKENDS[device] = MI355XBackend(device)
        return _BACKENDS[device]
    raise ValueError(f"Unknown backend: {name}. This package provides :mi355x only "
                     "(:cpu is the reference's own Julia path)")


def to_device(data, b: MI355XBackend):
    """``to_device`` (backends/backend.jl:68-71).  Arrays stay host-side numpy
    here; they are uploaded when a device model is built for a run."""
    return np.asarray(data)


def synchronize(b: MI355XBackend) -> None:
    b.synchronize()


# ---------------------------------------------------------------- structures
@dataclass(eq=False)
class Medium:
    """``Medium{T}`` (core/Structures.jl:66-83)."""
    nx: int
    nz: int
    dx: np.float32
    dz: np.float32
    x_max: np.float32
    z_max: np.float32
    M: int
    pad: int
    is_free_surface: bool
    rho_vx: np.ndarray
    rho_vz: np.ndarray
    lam: np.ndarray
    mu_txx: np.ndarray
    mu_txz: np.ndarray
    # device storage of coefficients and wavefields: "f32", "bf16", "f16",
    # "f16c" (FP16 coefficients, FP32 fields) or "f24" (FP32 coefficients,
    # fields rounded to a 16-bit significand); FP32 arithmetic in every case.
    # Not a field of the reference's Medium
    storage: str = "f32"


@dataclass(eq=False)
class HABCConfig:
    """``HABCConfig{T}`` (core/Structures.jl:94-105).  Weights are the (nz, nx)
    matrices of init_habc; ``None`` lets the library build the same ramps."""
    nbc: int
    qx: np.float32
    qz: np.float32
    qt_x: np.float32
    qt_z: np.float32
    qxt: np.float32
    w_vx: Optional[np.ndarray] = None
    w_vz: Optional[np.ndarray] = None
    w_tau: Optional[np.ndarray] = None


class SimParams:
    """``SimParams(dt, nt, dx, dz, fd_order)`` (core/Structures.jl:143-155):
    dt, dtx = Float32(dt/dx), dtz = Float32(dt/dz) in the arguments' own
    precision, M = fd_order / 2."""

    def __init__(self, dt, nt, dx, dz, fd_order):
        dt_, dx_, dz_ = _jl(dt), _jl(dx), _jl(dz)
        self.dt = F32(dt_)
        self.nt = int(nt)
        self.dtx = F32(dt_ / dx_)
        self.dtz = F32(dt_ / dz_)
        self.fd_order = int(fd_order)
        self.M = self.fd_order // 2

    def __repr__(self):
        return (f"SimParams(dt={self.dt}, nt={self.nt}, dtx={self.dtx}, dtz={self.dtz}, "
                f"fd_order={self.fd_order})")


@dataclass(eq=False)
class Source:
    """``Source{V,I}`` (core/Structures.jl:116-120) plus the source type of
    MultiShotConfig (shot_manager.jl:102): "pressure", "force_z", "force_x"."""
    i: int
    j: int
    wavelet: np.ndarray
    type: str = "pressure"


@dataclass(eq=False)
class Receivers:
    """``Receivers{T,I}`` (core/Structures.jl:127-132); ``type`` may also be a
    sequence with one type per receiver ("vz", "vx", "p")."""
    i: np.ndarray
    j: np.ndarray
    data: np.ndarray
    type: object = "vz"

    def type_codes(self) -> np.ndarray:
        n = len(self.i)
        if isinstance(self.type, str):
            return np.full(n, _lib.REC_TYPES[self.type.lstrip(":")], dtype=np.int32)
        codes = np.array([_lib.REC_TYPES[str(t).lstrip(":")] for t in self.type], dtype=np.int32)
        if codes.shape != (n,):
            raise ValueError("one receiver type per receiver")
        return codes


@dataclass
class TimeStepInfo:
    """``TimeStepInfo`` (simulation/time_stepper.jl:17-21)."""
    k: int
    t: np.float32
    nt: int


@dataclass(eq=False)
class ShotConfig:
    source_x: np.float32
    source_z: np.float32
    shot_id: int


class MultiShotConfig:
    """``MultiShotConfig(x_positions, z_positions, wavelet; source_type)``
    (shot_manager.jl:99-119): shot i at (Float32(x[i]), Float32(z[i])), id i."""

    def __init__(self, x_positions, z_positions, wavelet, source_type="pressure"):
        if len(x_positions) != len(z_positions):
            raise ValueError("x and z must have same length")
        self.shots = [ShotConfig(F32(x), F32(z), i + 1)
                      for i, (x, z) in enumerate(zip(x_positions, z_positions))]
        self.wavelet = np.asarray(wavelet, dtype=F32)
        self.source_type = str(source_type).lstrip(":")


@dataclass(eq=False)
class ShotResult:
    """``ShotResult`` (shot_manager.jl:17-28): gather is [nt x n_rec]."""
    gather: np.ndarray
    shot_id: int
    src_i: int
    src_j: int
    rec_i: list
    rec_j: list


# ---------------------------------------------------------------- setup (init.jl)
def init_medium(vp, vs, rho, dx, dz, nbc: int, fd_order: int, b=None,
                free_surface: bool = True, storage: str = "f32") -> Medium:
    """``init_medium`` (utils/init.jl:68-99).  ``storage`` selects how the
    device keeps the coefficients and wavefields ("f32", "bf16", "f16", "f16c",
    "f24")."""
    if fd_order < 2 or fd_order % 2:
        raise ValueError(f"Unsupported FD order: {fd_order}")
    if storage not in _lib.STORAGE:
        raise ValueError(f"storage must be one of {sorted(_lib.STORAGE)}")
    M = fd_order // 2
    pad = nbc + M
    vp = np.asarray(vp)
    nx_in, nz_in = vp.shape
    nx, nz = nx_in + 2 * pad, nz_in + 2 * pad
    x_max = F32((nx_in - 1) * _jl(dx))
    z_max = F32((nz_in - 1) * _jl(dz))
    lam, mu, mu_txz, rho_vx, rho_vz = _compute_staggered_params(
        _pad_array(vp, pad), _pad_array(np.asarray(vs), pad), _pad_array(np.asarray(rho), pad))
    return Medium(nx, nz, F32(dx), F32(dz), x_max, z_max, M, pad, bool(free_surface),
                  _host_f32(rho_vx), _host_f32(rho_vz), _host_f32(lam), _host_f32(mu),
                  _host_f32(mu_txz), storage)


def init_medium_device(vp, vs, rho, dx, dz, nbc: int, fd_order: int, b=None,
                       free_surface: bool = True, storage: str = "f32") -> Medium:
    """``init_medium`` with the padding and staggered averages left to the
    device (``fomo_model_create_from_velocity``): the returned Medium has the
    grid scalars but no host coefficient arrays, and carries the unpadded
    (nx, nz) vp, vs, rho; the device model built from it is the one
    ``init_medium``'s arrays give."""
    if fd_order < 2 or fd_order % 2:
        raise ValueError(f"Unsupported FD order: {fd_order}")
    if storage not in _lib.STORAGE:
        raise ValueError(f"storage must be one of {sorted(_lib.STORAGE)}")
    M = fd_order // 2
    pad = nbc + M
    vp, vs, rho = (_host_f32(x) for x in (vp, vs, rho))
    if not (vp.shape == vs.shape == rho.shape):
        raise ValueError("vp, vs, rho must have one shape")
    nx_in, nz_in = vp.shape
    x_max = F32((nx_in - 1) * _jl(dx))
    z_max = F32((nz_in - 1) * _jl(dz))
    med = Medium(nx_in + 2 * pad, nz_in + 2 * pad, F32(dx), F32(dz), x_max, z_max, M, pad,
                 bool(free_surface), None, None, None, None, None, storage)
    med._velocity = (vp, vs, rho)
    return med


def init_habc(nx: int, nz: int, nbc: int, dt, dx, dz, v_ref, b=None,
              weights: bool = True) -> HABCConfig:
    """``init_habc`` (utils/init.jl:171-203).  ``weights=False`` skips building
    the three (nz, nx) matrices; the library then builds the same ramps."""
    q = habc_coefficients(v_ref, dt, dx, dz)
    if not weights:
        return HABCConfig(nbc, *q)
    w = [np.asfortranarray(habc_weight_matrix(nx, nz, nbc, off)) for off in (0.0, 0.5, 0.75)]
    return HABCConfig(nbc, *q, *w)


def setup_receivers(x, z, M: Medium, type="vz") -> Receivers:
    """``setup_receivers`` (utils/init.jl:214-227)."""
    i, j = receiver_indices(x, z, M.dx, M.dz, M.pad)
    return Receivers(i.astype(np.int64), j.astype(np.int64), np.zeros((1, len(i)), dtype=F32), type)


# ---------------------------------------------------------------- device objects
class _DeviceModel:
    """fomo_model: Medium + HABCConfig + SimParams + coefficients on the GPU."""

    def __init__(self, b: MI355XBackend, M: Medium, H: HABCConfig, a, p: SimParams):
        lib = b.lib
        self.b = b
        self.M, self.H, self.p = M, H, p
        self.fd_order = 2 * len(a)
        if self.fd_order != p.fd_order:
            raise ValueError("FD coefficient vector does not match SimParams.fd_order")
        self.a = np.ascontiguousarray(np.asarray(a, dtype=F32))
        d = _lib.ModelDesc(M.nx, M.nz, M.pad, H.nbc, p.fd_order, int(M.is_free_surface),
                           float(M.dx), float(M.dz), float(p.dt), float(p.dtx), float(p.dtz),
                           float(H.qx), float(H.qz), float(H.qt_x), float(H.qt_z), float(H.qxt),
                           _lib.STORAGE[getattr(M, "storage", "f32")])
        h = ctypes.c_void_p()
        vel = getattr(M, "_velocity", None)
        if M.rho_vx is None and vel is not None:
            if H.w_vx is not None:
                raise ValueError("a device-side medium uses the init_habc ramps (weights=False)")
            vp, vs, rho = vel
            rc = lib.fomo_model_create_from_velocity(b.ctx, ctypes.byref(d), _f32ptr(self.a),
                                                     _f32ptr(vp), _f32ptr(vs), _f32ptr(rho),
                                                     vp.shape[0], vp.shape[1], ctypes.byref(h))
            b.check(rc, "fomo_model_create_from_velocity")
            self.h = h
            return
        med = [_host_f32(x) for x in (M.rho_vx, M.rho_vz, M.lam, M.mu_txx, M.mu_txz)]
        for x in med:
            if x.shape != (M.nx, M.nz):
                raise ValueError("medium arrays must be (nx, nz)")
        if H.w_vx is not None:
            ws = [np.asfortranarray(np.asarray(w, dtype=F32)) for w in (H.w_vx, H.w_vz, H.w_tau)]
            wp = [_f32ptr(w) for w in ws]
        else:
            wp = [None, None, None]
        rc = lib.fomo_model_create(b.ctx, ctypes.byref(d), _f32ptr(self.a), *[_f32ptr(x) for x in med],
                                   *wp, ctypes.byref(h))
        b.check(rc, "fomo_model_create")
        self.h = h

    def close(self) -> None:
        """Frees the device model now (refused while a pool member uses it)."""
        if getattr(self, "h", None):
            _lib.check(_lib.load().fomo_model_destroy(self.h), self.b._ctx, "fomo_model_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _model(b: MI355XBackend, M: Medium, H: HABCConfig, a, p: SimParams) -> _DeviceModel:
    key = (id(M), id(H), np.asarray(a, dtype=F32).tobytes(), float(p.dt), float(p.dtx),
           float(p.dtz), p.fd_order, getattr(M, "storage", "f32"))
    dm = b._models.get(key)
    if dm is None or dm.M is not M or dm.H is not H:
        dm = _DeviceModel(b, M, H, a, p)
        b._models[key] = dm
    return dm


class _DeviceShot:
    def __init__(self, dm: _DeviceModel, src: Source, rec: Receivers, nt: int):
        lib = dm.b.lib
        self.dm = dm
        self.nt = int(nt)
        wav = np.ascontiguousarray(np.asarray(src.wavelet, dtype=F32))
        ri = np.ascontiguousarray(np.asarray(rec.i, dtype=np.int32))
        rj = np.ascontiguousarray(np.asarray(rec.j, dtype=np.int32))
        rt = np.ascontiguousarray(rec.type_codes())
        self.n_rec = len(ri)
        stype = _lib.SRC_TYPES[str(src.type).lstrip(":")]
        h = ctypes.c_void_p()
        rc = lib.fomo_shot_create(dm.h, int(src.i), int(src.j), stype, _f32ptr(wav), len(wav),
                                  self.nt, _i32ptr(ri), _i32ptr(rj), _i32ptr(rt), self.n_rec,
                                  ctypes.byref(h))
        dm.b.check(rc, "fomo_shot_create")
        self.h = h

    def field_scales(self) -> np.ndarray:
        """FP16 store scales of vx vz txx tzz txz (1 for FP32/BF16 storage)."""
        out = np.zeros(5, dtype=F32)
        self.dm.b.check(self.dm.b.lib.fomo_shot_field_scales(self.h, _f32ptr(out)), "field_scales")
        return out

    def gather(self) -> np.ndarray:
        out = np.zeros((self.nt, self.n_rec), dtype=F32, order="F")
        if self.n_rec:
            self.dm.b.check(self.dm.b.lib.fomo_shot_gather(self.h, _f32ptr(out)), "fomo_shot_gather")
        return out

    def close(self) -> None:
        """Frees the device shot now (refused while it is in a pool)."""
        if getattr(self, "h", None):
            _lib.check(_lib.load().fomo_shot_destroy(self.h), self.dm.b._ctx, "fomo_shot_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Wavefield:
    """``Wavefield(nx, nz, backend)`` (core/Structures.jl:18-55).

    Device storage is allocated when the wavefield is first used with a model;
    field attributes (``W.vx`` ...) download the current values as (nx, nz)
    Float32 arrays."""

    def __init__(self, nx: int, nz: int, b: MI355XBackend):
        self.nx, self.nz, self.b = int(nx), int(nz), b
        self._h = None
        self._dm = None
        self._bound = None  # (M, H, a, p) of the last run, for the per-stage API

    def _bind(self, dm: _DeviceModel):
        if (dm.M.nx, dm.M.nz) != (self.nx, self.nz):
            raise ValueError("wavefield and medium sizes differ")
        if self._dm is not dm:
            old = self._dump() if self._h is not None else None
            self._free()
            h = ctypes.c_void_p()
            self.b.check(self.b.lib.fomo_wavefield_create(dm.h, ctypes.byref(h)),
                         "fomo_wavefield_create")
            self._h, self._dm = h, dm
            if old is not None:
                for c, arr in enumerate(old):
                    self._upload(c, arr)
        return self._h

    def _dump(self):
        n = 10 if getattr(self._dm.M, "storage", "f32") == "f32" else 5
        return [self.download(c) for c in range(n)]

    def _free(self):
        if self._h is not None:
            _lib.check(_lib.load().fomo_wavefield_destroy(self._h), self.b._ctx,
                       "fomo_wavefield_destroy")
            self._h = None
            self._dm = None

    def close(self) -> None:
        """Frees the device state now (refused while the wavefield is in a pool)."""
        self._free()

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass

    def download(self, f) -> np.ndarray:
        c = _lib.FIELDS[f] if isinstance(f, str) else int(f)
        out = np.zeros((self.nx, self.nz), dtype=F32, order="F")
        if self._h is None:
            return out
        self.b.check(self.b.lib.fomo_wavefield_download(self._h, c, _f32ptr(out)), "download")
        return out

    def _upload(self, c: int, arr) -> None:
        a = _host_f32(arr)
        self.b.check(self.b.lib.fomo_wavefield_upload(self._h, c, _f32ptr(a)), "upload")

    def upload(self, f, arr) -> None:
        if self._h is None:
            raise RuntimeError("wavefield not bound to a model yet (run a step or bind first)")
        self._upload(_lib.FIELDS[f] if isinstance(f, str) else int(f), arr)

    def __getattr__(self, name):
        if name in _lib.FIELDS:
            return self.download(name)
        raise AttributeError(name)


def bind(b: MI355XBackend, W: Wavefield, M: Medium, H: HABCConfig, a, p: SimParams) -> _DeviceModel:
    """Associates W with the device model of (M, H, a, p); later per-stage calls on W use it."""
    dm = _model(b, M, H, a, p)
    W._bind(dm)
    W._bound = dm
    return dm


def _shot_for(dm: _DeviceModel, src: Source, rec: Receivers, nt: int) -> _DeviceShot:
    cache = getattr(rec, "_fomo_shots", None)
    if cache is None:
        cache = {}
        rec._fomo_shots = cache
    key = (id(dm), id(src), int(nt))
    ds = cache.get(key)
    if ds is None or ds.dm is not dm:
        cache.clear()
        ds = _DeviceShot(dm, src, rec, nt)
        ds._src = src
        cache[key] = ds
    return ds


def _bound(W: Wavefield) -> _DeviceModel:
    if W._bound is None:
        raise RuntimeError("no model bound to this wavefield: call bind(...) or time_step(...) first")
    return W._bound


# ---------------------------------------------------------------- kernels (per stage)
def reset(b: MI355XBackend, W: Wavefield) -> None:
    """``reset!`` (kernels/source_receiver.jl:115-129)."""
    if W._h is not None:
        b.check(b.lib.fomo_wavefield_reset(W._h), "reset")


def backup_boundary(b, W, H=None, M=None):
    dm = _bound(W)
    b.check(b.lib.fomo_backup_boundary(dm.h, W._h), "backup_boundary")


def inject_source(b, W, src: Source, k: int, dt=None, rec: Receivers | None = None):
    dm = _bound(W)
    rec = rec if rec is not None else Receivers(np.zeros(0, np.int64), np.zeros(0, np.int64),
                                                np.zeros((1, 0), F32))
    ds = _shot_for(dm, src, rec, max(int(k), len(src.wavelet), 1))
    b.check(b.lib.fomo_inject_source(dm.h, W._h, ds.h, int(k)), "inject_source")


def update_velocity(b, W, M=None, a=None, p=None):
    dm = _bound(W)
    b.check(b.lib.fomo_update_velocity(dm.h, W._h), "update_velocity")


def apply_habc_velocity(b, W, H=None, M=None):
    dm = _bound(W)
    b.check(b.lib.fomo_apply_habc_velocity(dm.h, W._h), "apply_habc_velocity")


def update_stress(b, W, M=None, a=None, p=None):
    dm = _bound(W)
    b.check(b.lib.fomo_update_stress(dm.h, W._h), "update_stress")


def apply_habc_stress(b, W, H=None, M=None):
    dm = _bound(W)
    b.check(b.lib.fomo_apply_habc_stress(dm.h, W._h), "apply_habc_stress")


def apply_free_surface(b, W, M=None):
    dm = _bound(W)
    b.check(b.lib.fomo_apply_free_surface(dm.h, W._h), "apply_free_surface")


def record_receivers(b, W, rec: Receivers, k: int, src: Source | None = None, nt: int | None = None):
    """``record_receivers!``: samples into the device gather of (src, rec) and
    refreshes ``rec.data`` ([nt x n_rec])."""
    dm = _bound(W)
    src = src if src is not None else Source(1, 1, np.zeros(0, F32))
    nt = int(nt if nt is not None else max(rec.data.shape[0], k))
    ds = _shot_for(dm, src, rec, nt)
    b.check(b.lib.fomo_record_receivers(dm.h, W._h, ds.h, int(k)), "record_receivers")
    rec.data = ds.gather()


# ---------------------------------------------------------------- time stepping
def time_step(b, W, M, H, a, src: Source, rec: Receivers, k: int, p: SimParams) -> None:
    """``time_step!`` (simulation/time_stepper.jl:39-63), one fused step."""
    dm = bind(b, W, M, H, a, p)
    ds = _shot_for(dm, src, rec, p.nt)
    b.check(b.lib.fomo_time_step(dm.h, W._h, ds.h, int(k)), "time_step")


def run_time_loop(b, W, M, H, a, src: Source, rec: Receivers, p: SimParams,
                  progress: bool = False, on_step: Optional[Callable] = None,
                  on_step_every: int = 1) -> None:
    """``run_time_loop!`` (simulation/time_stepper.jl:97-128).

    Without a callback the nt steps run as replayed HIP graphs.  With
    ``on_step(W, TimeStepInfo)`` the loop breaks into segments of
    ``on_step_every`` steps (default 1, the reference's per-step callback);
    returning False stops the simulation early.  A ``video.MultiFieldRecorder``
    as ``on_step`` keeps the graph replay: its frames are taken on the device
    after every ``skip``-th step.  ``rec.data`` receives the [nt x n_rec]
    gather."""
    from .video import MultiFieldRecorder
    dm = bind(b, W, M, H, a, p)
    ds = _shot_for(dm, src, rec, p.nt)
    lib = b.lib
    nt = p.nt
    if isinstance(on_step, MultiFieldRecorder):
        on_step.run_loop(b, dm, W, ds, nt, p.dt)
    elif on_step is None:
        b.check(lib.fomo_run_time_loop(dm.h, W._h, ds.h, 1, nt), "run_time_loop")
    else:
        every = max(1, int(on_step_every))
        k = 1
        while k <= nt:
            n = min(every, nt - k + 1)
            b.check(lib.fomo_run_time_loop(dm.h, W._h, ds.h, k, n), "run_time_loop")
            k += n
            kk = k - 1
            b.synchronize()
            if on_step(W, TimeStepInfo(kk, F32(kk * p.dt), nt)) is False:
                break
    b.synchronize()
    rec.data = ds.gather()


def run_shot(b, W, M, H, a, src: Source, rec: Receivers, p: SimParams, shot_id: int = 1,
             progress: bool = False, on_step=None) -> ShotResult:
    """``run_shot!`` (simulation/shot_manager.jl:46-68)."""
    dm = bind(b, W, M, H, a, p)
    reset(b, W)
    ds = _shot_for(dm, src, rec, p.nt)
    b.check(b.lib.fomo_shot_reset(ds.h), "shot_reset")
    run_time_loop(b, W, M, H, a, src, rec, p, progress=progress, on_step=on_step)
    return ShotResult(np.array(rec.data, copy=True), shot_id, int(src.i), int(src.j),
                      [int(v) for v in rec.i], [int(v) for v in rec.j])


def _create_source(M: Medium, shot: ShotConfig, wavelet, source_type: str) -> Source:
    """``_create_source`` (shot_manager.jl:196-206) with ``_coord_to_index``."""
    i = coord_to_index(shot.source_x, M.dx, M.pad)
    j = coord_to_index(shot.source_z, M.dz, M.pad)
    return Source(i, j, wavelet, source_type)


def run_shots(b, W, M, H, a, rec_template: Receivers, shot_config: MultiShotConfig, p: SimParams,
              on_shot_complete=None, on_step=None, progress: bool = False,
              batch: Optional[int] = None, tune: bool = False) -> list:
    """``run_shots!`` (simulation/shot_manager.jl:145-180).

    Results come back in shot order and ``on_shot_complete`` is called in shot
    order, as in the reference.  Unlike the reference's sequential loop, shots
    run ``batch`` at a time in one launch per kernel (a device work list over
    all their tiles), each on its own device wavefield; ``batch=1`` (or an
    ``on_step`` callback) gives the reference's one-shot-at-a-time order on W.
    Default batch: as many shots as fit in a quarter of the free GPU memory,
    at most 64.  ``tune=True`` chooses the tile shape per batch by timing
    candidates on the batch itself (``fomo_julia_amd.tune``); the choice is
    in ``run_shots.last_tuning``."""
    shots = list(shot_config.shots)
    n = len(shots)
    results = [None] * n
    if on_step is not None or batch == 1 or n == 1:
        for idx, shot in enumerate(shots):
            src = _create_source(M, shot, shot_config.wavelet, shot_config.source_type)
            rec = Receivers(np.array(rec_template.i), np.array(rec_template.j),
                            np.zeros((p.nt, len(rec_template.i)), F32), rec_template.type)
            results[idx] = run_shot(b, W, M, H, a, src, rec, p, shot_id=shot.shot_id,
                                    progress=progress, on_step=on_step)
            if on_shot_complete is not None:
                on_shot_complete(results[idx])
        return results
    dm = bind(b, W, M, H, a, p)
    if batch is None:
        per_shot = int(b.lib.fomo_wavefield_device_bytes(W._h)) + 4 * p.nt * len(rec_template.i)
        free = _free_device_bytes()
        batch = max(1, min(64, n, int(free // 4 // max(per_shot, 1)) if free else 8))
    lib = b.lib
    done = 0
    while done < n:
        group = shots[done:done + batch]
        wfs, dss, srcs = [], [], []
        for shot in group:
            src = _create_source(M, shot, shot_config.wavelet, shot_config.source_type)
            rec = Receivers(np.array(rec_template.i), np.array(rec_template.j),
                            np.zeros((p.nt, len(rec_template.i)), F32), rec_template.type)
            w = Wavefield(M.nx, M.nz, b)
            w._bind(dm)
            wfs.append(w)
            srcs.append((src, rec))
            dss.append(_DeviceShot(dm, src, rec, p.nt))
        k = len(group)
        if tune:
            from . import tune as T
            run_shots.last_tuning = T.autotune_batch(b, [dm] * k, wfs, dss)
        arr = lambda hs: (ctypes.c_void_p * k)(*[h.value for h in hs])
        bh = ctypes.c_void_p()
        rc = lib.fomo_batch_create(b.ctx, k, arr([dm.h] * k), arr([w._h for w in wfs]),
                                   arr([d.h for d in dss]), ctypes.byref(bh))
        b.check(rc, "fomo_batch_create")
        try:
            b.check(lib.fomo_batch_run(bh, 1, p.nt), "fomo_batch_run")
            b.synchronize()
        finally:
            lib.fomo_batch_destroy(bh)
        for t, shot in enumerate(group):
            src, rec = srcs[t]
            rec.data = dss[t].gather()
            res = ShotResult(rec.data, shot.shot_id, int(src.i), int(src.j),
                             [int(v) for v in rec.i], [int(v) for v in rec.j])
            results[done + t] = res
            if on_shot_complete is not None:
                on_shot_complete(res)
        done += k
    return results


run_shots.last_tuning = None


def _free_device_bytes() -> int:
    try:
        import torch
        if torch.cuda.is_available():
            free, _ = torch.cuda.mem_get_info()
            return int(free)
    except Exception:
        pass
    return 0
