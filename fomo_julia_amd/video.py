"""Wavefield frames for the on_step video callback (src/visualization/video_recorder.jl).

``MultiFieldRecorder(nx, nz, dt, VideoConfig(...))`` is the reference's
callback object: called as ``on_step(W, info)`` it takes, on every step with
``info.k % skip == 0``, one decimated frame per field (p, vel, vx, vz) of the
state after step k.  Here the frames are produced on the GPU by the library's
frame store (``fomo_frames``): called directly the recorder extracts the
current state of W (one small kernel and a copy of the decimated frame);
handed to ``run_time_loop``/``run_shot``/``run_shots`` as ``on_step`` it is
recognised and the whole loop still runs as replayed HIP graphs, with the
frame kernel captured after every step writing into a ring of device frames
that the host drains once per ring.

Rendering (CairoMakie heatmaps, MP4 encoding) is plotting and not part of
this package: ``close()`` writes each field's frames as raw Float32
``<prefix>_<field>.f32`` ([frames][nz_d][nx_d], x fastest, as Julia holds the
(nx_d, nz_d) frame buffer) with a JSON sidecar holding the steps, times,
titles, the colour range the reference's auto-scaling picks and the video
frame rate.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass, field as dfield
from typing import Optional

import numpy as np

from . import _lib

F32 = np.float32
FIELDS = {"p": 0, "vel": 1, "vx": 2, "vz": 3}
_NAMES = {"p": "Pressure (Txx+Tzz)/2", "vel": "Velocity Magnitude",
          "vx": "Horizontal Velocity Vx", "vz": "Vertical Velocity Vz"}
_LABELS = {"p": "Pressure", "vel": "|V|", "vx": "Vx", "vz": "Vz"}
# device memory the frame ring of one run may take
RING_BYTES = 2 << 30


def _sym(x) -> str:
    return str(x).lstrip(":")


@dataclass
class VideoConfig:
    """``VideoConfig`` (video_recorder.jl:32-81), same keywords and defaults."""
    fields: list = dfield(default_factory=lambda: ["p"])
    output_dir: str = "."
    prefix: str = "wavefield"
    time_scale: float = 0.01
    skip: int = 1
    downsample: int = 1
    clim: Optional[tuple] = None
    colormap: str = "balance"

    def __post_init__(self):
        self.fields = [_sym(f) for f in self.fields]
        for f in self.fields:
            if f not in FIELDS:
                raise ValueError(f"Unknown field: {f}. Use :p, :vel, :vx, or :vz")
        if int(self.skip) < 1 or int(self.downsample) < 1:
            raise ValueError("skip and downsample must be >= 1")
        self.skip, self.downsample = int(self.skip), int(self.downsample)
        self.time_scale = F32(self.time_scale)
        if self.clim is not None:
            self.clim = (F32(self.clim[0]), F32(self.clim[1]))


def field_title(field: str, t, k: int, nt: int) -> str:
    """``_field_title`` with ``_field_name`` (video_recorder.jl:239-250)."""
    return "%s | Time: %.4f s | Step: %d/%d" % (_NAMES.get(field, field), float(t), k, nt)


def field_label(field: str) -> str:
    """``_field_label`` (video_recorder.jl:252-258)."""
    return _LABELS.get(field, field)


def video_fps(dt, skip: int, time_scale) -> int:
    """Frames per second of the video (video_recorder.jl:178-183):
    round(Int, 1.0 / (dt * skip * time_scale)) clamped to 1..120; the
    product is Float32, the quotient Float64, ties round to even."""
    prod = F32(F32(F32(dt) * F32(skip)) * F32(time_scale))
    return int(min(max(round(1.0 / float(prod)), 1), 120))


class _FieldFrames:
    """One field's frames and the state ``_SingleRecorder`` (video_recorder.jl:150-228)
    keeps besides its plot."""

    def __init__(self, field, cfg: VideoConfig, dt, nx_d, nz_d):
        self.field = field
        self.filename = os.path.join(cfg.output_dir, f"{cfg.prefix}_{field}.f32")
        self.fps = video_fps(dt, cfg.skip, cfg.time_scale)
        self.shape = (nx_d, nz_d)
        self.clim = cfg.clim
        self.colorrange = cfg.clim if cfg.clim is not None else (F32(-0.01), F32(0.01))
        self.auto_clim_done = False
        self.frames: list = []   # (nx_d, nz_d) Float32 arrays (Fortran order)
        self.steps: list = []
        self.titles: list = []

    def add(self, frame: np.ndarray, k: int, t, nt: int) -> None:
        # auto colour limits after five frames (video_recorder.jl:214-221)
        if self.clim is None and not self.auto_clim_done and len(self.frames) >= 5:
            rms = float(np.sqrt(np.sum(frame.astype(np.float64) ** 2) / frame.size))
            if rms > 1e-10:
                self.colorrange = (F32(-3.0 * rms), F32(3.0 * rms))
                self.auto_clim_done = True
        self.frames.append(frame)
        self.steps.append(int(k))
        self.titles.append(field_title(self.field, t, k, nt))


class MultiFieldRecorder:
    """``MultiFieldRecorder(nx, nz, dt, config)`` (video_recorder.jl:102-148)."""

    def __init__(self, nx: int, nz: int, dt, config: Optional[VideoConfig] = None):
        self.config = config if config is not None else VideoConfig()
        self.nx, self.nz = int(nx), int(nz)
        self.dt = F32(dt)
        ds = self.config.downsample
        self.nx_d, self.nz_d = -(-self.nx // ds), -(-self.nz // ds)
        self.recorders = {f: _FieldFrames(f, self.config, self.dt, self.nx_d, self.nz_d)
                          for f in self.config.fields}
        self._stores: dict = {}  # (id(backend), capacity) -> _FrameStore
        self.device_frames = 0   # frames taken inside graph-replayed loops

    # -- the reference's callback: the current state of W
    def __call__(self, W, info) -> bool:
        if info.k % self.config.skip != 0:
            return True
        b = W.b
        st = self._store(b, 1)
        b.check(b.lib.fomo_wavefield_frame(W._h, st.h, 0), "fomo_wavefield_frame")
        self._deliver(st.download(b, 0, 1)[0], info.k, info.t, info.nt)
        return True

    # -- run_time_loop's graph path
    def run_loop(self, b, dm, W, ds, nt: int, dt) -> None:
        """Steps 1..nt of (dm, W, ds) with every frame taken on the device."""
        skip = self.config.skip
        n_frames = nt // skip
        if n_frames == 0:
            b.check(b.lib.fomo_run_time_loop(dm.h, W._h, ds.h, 1, nt), "run_time_loop")
            return
        per = len(self.config.fields) * self.nx_d * self.nz_d * 4
        cap = max(1, min(n_frames, RING_BYTES // per))
        st = self._store(b, cap)
        k = 1
        while k <= nt:
            n = min(cap * skip, nt - k + 1)
            b.check(b.lib.fomo_run_time_loop_frames(dm.h, W._h, ds.h, k, n, st.h),
                    "fomo_run_time_loop_frames")
            first, last = -(-k // skip), (k + n - 1) // skip  # frame numbers (1-based) in the segment
            if last >= first:
                slots = [(f - 1) % cap for f in range(first, last + 1)]
                got = st.download(b, 0, cap)
                for f, s in zip(range(first, last + 1), slots):
                    kk = f * skip
                    self._deliver(got[s], kk, F32(F32(kk) * F32(dt)), nt)
                    self.device_frames += 1
            k += n

    def _store(self, b, capacity: int):
        key = (id(b), capacity)
        st = self._stores.get(key)
        if st is None:
            # one ring per backend: drop the others
            for k2 in [k2 for k2 in self._stores if k2[0] == id(b)]:
                del self._stores[k2]
            st = _FrameStore(b, self.nx, self.nz, [FIELDS[f] for f in self.config.fields],
                             self.config.skip, self.config.downsample, capacity)
            self._stores[key] = st
        return st

    def _deliver(self, frames_of_slot: np.ndarray, k: int, t, nt: int) -> None:
        # [n_fields][nz_d][nx_d] -> per field (nx_d, nz_d) as Julia's buffer
        for q, f in enumerate(self.config.fields):
            self.recorders[f].add(np.asfortranarray(frames_of_slot[q].T), k, t, nt)

    def frames(self, field: str) -> np.ndarray:
        """[n_frames, nx_d, nz_d] of one field."""
        r = self.recorders[_sym(field)]
        if not r.frames:
            return np.zeros((0, self.nx_d, self.nz_d), F32)
        return np.stack(r.frames)

    def close(self) -> None:
        """``close(rec)`` (video_recorder.jl:141-148): writes every field's frames."""
        os.makedirs(self.config.output_dir or ".", exist_ok=True)
        for f, r in self.recorders.items():
            data = self.frames(f)
            # [frames][nz_d][nx_d], x fastest
            np.ascontiguousarray(data.transpose(0, 2, 1)).astype("<f4").tofile(r.filename)
            meta = {"field": f, "label": field_label(f), "nx_d": self.nx_d, "nz_d": self.nz_d,
                    "frames": len(r.frames), "steps": r.steps, "titles": r.titles,
                    "fps": r.fps, "colorrange": [float(r.colorrange[0]), float(r.colorrange[1])],
                    "colormap": self.config.colormap, "layout": "[frame][nz_d][nx_d] float32 LE"}
            with open(r.filename[:-4] + ".json", "w") as fh:
                json.dump(meta, fh, indent=1)
        self._stores.clear()


def VideoRecorder(filename: str, nx: int, nz: int, field="p", dt=0.001, time_scale=0.01,
                  skip: int = 1, downsample: int = 1, clim=None, colormap="balance"):
    """``VideoRecorder`` (video_recorder.jl:300-333): a one-field MultiFieldRecorder."""
    base = os.path.basename(filename)
    cfg = VideoConfig(fields=[field], output_dir=os.path.dirname(filename),
                      prefix=os.path.splitext(base)[0], time_scale=time_scale, skip=skip,
                      downsample=downsample, clim=clim, colormap=colormap)
    return MultiFieldRecorder(nx, nz, dt, cfg)


class _FrameStore:
    """fomo_frames: a device ring of `capacity` frames."""

    def __init__(self, b, nx, nz, fields, skip, ds, capacity):
        self.lib = b.lib
        arr = (ctypes.c_int32 * len(fields))(*fields)
        h = ctypes.c_void_p()
        b.check(b.lib.fomo_frames_create(b.ctx, nx, nz, arr, len(fields), skip, ds, capacity,
                                         ctypes.byref(h)), "fomo_frames_create")
        self.h = h
        nxd, nzd = ctypes.c_int32(), ctypes.c_int32()
        b.check(b.lib.fomo_frames_shape(h, ctypes.byref(nxd), ctypes.byref(nzd)), "fomo_frames_shape")
        self.shape = (len(fields), nzd.value, nxd.value)
        self.capacity = capacity

    def download(self, b, first: int, count: int) -> np.ndarray:
        out = np.empty((count,) + self.shape, F32)
        b.check(self.lib.fomo_frames_download(self.h, first, count,
                                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                "fomo_frames_download")
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().fomo_frames_destroy(self.h)
        except Exception:
            pass
