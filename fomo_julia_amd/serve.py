"""Serving loop: a stream of shot requests on one GPU with continuous batching.

The reference runs one shot at a time (``run_shots!``,
simulation/shot_manager.jl:145-180); this module adds a request loop on top of
the same per-shot objects (medium, HABC, source, receivers, gather):

* requests arrive over time (any grid size, length, FD order and storage);
* each is prepared on the host in a worker thread (medium, ``init_medium``,
  HABC constants, wavelet, receivers) while the GPU keeps running;
* it is admitted when its device memory (model + wavefield + gather) fits in
  the HBM budget, first come first served among those that fit;
* admitted shots of the same FD order and storage run together in one
  ``fomo_pool`` (one launch per kernel for all of them, one captured graph
  per pool); shots join and leave the pool between chunks of steps while
  the others keep their wavefields resident;
* a finished shot's gather is copied back and its memory freed.

Every shot's arithmetic is the same as alone (the kernels treat each shot
of a pool independently), so a result is identical alone and in the mix.

Threads: worker threads prepare requests and create their device objects;
the thread that calls ``Server.serve`` owns the pools.  Every pool call, and
the freeing of a finished shot's device objects, happens on that thread
(``Pool`` refuses calls from any other).  A request whose preparation or
device objects fail is reported in ``Server.failed`` and the loop goes on.
"""
from __future__ import annotations

import ctypes
import math
import threading
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor, wait
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from . import elastic2d as E
from . import init as I
from . import media

F32 = np.float32


@dataclass
class ShotRequest:
    """One shot to serve.  The medium is generated from `seed` (layered media
    with Gaussian lenses and a water layer); the source is a Ricker pressure
    source at the surface centre, the receivers every `rec_every` columns on
    the surface recording vz."""
    rid: int
    arrival: float
    nx: int
    nz: int
    nt: int
    fd_order: int
    storage: str = "f32"
    seed: int = 0
    dx: float = 10.0
    f0: float = 15.0
    nbc: int = 40
    rec_every: int = 2
    cfl: float = 0.4


@dataclass
class Prepared:
    req: ShotRequest
    M: E.Medium
    H: E.HABCConfig
    p: E.SimParams
    a: np.ndarray
    src: E.Source
    rec: E.Receivers
    bytes: int
    prep_s: float


@dataclass
class Failed:
    rid: int
    arrival: float
    stage: str   # "prepare" or "admit"
    error: str


@dataclass
class Completed:
    rid: int
    arrival: float
    admitted: float
    done: float
    gather: np.ndarray
    storage: str
    fd_order: int
    cell_updates: int

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def c5_stream(n: int, seed: int = 5, rate: Optional[float] = None, storages=("f32", "bf16"),
              orders=(4, 8, 12), nx=(300, 2000), nz=(300, 1200), nt=(1000, 6000)) -> list:
    """The served-stream workload: grids uniform in nx and nz, lengths uniform
    in nt, FD orders from `orders`, storage alternating over `storages`,
    Poisson arrivals at `rate` per second (all at t = 0 when rate is None)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        if rate:
            t += rng.exponential(1.0 / rate)
        out.append(ShotRequest(
            rid=k, arrival=t if rate else 0.0,
            nx=int(rng.integers(nx[0], nx[1] + 1)), nz=int(rng.integers(nz[0], nz[1] + 1)),
            nt=int(rng.integers(nt[0], nt[1] + 1)) // 2 * 2,
            fd_order=int(rng.choice(orders)), storage=storages[k % len(storages)],
            seed=int(rng.integers(1 << 30))))
    return out


def with_arrivals(reqs: list, rate: Optional[float], seed: int = 11) -> list:
    """The same requests with Poisson arrival times at `rate` per second."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for r in reqs:
        if rate:
            t += rng.exponential(1.0 / rate)
        out.append(ShotRequest(**{**r.__dict__, "arrival": t if rate else 0.0}))
    return out


# ---------------------------------------------------------------- device bytes
def _pitch(nx: int) -> int:
    return (nx + 272 + 255) // 256 * 256


def device_bytes(nx_pad: int, nz_pad: int, storage: str, nt: int, n_rec: int) -> int:
    """Device memory of one shot: coefficient blocks, wavefield state and
    gather, with the library's layout (pitched rows, zero guard rows)."""
    stride = (nz_pad + 36) * _pitch(nx_pad) + 512
    ces = 4 if storage in ("f32", "f24") else 2
    # f24: 2 bytes per field cell plus the byte-plane image (1 byte)
    fes = {"bf16": 2, "f16": 2, "f24": 3}.get(storage, 4)
    fields = 15 if storage == "f32" else 10
    # FP32 coefficients also hold rho_vx and mu_txx for the derived-coefficient fused kernel
    coef = 7 if storage in ("f32", "f24") else 5
    return coef * stride * ces + fields * stride * fes + nt * max(n_rec, 1) * 4 + 3 * 4096


def prepare(req: ShotRequest) -> Prepared:
    """Host setup of a request (what the client would hand over): medium,
    init_medium, HABC constants, wavelet, source and receivers."""
    t0 = time.perf_counter()
    dx = F32(req.dx)
    vp, vs, rho = media.layered_with_lenses(req.nx, req.nz, float(dx), float(dx), seed=req.seed,
                                            water_depth=5 * float(dx))
    dt = media.cfl_dt(vp, dx, req.cfl)
    # padding, staggered averages and coefficients are computed on the GPU at admission
    M = E.init_medium_device(vp, vs, rho, dx, dx, req.nbc, req.fd_order, None, free_surface=True,
                             storage=req.storage)
    v_ref = F32(np.mean(vp, dtype=np.float64))  # served requests: numpy's mean
    H = E.init_habc(M.nx, M.nz, req.nbc, dt, dx, dx, v_ref, None, weights=False)
    p = E.SimParams(dt, req.nt, dx, dx, req.fd_order)
    a = I.get_fd_coefficients(req.fd_order)
    wav = I.ricker_wavelet(F32(req.f0), dt, req.nt)
    src = E.Source(I.coord_to_index(F32(M.x_max / F32(2)), M.dx, M.pad),
                   I.coord_to_index(F32(0.0), M.dz, M.pad), wav, "pressure")
    xr = np.arange(0, req.nx, req.rec_every, dtype=F32) * dx
    rec = E.setup_receivers(xr, np.zeros(len(xr), F32), M, type="vz")
    nbytes = device_bytes(M.nx, M.nz, req.storage, req.nt, len(xr))
    return Prepared(req, M, H, p, a, src, rec, nbytes, time.perf_counter() - t0)


# ---------------------------------------------------------------- pools
# Tile shape (columns per lane, rows per tile) of the served pools and of the
# serial baseline, per FD order, fixed so that a pool does not inherit whatever
# shape an earlier autotune left on the context.  On the 40-request stream
# (pool tiles of 12-16 rows), two columns per lane for orders 8 and 12 served
# 0.6 % fewer shots per hour than one column, and two columns for every order
# 1.1 % fewer: the served grids are small, and one column per lane gives more
# tiles to fill the GPU.
SERVE_TILING = {4: (1, 0), 8: (1, 0), 12: (1, 0)}


def _tiling_for(fd_order: int) -> tuple:
    return SERVE_TILING.get(int(fd_order), (1, 0))


class _Tiling:
    """Sets the backend's tile shape for a block and restores it after."""

    def __init__(self, b: E.MI355XBackend, fd_order: int):
        self.b, self.shape = b, _tiling_for(fd_order)

    def __enter__(self):
        self.saved = self.b.tiling
        self.b.set_tiling(self.shape[0], self.shape[1], self.saved[2])

    def __exit__(self, *exc):
        self.b.set_tiling(*self.saved)


class Pool:
    """fomo_pool: continuous batch of one FD order and storage type.  Its
    owner is the thread that creates it; calls from any other thread raise."""

    def __init__(self, b: E.MI355XBackend, fd_order: int, storage: str, capacity: int):
        self.b = b
        self.owner = threading.get_ident()
        self.key = (fd_order, storage)
        h = ctypes.c_void_p()
        with _Tiling(b, fd_order):  # the pool keeps the shape current at its creation
            b.check(b.lib.fomo_pool_create(b.ctx, fd_order, _lib.STORAGE[storage], capacity,
                                           ctypes.byref(h)), "fomo_pool_create")
        self.h = h
        self.capacity = capacity
        self.members: dict = {}  # rid -> _Member

    def _own(self, what: str) -> None:
        if threading.get_ident() != self.owner:
            raise RuntimeError(f"Pool.{what} called from a thread that does not own the pool")

    def add(self, m) -> None:
        self._own("add")
        self.b.check(self.b.lib.fomo_pool_add(self.h, m.dm.h, m.W._h, m.ds.h), "fomo_pool_add")
        self.members[m.rid] = m

    def remove(self, m) -> None:
        self._own("remove")
        self.b.check(self.b.lib.fomo_pool_remove(self.h, m.ds.h), "fomo_pool_remove")
        del self.members[m.rid]

    def run(self, n_steps: int) -> None:
        self._own("run")
        self.b.check(self.b.lib.fomo_pool_run(self.h, n_steps), "fomo_pool_run")

    def synchronize(self) -> None:
        self._own("synchronize")
        self.b.check(self.b.lib.fomo_pool_synchronize(self.h), "fomo_pool_synchronize")

    def next_step(self, m) -> int:
        self._own("next_step")
        k = ctypes.c_int32()
        self.b.check(self.b.lib.fomo_pool_next_step(self.h, m.ds.h, ctypes.byref(k)), "next_step")
        return k.value

    def close(self) -> None:
        """Destroys the pool (its members leave it; their objects stay)."""
        self._own("close")
        if getattr(self, "h", None):
            _lib.load().fomo_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().fomo_pool_destroy(self.h)
        except Exception:
            pass


class _Member:
    """A request's device objects: model, wavefield and shot.  Built on a
    worker thread; freed with close() on the pools' owner thread."""

    def __init__(self, b: E.MI355XBackend, prep: Prepared, admitted: float):
        self.rid = prep.req.rid
        self.prep = prep
        self.admitted = admitted
        self.dm = self.W = self.ds = None
        try:
            self.dm = E._DeviceModel(b, prep.M, prep.H, prep.a, prep.p)
            self.W = E.Wavefield(prep.M.nx, prep.M.nz, b)
            self.W._bind(self.dm)
            self.ds = E._DeviceShot(self.dm, prep.src, prep.rec, prep.p.nt)
        except Exception:
            self.close()
            raise
        self.nt = prep.p.nt

    def close(self) -> None:
        for obj in (self.ds, self.W, self.dm):
            if obj is not None:
                obj.close()
        self.dm = self.W = self.ds = None


class Server:
    """Continuous-batching shot server on one MI355X.

    ``budget_bytes``: device memory for resident shots; ``chunk``: steps per
    pool run between membership changes (even); ``capacity``: shots per pool."""

    def __init__(self, b: Optional[E.MI355XBackend] = None, budget_bytes: int = 64 << 30,
                 chunk: int = 64, capacity: int = 64, workers: int = 4):
        if chunk <= 0 or chunk % 2:
            raise ValueError("chunk must be a positive even number of steps")
        self.b = b or E.backend("mi355x")
        self.budget = int(budget_bytes)
        self.chunk = chunk
        self.capacity = capacity
        self.workers = workers
        self.pools: dict = {}
        self.failed: list = []  # Failed records of requests that could not run
        # seconds spent: preparing requests (worker threads, summed), creating
        # device objects, waiting for the GPU; chunks launched
        self.stats = {"prepare_s": 0.0, "admit_s": 0.0, "gpu_wait_s": 0.0, "chunks": 0,
                      "max_resident": 0}

    def _pool(self, key) -> Pool:
        if key not in self.pools:
            self.pools[key] = Pool(self.b, key[0], key[1], self.capacity)
        return self.pools[key]

    def serve(self, requests: list, realtime: bool = True, on_complete=None) -> list:
        """Serves `requests` (their `arrival` times are seconds from the start;
        with realtime=False every request is available at once).  Returns the
        completed shots in completion order."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        pending = deque(reqs)
        waiting: deque = deque()  # (request, future)
        active: dict = {}         # rid -> (pool, member)
        used = 0
        done: list = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        st = self.stats

        creating: deque = deque()  # (pool, future of _Member): memory reserved, objects being built
        slots: dict = {}           # pool key -> members + members being created

        def make_member(prep, t_admit):
            ta = time.perf_counter()
            m = _Member(self.b, prep, t_admit)
            m.create_s = time.perf_counter() - ta
            return m

        def admit(ex):
            # first come first served among the prepared requests that fit; the
            # device objects are built in a worker thread while the GPU runs
            nonlocal used, waiting
            keep = deque()
            while waiting:
                r, fut = waiting.popleft()
                if fut.done():
                    try:
                        prep = fut.result()
                    except Exception as e:  # this request cannot run; the others go on
                        self.failed.append(Failed(r.rid, r.arrival, "prepare", repr(e)))
                        continue
                    key = (r.fd_order, r.storage)
                    pool = self._pool(key)
                    fits = used + prep.bytes <= self.budget or not (active or creating)
                    if fits and slots.get(key, 0) < pool.capacity:
                        used += prep.bytes
                        slots[key] = slots.get(key, 0) + 1
                        st["prepare_s"] += prep.prep_s
                        creating.append((pool, prep, ex.submit(make_member, prep, now())))
                        continue
                keep.append((r, fut))
            waiting = keep
            # members whose device objects are ready join their pools
            still = deque()
            while creating:
                pool, prep, fut = creating.popleft()
                if fut.done():
                    try:
                        m = fut.result()
                    except Exception as e:  # device objects failed: release the reservation
                        used -= prep.bytes
                        slots[pool.key] -= 1
                        self.failed.append(Failed(prep.req.rid, prep.req.arrival, "admit", repr(e)))
                        continue
                    st["admit_s"] += m.create_s
                    pool.add(m)
                    active[m.rid] = (pool, m)
                else:
                    still.append((pool, prep, fut))
            creating.extend(still)
            st["max_resident"] = max(st["max_resident"], len(active))

        def arrivals(ex):
            t = now()
            while pending and (not realtime or pending[0].arrival <= t):
                r = pending.popleft()
                waiting.append((r, ex.submit(prepare, r)))

        with ThreadPoolExecutor(max_workers=self.workers) as ex:
            while pending or waiting or active or creating:
                arrivals(ex)
                if not active:
                    admit(ex)
                    if not active:
                        if creating:
                            wait([creating[0][2]])  # device objects being built
                        elif waiting:
                            wait([waiting[0][1]])  # a preparation
                        elif pending and realtime:
                            time.sleep(max(0.0, min(pending[0].arrival - now(), 0.05)))
                        continue
                # one chunk for every pool with members (asynchronous launches);
                # requests admitted meanwhile join at the next chunk
                for pool in self.pools.values():
                    if pool.members:
                        pool.run(self.chunk)
                        st["chunks"] += 1
                finished = [(pool, m) for pool, m in active.values()
                            if pool.next_step(m) > m.nt]
                arrivals(ex)
                admit(ex)
                tw = time.perf_counter()
                for pool in self.pools.values():
                    pool.synchronize()
                st["gpu_wait_s"] += time.perf_counter() - tw
                t_done = now()
                for pool, m in finished:
                    pool.remove(m)
                    g = m.ds.gather()
                    m.close()  # the device objects are freed here, on the pools' thread
                    c = Completed(m.rid, m.prep.req.arrival, m.admitted, t_done, g,
                                  m.prep.req.storage, m.prep.req.fd_order,
                                  m.prep.M.nx * m.prep.M.nz * m.nt)
                    done.append(c)
                    if on_complete:
                        on_complete(c)
                    del active[m.rid]
                    used -= m.prep.bytes
                    slots[pool.key] -= 1
                    del m
        return done


def serial(requests: list, b: Optional[E.MI355XBackend] = None, storage: Optional[str] = "f32",
           prep_times: Optional[list] = None, on_complete=None) -> list:
    """The same requests one at a time with the plain single-shot path
    (``run_shot``), storage forced to `storage` (None keeps each request's).
    Each request's host preparation runs in the loop, as a client's would;
    its seconds are appended to `prep_times` when given."""
    b = b or E.backend("mi355x")
    out = []
    t0 = time.perf_counter()
    for r in requests:
        rr = ShotRequest(**{**r.__dict__, "storage": storage or r.storage})
        ta = time.perf_counter() - t0
        prep = prepare(rr)
        if prep_times is not None:
            prep_times.append(prep.prep_s)
        W = E.Wavefield(prep.M.nx, prep.M.nz, b)
        with _Tiling(b, rr.fd_order):
            res = E.run_shot(b, W, prep.M, prep.H, prep.a, prep.src, prep.rec, prep.p)
        out.append(Completed(r.rid, ta, ta, time.perf_counter() - t0, res.gather, rr.storage,
                             rr.fd_order, prep.M.nx * prep.M.nz * prep.p.nt))
        if on_complete:
            on_complete(out[-1])
        b._models.clear()
        del W
    return out


def summarize(done: list, wall: float) -> dict:
    """shots/hour, cell-updates/s and latency percentiles of a served run."""
    lat = np.array([c.latency for c in done]) if done else np.zeros(1)
    cells = sum(c.cell_updates for c in done)
    return {"shots": len(done), "wall_s": round(wall, 3),
            "shots_per_hour": round(len(done) / wall * 3600.0, 1) if wall > 0 else None,
            "mcell_updates_per_s": round(cells / wall / 1e6, 1) if wall > 0 else None,
            "latency_p50_s": round(float(np.percentile(lat, 50)), 3),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 3),
            "latency_max_s": round(float(lat.max()), 3)}
