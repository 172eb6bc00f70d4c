"""Serving loop: continuous batching in fomo_pool.

Every served shot must equal the same shot run alone through the plain
single-shot path (bit for bit: the pool runs each shot's arithmetic exactly
as the single-shot kernels do), whatever else is in the pool, whenever it
joined and whatever its storage and FD order."""
import numpy as np
import pytest

from fomo_julia_amd import serve as S

SMALL = dict(nx=(60, 140), nz=(50, 100), nt=(60, 160))


def test_stream_is_deterministic_and_in_range():
    a = S.c5_stream(40, seed=5, rate=2.0)
    b = S.c5_stream(40, seed=5, rate=2.0)
    assert [r.__dict__ for r in a] == [r.__dict__ for r in b]
    assert all(300 <= r.nx <= 2000 and 300 <= r.nz <= 1200 and 1000 <= r.nt <= 6000 for r in a)
    assert {r.fd_order for r in a} <= {4, 8, 12} and {r.storage for r in a} == {"f32", "bf16"}
    t = [r.arrival for r in a]
    assert t == sorted(t) and t[0] > 0
    assert all(r.arrival == 0.0 for r in S.c5_stream(5, rate=None))


def test_prepare_matches_request():
    r = S.c5_stream(1, seed=9, **SMALL)[0]
    p = S.prepare(r)
    pad = r.nbc + r.fd_order // 2
    assert (p.M.nx, p.M.nz) == (r.nx + 2 * pad, r.nz + 2 * pad)
    assert p.p.nt == r.nt and len(p.src.wavelet) == r.nt and p.M.storage == r.storage
    assert p.bytes == S.device_bytes(p.M.nx, p.M.nz, r.storage, r.nt, len(p.rec.i))


gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def b():
    from fomo_julia_amd import elastic2d as E
    bk = E.backend("mi355x")
    bk.set_tiling(1, 0, 16)
    bk.set_step_mode(1)
    return bk


def _alone(b, r):
    return S.serial([r], b, storage=None)[0].gather


@gpu
def test_served_equals_alone_mixed_orders_and_storage(b):
    reqs = S.c5_stream(10, seed=21, storages=("f32", "bf16", "f16", "f16c", "f24"), **SMALL)
    print("case served mix of", [(r.nx, r.nz, r.nt, r.fd_order, r.storage) for r in reqs], flush=True)
    srv = S.Server(b, budget_bytes=8 << 30, chunk=16, capacity=8)
    done = srv.serve(reqs, realtime=False)
    assert sorted(c.rid for c in done) == list(range(10))
    for c in done:
        want = _alone(b, reqs[c.rid])
        assert np.array_equal(c.gather, want), f"request {c.rid} differs from its solo run"


@gpu
def test_staggered_joins_and_small_budget(b):
    reqs = S.c5_stream(8, seed=33, orders=(8,), storages=("f32",), **SMALL)
    # joins at different chunk boundaries: arrivals spread over the run
    reqs = [S.ShotRequest(**{**r.__dict__, "arrival": 0.02 * k}) for k, r in enumerate(reqs)]
    print("case staggered joins", flush=True)
    one = max(S.prepare(r).bytes for r in reqs)
    srv = S.Server(b, budget_bytes=2 * one, chunk=8, capacity=8)
    seen = []
    done = srv.serve(reqs, realtime=True, on_complete=lambda c: seen.append(c.rid))
    assert len(done) == 8 and seen == [c.rid for c in done]
    for c in done:
        assert np.array_equal(c.gather, _alone(b, reqs[c.rid]))
        assert c.admitted >= c.arrival and c.done >= c.admitted


@gpu
def test_device_bytes_estimate_matches_library(b):
    from fomo_julia_amd import elastic2d as E
    for storage in ("f32", "bf16", "f16c", "f24"):
        r = S.c5_stream(1, seed=3, storages=(storage,), **SMALL)[0]
        p = S.prepare(r)
        dm = E._DeviceModel(b, p.M, p.H, p.a, p.p)
        W = E.Wavefield(p.M.nx, p.M.nz, b)
        W._bind(dm)
        lib = b.lib
        got = lib.fomo_model_device_bytes(dm.h) + lib.fomo_wavefield_device_bytes(W._h)
        gather = r.nt * len(p.rec.i) * 4
        assert abs(p.bytes - (got + gather)) <= 3 * 4096 + 4 * (p.M.nz + 36 + 8), storage


def _member_objects(b, r):
    from fomo_julia_amd import elastic2d as E
    p = S.prepare(r)
    dm = E._DeviceModel(b, p.M, p.H, p.a, p.p)
    W = E.Wavefield(p.M.nx, p.M.nz, b)
    W._bind(dm)
    ds = E._DeviceShot(dm, p.src, p.rec, p.p.nt)
    return p, dm, W, ds


@gpu
def test_reused_objects_join_after_reset_on_the_context_stream(b):
    """run alone, reset (queued on the context's stream), add to a pool, run:
    the pool's first steps wait for the zeroing, so the gather equals the
    solo run's bit for bit."""
    from fomo_julia_amd import _lib
    r = S.c5_stream(1, seed=41, orders=(8,), storages=("f32",), **SMALL)[0]
    print("case reset-then-join", r, flush=True)
    p, dm, W, ds = _member_objects(b, r)
    lib = b.lib
    with S._Tiling(b, r.fd_order):
        b.check(lib.fomo_run_time_loop(dm.h, W._h, ds.h, 1, p.p.nt), "solo")
        b.synchronize()
        solo = ds.gather()
        pool = S.Pool(b, r.fd_order, r.storage, 4)
    for _ in range(2):
        b.check(lib.fomo_wavefield_reset(W._h), "reset")
        b.check(lib.fomo_shot_reset(ds.h), "shot reset")
        b.check(lib.fomo_pool_add(pool.h, dm.h, W._h, ds.h), "add")
        steps = (p.p.nt + 1) // 2 * 2
        pool.run(steps)
        pool.synchronize()
        b.check(lib.fomo_pool_remove(pool.h, ds.h), "remove")
        assert np.array_equal(ds.gather(), solo)
    pool.close()
    for obj in (ds, W, dm):
        obj.close()


@gpu
def test_pool_runs_while_other_threads_create_and_free_objects(b):
    """Worker threads create and destroy models, wavefields and shots while
    the owner thread runs a pool: the library lock keeps allocations and
    frees out of the pool's graph capture, a pool member cannot be destroyed
    while in the pool, and every served gather equals its solo run."""
    import threading
    from fomo_julia_amd import _lib
    reqs = S.c5_stream(3, seed=52, orders=(8,), storages=("f32",), **SMALL)
    print("case frees from worker threads", [(r.nx, r.nz, r.nt) for r in reqs], flush=True)
    members = [_member_objects(b, r) for r in reqs]
    with S._Tiling(b, 8):
        pool = S.Pool(b, 8, "f32", 4)
    for p, dm, W, ds in members:
        b.check(b.lib.fomo_pool_add(pool.h, dm.h, W._h, ds.h), "add")
    # a member in the pool cannot be destroyed
    assert b.lib.fomo_shot_destroy(members[0][3].h) == _lib.FOMO_ERR_STATE
    assert b.lib.fomo_wavefield_destroy(members[0][2]._h) == _lib.FOMO_ERR_STATE
    assert b.lib.fomo_model_destroy(members[0][1].h) == _lib.FOMO_ERR_STATE
    stop = threading.Event()
    churn = {"n": 0, "err": None}
    spare = S.c5_stream(1, seed=60, orders=(4,), storages=("f32",), **SMALL)[0]

    def worker():
        try:
            while not stop.is_set():
                objs = _member_objects(b, spare)
                for obj in (objs[3], objs[2], objs[1]):
                    obj.close()
                churn["n"] += 1
        except Exception as e:  # reported to the owner thread
            churn["err"] = e

    th = [threading.Thread(target=worker) for _ in range(2)]
    for t in th:
        t.start()
    nt = max(r.nt for r in reqs)
    k = 0
    while k < nt:
        pool.run(16)
        pool.synchronize()
        k += 16
    stop.set()
    for t in th:
        t.join()
    assert churn["err"] is None, churn["err"]
    assert churn["n"] > 0
    for (p, dm, W, ds), r in zip(members, reqs):
        b.check(b.lib.fomo_pool_remove(pool.h, ds.h), "remove")
        assert np.array_equal(ds.gather(), _alone(b, r)), r.rid
    pool.close()
    for p, dm, W, ds in members:
        for obj in (ds, W, dm):
            obj.close()


@gpu
def test_served_stream_at_its_real_grid_sizes(b):
    """Requests drawn from the served stream's own ranges (nx 300-2000, nz
    300-1200, fd_order 4/8/12, FP32 and BF16 storage), nt cut to about 200:
    each served gather equals its solo run bit for bit, and one FP32 request
    equals the oracle's run of the same medium, source and receivers."""
    import oracle as O
    from fomo_julia_amd import elastic2d as E
    from fomo_julia_amd import init as I
    reqs = S.c5_stream(8, seed=5, nt=(200, 240))
    print("case C5 real sizes", [(r.nx, r.nz, r.nt, r.fd_order, r.storage) for r in reqs], flush=True)
    assert {r.storage for r in reqs} == {"f32", "bf16"}
    srv = S.Server(b, budget_bytes=32 << 30, chunk=64, capacity=16)
    done = srv.serve(reqs, realtime=False)
    assert not srv.failed and sorted(c.rid for c in done) == list(range(8))
    for c in done:
        assert np.array_equal(c.gather, _alone(b, reqs[c.rid])), f"request {c.rid} differs from its solo run"
    # one FP32 request against the oracle (host setup of the same medium)
    r = next(q for q in reqs if q.storage == "f32")
    p = S.prepare(r)
    vp, vs, rho = p.M._velocity
    Mh = E.init_medium(vp, vs, rho, p.M.dx, p.M.dz, r.nbc, r.fd_order, None, free_surface=True)
    w = [I.habc_weight_matrix(Mh.nx, Mh.nz, r.nbc, off) for off in (0.0, 0.5, 0.75)]
    H = p.H
    orc = O.OracleRun(nx=Mh.nx, nz=Mh.nz, pad=Mh.pad, nbc=r.nbc, fd_order=r.fd_order,
                      free_surface=True, dt=p.p.dt, dtx=p.p.dtx, dtz=p.p.dtz,
                      q=(H.qx, H.qz, H.qt_x, H.qt_z, H.qxt), a=p.a, rho_vx=Mh.rho_vx,
                      rho_vz=Mh.rho_vz, lam=Mh.lam, mu_txx=Mh.mu_txx, mu_txz=Mh.mu_txz,
                      w_vx=w[0], w_vz=w[1], w_tau=w[2], src_i=p.src.i, src_j=p.src.j, src_type=0,
                      wavelet=p.src.wavelet, rec_i=p.rec.i, rec_j=p.rec.j, rec_type=0, nt=r.nt,
                      dtype=np.float32, flags=O.FMA)
    orc.run(1, r.nt)
    served = next(c for c in done if c.rid == r.rid)
    assert np.array_equal(served.gather, orc.gather_nt_by_nrec()), "served FP32 request vs oracle"
    assert np.abs(served.gather).max() > 0
