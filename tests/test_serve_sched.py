"""The serving loop's scheduling, on the CPU: serve.Server driven with stand-in
pools and members that only count steps (no device), so its host logic is
checked on its own:
- every request completes once, with its own gather, and no earlier than
  nt steps of chunks after it joined;
- a pool never holds more members than its capacity;
- device memory of the resident shots stays within the budget, except for a
  lone request larger than the budget (admitted when nothing else runs);
- requests of different (fd_order, storage) go to different pools;
- with real-time arrivals nothing is admitted before it arrives;
- threads: members are built on worker threads while the pools run, but
  every pool call and every free of a member's device objects happens on
  the serving thread; a request whose preparation or device objects fail is
  reported and releases its budget, and the others complete."""
import threading
import time
from types import SimpleNamespace

import numpy as np
import pytest

from fomo_julia_amd import serve as S


class _World:
    def __init__(self):
        self.pool_threads = set()     # threads that made pool calls
        self.create_threads = set()   # threads that built members
        self.close_threads = []       # (rid, thread) of every member close
        self.fail_prepare = set()     # rids whose preparation raises
        self.fail_create = set()      # rids whose device objects raise
        self.resident = {}   # rid -> bytes
        self.max_resident_bytes = 0
        self.pools = []
        self.added = {}      # rid -> time added
        self.steps_run = {}  # rid -> steps its pool ran while it was a member


def _install(monkeypatch, world):
    class FakePool:
        def __init__(self, b, fd_order, storage, capacity):
            self.key = (fd_order, storage)
            self.capacity = capacity
            self.members = {}
            self.k = {}
            world.pools.append(self)

        def add(self, m):
            world.pool_threads.add(threading.get_ident())
            assert len(self.members) < self.capacity, "pool over capacity"
            assert m.prep.req.fd_order == self.key[0] and m.prep.req.storage == self.key[1]
            self.members[m.rid] = m
            self.k[m.rid] = 1
            world.resident[m.rid] = m.prep.bytes
            world.max_resident_bytes = max(world.max_resident_bytes, sum(world.resident.values()))
            world.added[m.rid] = time.perf_counter()
            world.steps_run[m.rid] = 0

        def remove(self, m):
            world.pool_threads.add(threading.get_ident())
            del self.members[m.rid]
            del world.resident[m.rid]

        def run(self, n):
            world.pool_threads.add(threading.get_ident())
            assert n > 0 and n % 2 == 0
            for rid, m in self.members.items():
                self.k[rid] = min(self.k[rid] + n, m.nt + 1)
                world.steps_run[rid] += n

        def synchronize(self):
            world.pool_threads.add(threading.get_ident())

        def next_step(self, m):
            world.pool_threads.add(threading.get_ident())
            return self.k[m.rid]

    class FakeMember:
        def __init__(self, b, prep, admitted):
            world.create_threads.add(threading.get_ident())
            time.sleep(0.002)  # device setup takes a while, overlapping the pools' runs
            if prep.req.rid in world.fail_create:
                raise RuntimeError("fomo_model_create_from_velocity failed (-2): out of memory")
            self.rid = prep.req.rid
            self.prep = prep
            self.admitted = admitted
            self.nt = prep.p.nt
            self.ds = SimpleNamespace(gather=lambda rid=self.rid, nt=self.nt:
                                      np.full((nt, 1), rid, np.float32))

        def close(self):
            world.close_threads.append((self.rid, threading.get_ident()))

    def fake_prepare(req):
        if req.rid in world.fail_prepare:
            raise ValueError("medium generation failed")
        nb = req.nbc + req.fd_order // 2
        M = SimpleNamespace(nx=req.nx + 2 * nb, nz=req.nz + 2 * nb)
        return SimpleNamespace(req=req, M=M, p=SimpleNamespace(nt=req.nt),
                               bytes=req.nx * req.nz * 100, prep_s=0.0)

    monkeypatch.setattr(S, "Pool", FakePool)
    monkeypatch.setattr(S, "_Member", FakeMember)
    monkeypatch.setattr(S, "prepare", fake_prepare)


def _requests(n, seed=3, arrival_step=0.0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        out.append(S.ShotRequest(rid=i, arrival=i * arrival_step,
                                 nx=int(rng.integers(100, 400)), nz=int(rng.integers(100, 300)),
                                 nt=int(rng.integers(50, 400)), fd_order=int(rng.choice([4, 8, 12])),
                                 storage=str(rng.choice(["f32", "bf16"]))))
    return out


@pytest.mark.parametrize("budget_shots,capacity", [(1000, 64), (3, 64), (1000, 2)])
def test_every_request_completes_within_budget_and_capacity(monkeypatch, budget_shots, capacity):
    world = _World()
    _install(monkeypatch, world)
    reqs = _requests(30)
    biggest = max(r.nx * r.nz * 100 for r in reqs)
    budget = budget_shots * biggest
    srv = S.Server(b=object(), budget_bytes=budget, chunk=32, capacity=capacity, workers=3)
    done = srv.serve(reqs, realtime=False)
    assert sorted(c.rid for c in done) == list(range(len(reqs)))
    for c in done:
        r = reqs[c.rid]
        assert np.all(c.gather == c.rid) and c.gather.shape == (r.nt, 1)
        assert (c.fd_order, c.storage) == (r.fd_order, r.storage)
        assert world.steps_run[c.rid] >= r.nt
        assert c.done >= c.admitted >= c.arrival
    assert world.max_resident_bytes <= budget
    assert not world.resident
    keys = {p.key for p in world.pools}
    assert keys == {(r.fd_order, r.storage) for r in reqs} and len(keys) == len(world.pools)
    assert srv.stats["max_resident"] <= min(len(reqs), capacity * len(keys))


def test_request_larger_than_the_budget_runs_alone(monkeypatch):
    world = _World()
    _install(monkeypatch, world)
    reqs = _requests(6)
    budget = min(r.nx * r.nz * 100 for r in reqs) - 1  # no request fits the budget
    srv = S.Server(b=object(), budget_bytes=budget, chunk=16, capacity=8, workers=2)
    done = srv.serve(reqs, realtime=False)
    assert sorted(c.rid for c in done) == list(range(6))
    assert srv.stats["max_resident"] == 1


def test_realtime_arrivals_are_not_admitted_early(monkeypatch):
    world = _World()
    _install(monkeypatch, world)
    reqs = _requests(8, arrival_step=0.02)
    srv = S.Server(b=object(), budget_bytes=1 << 40, chunk=64, capacity=64, workers=2)
    done = srv.serve(reqs, realtime=True)
    assert sorted(c.rid for c in done) == list(range(8))
    for c in done:
        assert c.admitted >= c.arrival - 1e-3
    s = S.summarize(done, 1.0)
    assert s["shots"] == 8 and s["latency_p99_s"] >= s["latency_p50_s"] >= 0


def test_pools_have_one_owner_and_failures_release_their_budget(monkeypatch):
    world = _World()
    _install(monkeypatch, world)
    reqs = _requests(24, seed=7)
    world.fail_prepare = {3, 17}
    world.fail_create = {5, 11, 20}
    biggest = max(r.nx * r.nz * 100 for r in reqs)
    budget = 4 * biggest
    srv = S.Server(b=object(), budget_bytes=budget, chunk=16, capacity=8, workers=4)
    me = threading.get_ident()
    done = srv.serve(reqs, realtime=False)
    failed = {f.rid: f.stage for f in srv.failed}
    assert failed == {3: "prepare", 17: "prepare", 5: "admit", 11: "admit", 20: "admit"}
    assert sorted(c.rid for c in done) == sorted(set(range(24)) - set(failed))
    # one owner: every pool call on the serving thread; members built on workers
    assert world.pool_threads == {me}
    assert world.create_threads and me not in world.create_threads
    # every completed member's device objects were freed, once, on the serving thread
    assert sorted(r for r, _ in world.close_threads) == sorted(c.rid for c in done)
    assert {t for _, t in world.close_threads} == {me}
    # the failed reservations were released: the budget held and nothing stayed resident
    assert world.max_resident_bytes <= budget and not world.resident


def test_real_pool_refuses_calls_from_another_thread():
    """serve.Pool's owner check runs before any library call."""
    pool = S.Pool.__new__(S.Pool)
    pool.owner = threading.get_ident()
    pool.h = None
    errors = []

    def other():
        for call in (lambda: pool.run(16), lambda: pool.synchronize(), lambda: pool.close()):
            try:
                call()
            except RuntimeError as e:
                errors.append(str(e))

    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert len(errors) == 3 and all("does not own" in e for e in errors)
