"""The tile autotuner's host logic, on the CPU: tune.autotune_batch with a
stand-in backend and timer.  It must time every candidate, skip one the
library rejects, keep the fastest, set it on the backend, reset the members
after every trial and the shots at the end, and reuse its choice for a batch
of the same class without timing again."""
from types import SimpleNamespace

import pytest

from fomo_julia_amd import _lib
from fomo_julia_amd import tune as T

TIMES = {(1, 0): 0.93, (2, 0): 0.83, (4, 0): 1.18, (2, 16): 0.99, (1, 48): 0.97}


class FakeBackend:
    def __init__(self):
        self.tiling = None
        self.calls = []
        self.lib = SimpleNamespace(fomo_shot_reset=lambda h: self.calls.append(("shot_reset", h)) or 0)

    def set_tiling(self, cpt, tz, gsteps):
        self.tiling = (cpt, tz, gsteps)

    def check(self, rc, what):
        assert rc == 0, what

    def synchronize(self):
        self.calls.append(("sync",))


def _batch(n=3, nx=400, nz=300, order=12, storage="f32"):
    M = SimpleNamespace(nx=nx, nz=nz, storage=storage)
    dms = [SimpleNamespace(M=M, p=SimpleNamespace(fd_order=order), h=None) for _ in range(n)]
    wfs = [SimpleNamespace(name=f"w{i}") for i in range(n)]
    dss = [SimpleNamespace(h=f"s{i}") for i in range(n)]
    return dms, wfs, dss


@pytest.fixture
def fake(monkeypatch):
    b = FakeBackend()
    timed, resets = [], []

    def fake_time(bk, dms, wfs, dss, steps, warm=4):
        cpt, tz, _ = bk.tiling
        timed.append((cpt, tz))
        assert steps % 2 == 0
        if (cpt, tz) == (4, 0):
            raise _lib.FomoError("rejected by the library")
        return TIMES[(cpt, tz)]

    monkeypatch.setattr(T, "time_batch", fake_time)
    monkeypatch.setattr(T.E, "reset", lambda bk, w: resets.append(w.name))
    monkeypatch.setattr(T, "_cache", {})
    return b, timed, resets


def test_picks_the_fastest_and_resets(fake):
    b, timed, resets = fake
    dms, wfs, dss = _batch()
    best = T.autotune_batch(b, dms, wfs, dss, steps=31, graph_steps=8)
    assert timed == list(T.CANDIDATES)
    assert best["tiling"] == (2, 0) and best["ms_per_step"] == TIMES[(2, 0)]
    assert b.tiling == (2, 0, 8)
    assert set(best["trials"]) == {"1x0", "2x0", "2x16", "1x48"}  # the rejected one is skipped
    # members reset after each timed trial, shots once at the end
    assert resets == [w.name for _ in range(4) for w in wfs]
    assert [c for c in b.calls if c[0] == "shot_reset"] == [("shot_reset", d.h) for d in dss]


def test_choice_is_cached_per_class(fake):
    b, timed, resets = fake
    dms, wfs, dss = _batch()
    first = T.autotune_batch(b, dms, wfs, dss)
    n_timed = len(timed)
    b.tiling = None
    again = T.autotune_batch(b, *_batch())  # same order, storage, members, cells
    assert again is first and len(timed) == n_timed and b.tiling[:2] == (2, 0)
    T.autotune_batch(b, *_batch(storage="bf16"))   # another class: timed again
    assert len(timed) == 2 * n_timed
    T.autotune_batch(b, *_batch(n=4))
    assert len(timed) == 3 * n_timed
