"""fomo_shot_gather_noalloc against fomo_shot_gather, through the C ABI.

The no-allocation entry point transposes the device gather through the
library's fixed 16 MiB device array in chunks of time steps.  Two calls in a
row, the second on a smaller shot: the first gather (4000 receivers x 1200
steps = 4.8 M samples) needs three chunks, the second fits in one; each must
equal fomo_shot_gather's result bit for bit.  A shot with more receivers than
the array holds for 64 steps is refused before any device work."""
import ctypes

import numpy as np
import pytest

from fomo_julia_amd import _lib
from fomo_julia_amd import elastic2d as E
from tests.helpers import Case

pytestmark = pytest.mark.gpu
F32 = np.float32


def _ptr(a):
    return a.ctypes.data_as(_lib.P_f32)


def _both(b, ds):
    want = np.zeros((ds.nt, ds.n_rec), F32, order="F")
    got = np.full((ds.nt, ds.n_rec), np.nan, F32, order="F")  # every sample must be written
    b.check(b.lib.fomo_shot_gather(ds.h, _ptr(want)), "fomo_shot_gather")
    b.check(b.lib.fomo_shot_gather_noalloc(ds.h, _ptr(got)), "fomo_shot_gather_noalloc")
    return want, got


def _shot(b, c, n_rec, seed):
    rng = np.random.default_rng(seed)
    xr = rng.uniform(0.0, float(c.M.x_max), n_rec).astype(F32)
    zr = rng.uniform(0.0, 20.0 * float(c.M.dz), n_rec).astype(F32)
    types = [("vz", "vx", "p")[k % 3] for k in range(n_rec)]
    rec = E.setup_receivers(xr, zr, c.M, type=types)
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    E.run_shot(b, W, c.M, c.H, c.a, c.src, rec, c.p)
    dm = E._model(b, c.M, c.H, c.a, c.p)
    return E._shot_for(dm, c.src, rec, c.nt), W


def test_noalloc_gather_equals_gather_on_two_calls_second_smaller():
    b = E.backend("mi355x")
    b.set_tiling(1, 0, 16)
    b.set_step_mode(1)
    big = Case(nx_in=120, nz_in=80, nbc=12, nt=1200)
    print("case noalloc gather: 4000 receivers x 1200 steps, then 23 x 70", flush=True)
    ds, W = _shot(b, big, 4000, 1)
    assert ds.nt * ds.n_rec > 4 << 20  # more than one chunk of the 16 MiB array
    want, got = _both(b, ds)
    assert np.abs(want).max() > 0
    assert np.array_equal(got, want)
    small = Case(nx_in=60, nz_in=40, nbc=8, nt=70)
    ds2, W2 = _shot(b, small, 23, 2)
    want2, got2 = _both(b, ds2)
    assert np.abs(want2).max() > 0
    assert np.array_equal(got2, want2)
    # the first shot again, after the array held the second one's chunk
    want3, got3 = _both(b, ds)
    assert np.array_equal(got3, want)


def test_noalloc_gather_refuses_more_receivers_than_its_array_holds():
    b = E.backend("mi355x")
    c = Case(nx_in=60, nz_in=40, nbc=8, nt=2)
    dm = E._model(b, c.M, c.H, c.a, c.p)
    n_rec = 65537
    rec = E.Receivers(np.full(n_rec, c.src.i, np.int64), np.full(n_rec, c.src.j, np.int64),
                      np.zeros((1, n_rec), F32), "vz")
    ds = E._DeviceShot(dm, c.src, rec, 2)
    out = np.zeros((2, n_rec), F32, order="F")
    assert b.lib.fomo_shot_gather_noalloc(ds.h, _ptr(out)) == _lib.FOMO_ERR_ARG
    msg = b.lib.fomo_ctx_last_error(b.ctx).decode()
    assert "65536" in msg and "fomo_shot_gather" in msg
    b.check(b.lib.fomo_shot_gather(ds.h, _ptr(out)), "fomo_shot_gather")
