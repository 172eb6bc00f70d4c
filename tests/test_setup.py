"""Device-side model setup (fomo_model_create_from_velocity).

The device path pads vp, vs, rho, forms the staggered averages and the
coefficients on the GPU; the host path (init_medium -> fomo_model_create)
builds the same arrays on the CPU.  Both must give the same model: the GPU
tests run one shot through each and compare gathers, all five wavefields and
the FP16 field scales bit for bit, for every storage type, every source type
(force sources read the staggered density at the source cell), dx != dz (the
seven-array coefficient layout), a fluid layer (mu = 0, so mu_txz divides by
zero) and edge-replicated padding on grids whose sizes are not multiples of
the tile."""
import numpy as np
import pytest

from fomo_julia_amd import elastic2d as E
from fomo_julia_amd import init as I

F32 = np.float32
gpu = pytest.mark.gpu


def _media(nx_in, nz_in, seed, water):
    rng = np.random.default_rng(seed)
    vp = (2400 + 900 * rng.random((nx_in, nz_in))).astype(F32)
    vs = (vp / F32(1.8) - rng.uniform(0, 50, (nx_in, nz_in))).astype(F32)
    rho = (1900 + 500 * rng.random((nx_in, nz_in))).astype(F32)
    if water:
        vp[:, :4] = 1500.0
        vs[:, :4] = 0.0
        rho[:, :4] = 1000.0
    return vp, vs, rho


def _build(device, vp, vs, rho, dx, dz, nbc, fd_order, storage, src_type, nt):
    init = E.init_medium_device if device else E.init_medium
    M = init(vp, vs, rho, dx, dz, nbc, fd_order, None, free_surface=True, storage=storage)
    dt = F32(0.4 * min(float(dx), float(dz)) / float(vp.max()))
    H = E.init_habc(M.nx, M.nz, nbc, dt, dx, dz, I.julia_mean_f32(vp), None, weights=False)
    p = E.SimParams(dt, nt, dx, dz, fd_order)
    a = I.get_fd_coefficients(fd_order)
    wav = I.ricker_wavelet(F32(20.0), dt, nt)
    si = I.coord_to_index(F32(M.x_max / F32(3)), M.dx, M.pad)
    sj = I.coord_to_index(F32(6 * float(dz)), M.dz, M.pad)
    src = E.Source(si, sj, wav, src_type)
    nx_in = vp.shape[0]
    xr = np.arange(0, nx_in, 4, dtype=F32) * F32(dx)
    types = [("vz", "vx", "p")[k % 3] for k in range(len(xr))]
    rec = E.setup_receivers(xr, np.full(len(xr), F32(2 * float(dz)), F32), M, type=types)
    return M, H, p, a, src, rec


def test_device_medium_scalars_match_init_medium():
    vp, vs, rho = _media(37, 29, 0, False)
    for fd_order in (2, 8):
        host = E.init_medium(vp, vs, rho, F32(7.5), F32(5.0), 10, fd_order, None)
        dev = E.init_medium_device(vp, vs, rho, F32(7.5), F32(5.0), 10, fd_order, None)
        for f in ("nx", "nz", "dx", "dz", "x_max", "z_max", "M", "pad", "is_free_surface", "storage"):
            assert getattr(dev, f) == getattr(host, f), f
        assert dev.rho_vx is None and dev.lam is None


def test_device_medium_rejects_bad_input():
    vp, vs, rho = _media(20, 20, 0, False)
    with pytest.raises(ValueError):
        E.init_medium_device(vp, vs[:, :10], rho, 5.0, 5.0, 10, 8)
    with pytest.raises(ValueError):
        E.init_medium_device(vp, vs, rho, 5.0, 5.0, 10, 7)
    with pytest.raises(ValueError):
        E.init_medium_device(vp, vs, rho, 5.0, 5.0, 10, 8, storage="f8")


@pytest.fixture(scope="module")
def b():
    bk = E.backend("mi355x")
    bk.set_tiling(1, 0, 16)
    bk.set_step_mode(1)
    return bk


def _run(b, parts, nt):
    M, H, p, a, src, rec = parts
    W = E.Wavefield(M.nx, M.nz, b)
    res = E.run_shot(b, W, M, H, a, src, rec, p)
    dm = E._model(b, M, H, a, p)
    scales = E._shot_for(dm, src, rec, nt).field_scales()
    return res.gather, [W.download(f) for f in range(5)], list(scales)


@gpu
@pytest.mark.parametrize("storage", ["f32", "bf16", "f16", "f16c"])
@pytest.mark.parametrize("src_type,dz,water", [("pressure", 10.0, True), ("force_z", 7.5, False),
                                               ("force_x", 10.0, True)])
def test_device_setup_matches_host_setup(b, storage, src_type, dz, water):
    fd_order = 4 if src_type == "force_z" else 8
    nt = 90
    vp, vs, rho = _media(83, 61, 3, water)
    print(f"case setup storage={storage} src={src_type} dz={dz} water={water}", flush=True)
    args = (vp, vs, rho, F32(10.0), F32(dz), 12, fd_order, storage, src_type, nt)
    g0, f0, s0 = _run(b, _build(False, *args), nt)
    g1, f1, s1 = _run(b, _build(True, *args), nt)
    assert np.abs(g0).max() > 0
    assert s1 == s0
    assert np.array_equal(g1, g0), int((g1 != g0).sum())
    for k in range(5):
        assert np.array_equal(f1[k], f0[k]), (k, int((f1[k] != f0[k]).sum()))


@gpu
def test_device_setup_large_grid_runs():
    """A grid whose setup spans many blocks of the two-pass scan (the
    coefficient maxima and the impedance range are reductions over blocks)."""
    bk = E.backend("mi355x")
    vp, vs, rho = _media(1201, 700, 5, True)
    args = (vp, vs, rho, F32(10.0), F32(10.0), 30, 8, "f16", "pressure", 12)
    g0, f0, s0 = _run(bk, _build(False, *args), 12)
    g1, f1, s1 = _run(bk, _build(True, *args), 12)
    assert s1 == s0
    assert np.array_equal(g1, g0)
    for k in range(5):
        assert np.array_equal(f1[k], f0[k]), k
