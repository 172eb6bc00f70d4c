"""The data formats on either side of the path, fed by and feeding the GPU:
- gathers of a batched MI355X run_shots go through save_gather / load_gather
  (src/io/output.jl:18-33,60-70) unchanged, bit for bit, and create_geometry
  (src/io/geometry_io.jl:85-130) gives the grid positions the shots used;
- a model file read by load_model (src/io/model_loader.jl:88-106,199-248:
  a SEG-Y vp with the reference's vs and Gardner rho estimates) drives a
  device-side setup (init_medium_device) and a GPU shot equal to the oracle's
  run on init_medium's host arrays of the same model, bit for bit."""
import numpy as np
import pytest

from fomo_julia_amd import elastic2d as E
from fomo_julia_amd import geometry as G
from fomo_julia_amd import init as I
from fomo_julia_amd import model_loader as ML
from fomo_julia_amd import output as OUT
from tests.helpers import Case

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def b():
    bk = E.backend("mi355x")
    bk.set_tiling(1, 0, 16)
    bk.set_step_mode(1)
    return bk


def test_gathers_and_geometry_of_gpu_shots(b, tmp_path):
    c = Case(nx_in=90, nz_in=60, nbc=12, nt=90)
    xs = np.array([100.0, 333.3, 560.0], F32)
    zs = np.array([30.0, 41.0, 25.0], F32)
    cfg = E.MultiShotConfig(xs, zs, c.src.wavelet, source_type="pressure")
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    print("case gather files of a 3-shot batch", flush=True)
    results = E.run_shots(b, W, c.M, c.H, c.a, c.rec, cfg, c.p, batch=3)
    assert [r.shot_id for r in results] == [1, 2, 3]
    geom = G.create_geometry(results, c.M, c.p)
    pad, dx = c.M.pad, F32(c.M.dx)
    for k, r in enumerate(results):
        assert r.gather.shape == (c.nt, len(c.rec.i)) and np.abs(r.gather).max() > 0
        path = str(tmp_path / f"shot_{k + 1}.bin")
        OUT.save_gather(r, path)
        raw = np.fromfile(path, dtype="<f4")
        assert raw.size == r.gather.size
        assert np.array_equal(raw, np.asarray(r.gather, F32).ravel(order="F"))  # column-major
        assert np.array_equal(OUT.load_gather(path, c.nt, len(c.rec.i)), r.gather)
        s = geom.shots[k]
        assert np.array_equal(OUT.load_gather(path, s), r.gather)
        # the source sits where _coord_to_index put it: round(x/dx) cells from the origin
        assert (r.src_i, r.src_j) == (I.coord_to_index(xs[k], c.M.dx, pad),
                                      I.coord_to_index(zs[k], c.M.dz, pad))
        assert s.src_x == F32(F32(r.src_i - pad - 1) * dx)
        assert s.src_x == F32(np.rint(xs[k] / dx)) * dx and s.src_z == F32(np.rint(zs[k] / dx)) * dx
        assert s.rec_i_idx == [i - pad for i in r.rec_i]
        assert np.array_equal(s.rec_x, np.array([F32(i - pad - 1) * dx for i in r.rec_i], F32))
        assert (s.nt, s.n_rec, s.nx, s.nz) == (c.nt, len(c.rec.i), c.M.nx - 2 * pad, c.M.nz - 2 * pad)
        assert s.t_max == F32(F32(c.p.dt) * F32(c.nt))
    # the batched results equal the one-at-a-time loop
    seq = E.run_shots(b, W, c.M, c.H, c.a, c.rec, cfg, c.p, batch=1)
    for r, q in zip(results, seq):
        assert np.array_equal(r.gather, q.gather)


def test_segy_model_to_device_setup_to_gpu_shot(b, tmp_path):
    import oracle as O
    from tests.test_model_loader import _write_segy
    rng = np.random.default_rng(4)
    nx_in, nz_in = 96, 64
    vp = np.full((nx_in, nz_in), 2400.0, F32)
    vp[:, 30:] = 3100.0
    vp += F32(8.0) * rng.integers(0, 20, vp.shape).astype(F32)  # exact in SEG-Y IEEE floats
    path = str(tmp_path / "vp.segy")
    _write_segy(path, vp, 5)
    m = ML.load_model(path, dx=10.0)
    assert np.array_equal(m.vp, vp) and m.dx == F32(10.0)
    nbc, order, nt = 15, 8, 110
    dt = F32(0.4 * float(m.dx) / float(m.vp.max()))
    Md = E.init_medium_device(m.vp, m.vs, m.rho, m.dx, m.dz, nbc, order, b, free_surface=True)
    Mh = E.init_medium(m.vp, m.vs, m.rho, m.dx, m.dz, nbc, order, None, free_surface=True)
    H = E.init_habc(Md.nx, Md.nz, nbc, dt, m.dx, m.dz, I.julia_mean_f32(m.vp), None, weights=False)
    p = E.SimParams(dt, nt, m.dx, m.dz, order)
    a = I.get_fd_coefficients(order)
    src = E.Source(I.coord_to_index(F32(Md.x_max / F32(2)), Md.dx, Md.pad),
                   I.coord_to_index(F32(20.0), Md.dz, Md.pad), I.ricker_wavelet(F32(20.0), dt, nt),
                   "pressure")
    xr = np.arange(0, nx_in, 2, dtype=F32) * m.dx
    rec = E.setup_receivers(xr, np.full(len(xr), 10.0, F32), Md, type="vz")
    print(f"case SEG-Y model {nx_in}x{nz_in} -> device setup -> shot", flush=True)
    W = E.Wavefield(Md.nx, Md.nz, b)
    res = E.run_shot(b, W, Md, H, a, src, rec, p)
    w = [I.habc_weight_matrix(Mh.nx, Mh.nz, nbc, off) for off in (0.0, 0.5, 0.75)]
    r = O.OracleRun(nx=Mh.nx, nz=Mh.nz, pad=Mh.pad, nbc=nbc, fd_order=order, free_surface=True,
                    dt=p.dt, dtx=p.dtx, dtz=p.dtz, q=(H.qx, H.qz, H.qt_x, H.qt_z, H.qxt), a=a,
                    rho_vx=Mh.rho_vx, rho_vz=Mh.rho_vz, lam=Mh.lam, mu_txx=Mh.mu_txx,
                    mu_txz=Mh.mu_txz, w_vx=w[0], w_vz=w[1], w_tau=w[2], src_i=src.i, src_j=src.j,
                    src_type=0, wavelet=src.wavelet, rec_i=rec.i, rec_j=rec.j, rec_type=0, nt=nt,
                    dtype=F32, flags=O.FMA)
    r.run(1, nt)
    assert np.abs(res.gather).max() > 0
    assert np.array_equal(res.gather, r.gather_nt_by_nrec())
    for f in range(5):
        assert np.array_equal(W.download(f), r.field(f)), f
