"""Reduced-precision storage (BF16, FP16, FP16 coefficients with FP32 fields,
F24 fields with FP32 coefficients).

The reference is FP32 only, so parity here has two parts:
  * the kernels against the oracle's sequential restatement of the storage
    contract (coefficients dither-quantized once, fields rounded to nearest
    even whenever a stage writes them): bit for bit;
  * the accuracy of each storage mode against the oracle's Float64 run of the
    same scheme: relative L2 and peak-amplitude error per receiver component
    (vx, vz, pressure), bounds written below.

The rounding functions the oracle uses are pinned against numpy's float16 and
torch's bfloat16 conversions (both round to nearest, ties to even), and F24's
against the integer rounding of the FP32 bit pattern and float64 arithmetic."""
import numpy as np
import pytest

import oracle as O
from fomo_julia_amd import elastic2d as E
from tests.helpers import C1Case, Case, rel_l2

F32 = np.float32

# Relative L2 bounds of the C1 gathers against Float64 (measured values in the
# comments; nt = 1000).
C1_BOUNDS = {"f32": 1e-5,    # ~1.0e-6
             "f16c": 1e-3,   # ~8.2e-4 (vx), 5.3e-4 (vz), 5.6e-4 (p)
             "f16": 4e-3,    # ~2.7e-3 (vx), 1.9e-3 (vz), 3.0e-3 (p)
             "bf16": 3e-2,   # ~2.2e-2 (vx), 1.7e-2 (vz), 2.4e-2 (p)
             "f24": 3e-4}    # ~8.1e-5 (vx), 5.7e-5 (vz), 9.2e-5 (p) (oracle emulation)


def _values(rng, n):
    """Floats across FP16's normal and subnormal ranges, plus exact ties."""
    mags = np.exp2(rng.uniform(-26, 15.9, n))
    x = (mags * rng.choice([-1.0, 1.0], n)).astype(F32)
    # halfway points between neighbouring FP16 values (ties go to even)
    h = np.float16(rng.uniform(-1000, 1000, 200)).astype(np.float64)
    nxt = np.nextafter(np.float16(h), np.float16(np.inf)).astype(np.float64)
    ties = ((h + nxt) / 2).astype(F32)
    return np.concatenate([x, ties, np.array([0.0, -0.0, 6.1e-5, 5.96e-8, 65504.0], F32)])


def test_f16_rounding_matches_numpy():
    O.build()
    rng = np.random.default_rng(1)
    for scale in (1.0, 2.0 ** 12, 2.0 ** -20):
        for x in _values(rng, 3000):
            if abs(float(x) * scale) >= 65504.0:  # beyond FP16's range: not stored
                continue
            want = np.float64(np.float16(np.float32(x * scale))) / scale
            got = O.round_storage("f16", float(x), scale)
            assert got == np.float32(want), (x, scale, got, want)


def test_bf16_rounding_matches_torch_and_bits():
    torch = pytest.importorskip("torch")
    O.build()
    rng = np.random.default_rng(2)
    x = np.concatenate([_values(rng, 3000),
                        (np.exp2(rng.uniform(-140, 100, 2000)) *
                         rng.choice([-1.0, 1.0], 2000)).astype(F32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    u = x.view(np.uint32).astype(np.uint64)
    bits = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(F32)
    assert np.array_equal(want, bits)
    got = np.array([O.round_storage("bf16", float(v)) for v in x], F32)
    assert np.array_equal(got, want)


def _f24_bits(x):
    """x rounded to a 16-bit significand by the integer rounding of its FP32
    bit pattern (add 0x7F plus the kept lowest bit, clear the low byte)."""
    u = np.asarray(x, F32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7F + ((u >> 8) & 1)) & 0xFFFFFF00).astype(np.uint32).view(F32)


def test_f24_rounding_matches_bits_and_float64():
    """F24 (16-bit significand, FP32's exponent range): the oracle's rounding
    equals the bit-pattern rounding the kernels use, and both equal rounding
    in float64 to the grid 2^(e-15) (|x| in [2^(e-1), 2^e)), nearest, ties to
    even; exact ties, subnormals, the largest finite values and FP32's
    extremes included."""
    O.build()
    rng = np.random.default_rng(3)
    mags = np.exp2(rng.uniform(-149, 127.9, 6000))
    x = (mags * rng.choice([-1.0, 1.0], mags.size)).astype(F32)
    base = (rng.integers(1 << 15, 1 << 16, 800).astype(np.float64) *
            np.exp2(rng.integers(-140, 100, 800) - 15.0))
    ties = np.concatenate([base + np.exp2(np.floor(np.log2(base)) - 16),     # halfway up
                           base - np.exp2(np.floor(np.log2(base)) - 16)]).astype(F32)
    special = np.array([0.0, -0.0, 1.0, 1.0 + 2.0 ** -16, 1.0 + 3 * 2.0 ** -16, 2.0 ** -149,
                        2.0 ** -126, 3.4e38, -3.4e38, np.finfo(F32).max], F32)
    x = np.concatenate([x, ties, special])
    got = np.array([O.round_storage("f24", float(v)) for v in x], F32)
    assert np.array_equal(got.view(np.uint32), _f24_bits(x).view(np.uint32))
    want = []
    for v in x.astype(np.float64):
        if v == 0:
            want.append(v)
            continue
        e = int(np.floor(np.log2(abs(v)))) + 1  # |v| in [2^(e-1), 2^e)
        sp = 2.0 ** (max(e - 1, -126) - 15)
        want.append(np.round(v / sp) * sp)  # numpy rounds half to even
    with np.errstate(over="ignore"):
        want = np.array(want).astype(F32)
    assert np.array_equal(got, want)
    # ties went to the even neighbour, and every result has its low byte clear
    assert float(O.round_storage("f24", 1.0 + 2.0 ** -16)) == 1.0
    assert float(O.round_storage("f24", 1.0 + 3 * 2.0 ** -16)) == 1.0 + 2.0 ** -14
    assert not np.any(got.view(np.uint32) & 0xFF)


@pytest.fixture(scope="module")
def c1_f64():
    O.build()
    c = C1Case()
    r64 = c.oracle(np.float64)
    r64.run(1, c.nt)
    return c, c.split(r64.gather_nt_by_nrec())


def test_emulated_storage_accuracy_c1_cpu(c1_f64):
    """The oracle's storage emulation on the homogeneous 200 x 200 case: FP16
    coefficients with FP32 fields stay within 1e-3 of Float64 because the
    coefficients are dithered (rounding them to nearest instead gives a
    systematic ~3e-3 error in a homogeneous medium)."""
    c, g64 = c1_f64
    r = c.oracle(storage="f16c")
    r.run(1, c.nt)
    g = c.split(r.gather_nt_by_nrec())
    for t in g:
        assert rel_l2(g[t], g64[t]) < C1_BOUNDS["f16c"], t


def test_block_format_rounding():
    """The block floating-point segment (storage "bfp16", a design study that
    no kernel implements) against hand-computed values: the largest magnitude
    3 puts the spacing at 2^-13, so 2^-20 vanishes, ties go to even, and a
    mantissa that would round to 2^15 is clamped to 32767."""
    O.build()
    F = np.float32
    assert np.array_equal(O.round_block([3.0, 1.0, 2.0**-20, -1.5]), np.array([3.0, 1.0, 0.0, -1.5], F))
    # 1 + 2^-14 is 8192.5 spacings (tie: to 8192), 1 + 3 * 2^-14 is 8193.5 (to 8194)
    got = O.round_block([3.0, 1 + 2.0**-14, 1 + 3 * 2.0**-14, -(1 + 2.0**-14)])
    assert np.array_equal(got, np.array([3.0, 1.0, 1 + 2.0**-12, -1.0], F))
    assert O.round_block([4 - 2.0**-20])[0] == F(4 - 2.0**-13)  # 32768 clamped to 32767
    assert np.array_equal(O.round_block(np.zeros(64)), np.zeros(64, F))
    rng = np.random.default_rng(11)
    for _ in range(50):
        x = (rng.standard_normal(64) * 10.0 ** rng.uniform(-8, 8)).astype(F)
        y = O.round_block(x)
        e = np.frexp(np.abs(x.astype(np.float64)).max())[1]
        spacing = 2.0 ** (e - 15)
        assert np.abs(y.astype(np.float64) - x).max() <= spacing / 2
        assert np.array_equal(np.rint(y / spacing) * spacing, y)  # on the segment's grid
        assert np.array_equal(O.round_block(y), y)                # idempotent


@pytest.mark.parametrize("storage", ["bfp16", "bfp16c"])
def test_block_format_accuracy_c1_cpu(c1_f64, storage):
    """The design study's 16-bit field format meets 1e-3 of Float64 on the
    homogeneous 200 x 200 case, where FP16 fields give 2.7e-3 and BF16 2.2e-2:
    about 5.1e-4 / 3.5e-4 / 6.0e-4 for vx / vz / p with FP32 coefficients
    ("bfp16"), 9.6e-4 / 6.5e-4 / 8.1e-4 with dithered FP16 coefficients
    ("bfp16c")."""
    c, g64 = c1_f64
    r = c.oracle(storage=storage)
    r.run(1, c.nt)
    g = c.split(r.gather_nt_by_nrec())
    for t in g:
        assert rel_l2(g[t], g64[t]) < 1e-3, t


# ---------------------------------------------------------------- GPU
gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def b():
    bk = E.backend("mi355x")
    bk.set_tiling(1, 0, 16)
    bk.set_step_mode(1)
    return bk


def _shot(b, c):
    dm = E._model(b, c.M, c.H, c.a, c.p)
    return E._shot_for(dm, c.src, c.rec, c.nt)


def _run_and_emulate(b, c, storage):
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    res = E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p)
    scales = _shot(b, c).field_scales()
    r = c.oracle(storage=storage, fscale=[float(s) for s in scales])
    r.run(1, c.nt)
    return res, W, r, scales


def _assert_same(res, W, r, what):
    g, want = res.gather, r.gather_nt_by_nrec()
    if not np.array_equal(g, want):
        bad = np.argwhere(g != want)
        k, q = bad[0]
        raise AssertionError(f"{what}: gather differs in {len(bad)} samples, first (k={k}, r={q}) "
                             f"got {g[k, q]!r} want {want[k, q]!r}")
    for f in range(5):
        got, ref = W.download(f), r.field(f)
        if not np.array_equal(got, ref):
            bad = np.argwhere(got != ref)
            i, j = bad[0]
            raise AssertionError(f"{what}: field {f} differs in {len(bad)} cells, first (i={i}, j={j}) "
                                 f"got {got[i, j]!r} want {ref[i, j]!r}")


@gpu
@pytest.mark.parametrize("storage", ["bf16", "f16", "f16c", "f24"])
@pytest.mark.parametrize("fd_order,fs,src_type", [(8, True, "pressure"), (4, False, "force_z"),
                                                  (12, True, "force_x")])
def test_reduced_storage_bit_exact_vs_emulation(b, storage, fd_order, fs, src_type):
    c = Case(fd_order=fd_order, fs=fs, src_type=src_type, nt=120, water=fs, storage=storage)
    print(f"case storage={storage} order={fd_order} fs={fs} src={src_type}", flush=True)
    res, W, r, scales = _run_and_emulate(b, c, storage)
    assert all(float(s) == 2.0 ** np.round(np.log2(float(s))) for s in scales)
    assert np.abs(res.gather).max() > 0
    _assert_same(res, W, r, storage)


@gpu
@pytest.mark.parametrize("storage", ["bf16", "f16", "f24"])
@pytest.mark.parametrize("cpt,tz", [(2, 17), (4, 0)])
def test_reduced_storage_tilings(b, storage, cpt, tz):
    c = Case(nx_in=150, nz_in=90, nbc=20, nt=71, storage=storage)
    print(f"case storage={storage} cpt={cpt} tz={tz}", flush=True)
    b.set_tiling(cpt, tz, 6)
    try:
        res, W, r, _ = _run_and_emulate(b, c, storage)
    finally:
        b.set_tiling(1, 0, 16)
    _assert_same(res, W, r, f"{storage} cpt={cpt}")


@gpu
@pytest.mark.parametrize("storage", ["bf16", "f16", "f16c", "f24"])
@pytest.mark.parametrize("cpt", [2, 4])
def test_reduced_storage_order12_wide_tiles(b, storage, cpt):
    """fd_order 12 with 2 and 4 columns per lane in each reduced storage: the
    fully unrolled z-march of the widest reduced-storage kernels."""
    c = Case(nx_in=150, nz_in=70, nbc=20, fd_order=12, nt=60, storage=storage)
    print(f"case storage={storage} order=12 cpt={cpt}", flush=True)
    b.set_tiling(cpt, 0, 6)
    try:
        res, W, r, _ = _run_and_emulate(b, c, storage)
    finally:
        b.set_tiling(1, 0, 16)
    _assert_same(res, W, r, f"{storage} order 12 cpt={cpt}")


@gpu
@pytest.mark.parametrize("storage", ["f32", "f16c", "f16", "bf16", "f24"])
def test_c1_accuracy_vs_float64(b, storage):
    """C1 with each storage: bit-exact against the emulation and within the
    stated relative L2 of the Float64 oracle, per component."""
    c = C1Case(storage=storage)
    print(f"case C1 storage={storage}", flush=True)
    res, W, r, _ = _run_and_emulate(b, c, storage)
    if storage != "f32":
        _assert_same(res, W, r, storage)
    r64 = c.oracle(np.float64)
    r64.run(1, c.nt)
    g64 = c.split(r64.gather_nt_by_nrec())
    g = c.split(res.gather)
    for t in g:
        e = rel_l2(g[t], g64[t])
        peak = float(np.abs(g[t] - g64[t]).max() / np.abs(g64[t]).max())
        print(f"  {storage} {t}: rel L2 {e:.3e}, peak {peak:.3e}", flush=True)
        assert e < C1_BOUNDS[storage], (t, e)


@gpu
@pytest.mark.parametrize("storage,mode,nts", [("bf16", 1, (50, 50, 50)), ("f24", 1, (50, 50, 50)),
                                              ("f24", 2, (50, 50, 50)), ("bf16", 1, (50, 37, 44)),
                                              ("f24", 1, (31, 50, 44))])
def test_reduced_batch_uneven_matches_single(b, storage, mode, nts):
    """A batch of uneven grids equals each shot run alone (and, for F24, the
    oracle's emulation), in the whole-grid and the split step; whole-grid
    batches step as two halves on two streams, and shots that end before
    the batch's run (nts) keep the state of their own last step."""
    cases = [Case(nx_in=90, nz_in=60, nt=nts[0], storage=storage, seed=1),
             Case(nx_in=130, nz_in=45, nt=nts[1], storage=storage, seed=2, src_type="force_z"),
             Case(nx_in=200, nz_in=110, nt=nts[2], storage=storage, seed=3, src_type="force_x")]
    print(f"case uneven batch {storage} mode {mode}", flush=True)
    b.set_step_mode(mode)
    b.set_split(7, 1)
    singles = []
    for c in cases:
        W = E.Wavefield(c.M.nx, c.M.nz, b)
        singles.append(E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p).gather)
    lib = b.lib
    import ctypes
    hs = [(E._model(b, c.M, c.H, c.a, c.p), c) for c in cases]
    ws, shots, models, keep = [], [], [], []
    for dm, c in hs:
        W = E.Wavefield(c.M.nx, c.M.nz, b)
        W._bind(dm)
        E.reset(b, W)
        ds = E._DeviceShot(dm, c.src, c.rec, c.nt)
        keep += [W, ds]
        ws.append(W._h)
        shots.append(ds.h)
        models.append(dm.h)
    arr = lambda v: (ctypes.c_void_p * len(v))(*[x.value for x in v])
    hb = ctypes.c_void_p()
    b.check(lib.fomo_batch_create(b.ctx, len(cases), arr(models), arr(ws), arr(shots),
                                  ctypes.byref(hb)), "batch_create")
    try:
        b.check(lib.fomo_batch_run(hb, 1, max(nts)), "batch_run")
        b.synchronize()
    finally:
        lib.fomo_batch_destroy(hb)
        b.set_step_mode(1)
        b.set_split(16, 0)
    for q, want in enumerate(singles):
        assert np.array_equal(keep[2 * q + 1].gather(), want)
    if len(set(nts)) > 1:  # each shot's final state against its solo run's
        for q, c in enumerate(cases):
            W = E.Wavefield(c.M.nx, c.M.nz, b)
            E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p)
            for f in range(5):
                assert np.array_equal(keep[2 * q].download(f), W.download(f)), (q, f)
    if storage == "f24":
        for q, c in enumerate(cases):
            r = c.oracle(storage="f24")
            r.run(1, c.nt)
            assert np.array_equal(keep[2 * q + 1].gather(), r.gather_nt_by_nrec()), q
            same = [np.array_equal(keep[2 * q].download(f), r.field(f)) for f in range(5)]
            assert all(same), (q, same)


@gpu
def test_reduced_storage_rejections(b):
    c = Case(nt=20, storage="bf16")
    W = E.Wavefield(c.M.nx, c.M.nz, b)
    E.run_shot(b, W, c.M, c.H, c.a, c.src, c.rec, c.p)
    with pytest.raises(E._lib.FomoError):
        E.update_velocity(b, W)
    with pytest.raises(E._lib.FomoError):
        W.download("vx_old")
    odd = Case(nt=20, fd_order=6, storage="f16")
    W2 = E.Wavefield(odd.M.nx, odd.M.nz, b)
    with pytest.raises(E._lib.FomoError):
        E.run_shot(b, W2, odd.M, odd.H, odd.a, odd.src, odd.rec, odd.p)
    # an FP16 state continued with a shot whose scales differ
    c16 = Case(nt=20, storage="f16")
    W3 = E.Wavefield(c16.M.nx, c16.M.nz, b)
    E.run_shot(b, W3, c16.M, c16.H, c16.a, c16.src, c16.rec, c16.p)
    loud = E.Source(c16.src.i, c16.src.j, c16.src.wavelet * F32(1024.0), c16.src.type)
    with pytest.raises(E._lib.FomoError):
        E.run_time_loop(b, W3, c16.M, c16.H, c16.a, loud, c16.rec, c16.p)
