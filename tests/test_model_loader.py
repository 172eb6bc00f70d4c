"""Velocity-model ingestion (model_loader.jl): formats, layouts, estimates.

The SEG-Y reader is checked on files written here from the SEG-Y rev 1
layout (textual header, binary header, 240-byte trace headers) and on IBM
float words whose values are fixed by the format's definition (the
classic 0xC276A000 = -118.625 among them)."""
import struct

import numpy as np
import pytest

from fomo_julia_amd import model_loader as ML

F32 = np.float32


def _ibm_word(v: float) -> int:
    """IBM single precision word of v (v a short binary fraction, so exact)."""
    if v == 0:
        return 0
    s = 0x80000000 if v < 0 else 0
    a = abs(v)
    e = 0
    while a >= 1.0:
        a /= 16.0
        e += 1
    while a < 1.0 / 16.0:
        a *= 16.0
        e -= 1
    frac = int(round(a * (1 << 24)))
    assert frac == a * (1 << 24)
    return s | ((e + 64) << 24) | frac


def _write_segy(path, data_ns_by_ntr, fmt, n_ext=0):
    ns, ntr = data_ns_by_ntr.shape
    head = bytearray(3600)
    head[0:3200] = b" " * 3200
    head[3216:3218] = struct.pack(">H", 1000)  # sample interval (unused)
    head[3220:3222] = struct.pack(">H", ns)
    head[3224:3226] = struct.pack(">H", fmt)
    head[3504:3506] = struct.pack(">h", n_ext)
    body = bytearray()
    body += b" " * (3200 * n_ext)
    for t in range(ntr):
        th = bytearray(240)
        th[114:116] = struct.pack(">H", ns)
        col = data_ns_by_ntr[:, t]
        if fmt == 1:
            smp = b"".join(struct.pack(">I", _ibm_word(float(v))) for v in col)
        elif fmt == 5:
            smp = col.astype(">f4").tobytes()
        elif fmt == 3:
            smp = col.astype(">i2").tobytes()
        elif fmt == 2:
            smp = col.astype(">i4").tobytes()
        else:
            smp = col.astype("i1").tobytes()
        body += th + smp
    with open(path, "wb") as f:
        f.write(bytes(head) + bytes(body))


def test_ibm_float_words():
    words = np.array([0x42640000, 0xC276A000, 0x41100000, 0x00000000, 0x40800000,
                      0x3F800000, 0x7FFFFFFF & 0x43100000], dtype=np.uint32)
    want = np.array([100.0, -118.625, 1.0, 0.0, 0.5, 0.03125, 256.0], F32)
    assert np.array_equal(ML.ibm_to_float32(words), want)
    for v in (1500.0, 4500.0, -2.25, 3000.5, 1.0 / 1024):
        assert ML.ibm_to_float32(np.array([_ibm_word(v)], np.uint32))[0] == F32(v)


@pytest.mark.parametrize("fmt", [1, 2, 3, 5, 8])
def test_segy_roundtrip(tmp_path, fmt):
    rng = np.random.default_rng(fmt)
    if fmt in (1, 5):
        data = (1500 + 8 * rng.integers(0, 400, (13, 9))).astype(F32)  # exact in IBM too
        data[0, 0] = -2.5
    else:
        data = rng.integers(-100, 100, (13, 9)).astype(F32)
    p = tmp_path / "vp.sgy"
    _write_segy(p, data, fmt, n_ext=1 if fmt == 5 else 0)
    got = ML.read_segy(str(p))
    assert got.shape == (13, 9) and got.dtype == F32
    assert np.array_equal(got, data)
    with pytest.raises(ValueError):  # Gardner's vp^0.25 of a negative sample
        ML.load_model(str(p), dx=12.5)
    q = tmp_path / "pos.segy"
    _write_segy(q, np.abs(data), fmt)
    m = ML.load_model(str(q), dx=12.5)
    assert (m.nx, m.nz) == (13, 9) and m.dx == F32(12.5) and m.dz == F32(12.5)
    assert np.array_equal(m.rho, ML.gardner_rho(np.abs(data)))


def test_segy_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.segy"
    _write_segy(p, np.ones((4, 3), F32), 5)
    with open(p, "ab") as f:
        f.write(b"\0" * 7)  # not a whole trace
    with pytest.raises(ValueError):
        ML.read_segy(str(p))
    q = tmp_path / "fmt.segy"
    _write_segy(q, np.ones((4, 3), F32), 5)
    raw = bytearray(q.read_bytes())
    raw[3224:3226] = struct.pack(">H", 4)
    q.write_bytes(bytes(raw))
    with pytest.raises(ValueError):
        ML.read_segy(str(q))


def test_binary_layouts(tmp_path):
    nx, nz = 5, 3
    vp = np.arange(nx * nz, dtype=F32).reshape((nx, nz), order="F") + 1000
    col = tmp_path / "vp.bin"
    vp.ravel(order="F").tofile(col)  # Julia's write(io, vp)
    m = ML.load_model(str(col), nx=nx, nz=nz, dx=10.0)
    assert np.array_equal(m.vp, vp) and (m.nx, m.nz) == (nx, nz)
    row = tmp_path / "vp_row.bin"
    vp.T.ravel(order="F").tofile(row)  # (nz, nx) column-major = vp row-major
    m2 = ML.load_model(str(row), nx=nx, nz=nz, order="row_major")
    assert np.array_equal(m2.vp, vp)
    assert np.array_equal(ML.load_model(str(col), nx=nx, nz=nz, single=True), vp)
    with pytest.raises(ValueError):
        ML.load_model(str(col))
    with pytest.raises(EOFError):
        ML.load_model(str(col), nx=nx, nz=nz + 1)


def test_estimates_closed_forms():
    vp = np.array([[4096.0, 81.0], [1.0, 0.0]], F32)
    assert np.array_equal(ML.gardner_rho(vp), np.array([[2480.0, 930.0], [310.0, 0.0]], F32))
    vs = ML.estimate_vs(np.array([[0.0, 1.73]], F32))
    assert vs[0, 0] == 0.0 and vs[0, 1] == 1.0


def test_npy_mat_npz_and_files(tmp_path):
    rng = np.random.default_rng(0)
    vp = (2000 + 1000 * rng.random((6, 4))).astype(F32)
    vs = (vp / 2).astype(F32)
    np.save(tmp_path / "vp.npy", vp)
    np.save(tmp_path / "vs.npy", vs)
    m = ML.load_model(str(tmp_path / "vp.npy"), dx=5.0, vs=str(tmp_path / "vs.npy"))
    assert np.array_equal(m.vs, vs) and np.array_equal(m.rho, ML.gardner_rho(vp))
    import scipy.io
    scipy.io.savemat(tmp_path / "m.mat", {"Vp": vp, "rho": vs})
    mm = ML.load_model(str(tmp_path / "m.mat"), dx=2.0)
    assert np.array_equal(mm.vp, vp) and np.array_equal(mm.rho, vs)
    assert np.array_equal(mm.vs, ML.estimate_vs(vp))
    ML.save_model(str(tmp_path / "out.npz"), mm)
    back = ML.load_model(str(tmp_path / "out.npz"))
    assert np.array_equal(back.vp, vp) and back.dx == F32(2.0)
    vp.ravel(order="F").tofile(tmp_path / "rho.bin")
    mf = ML.load_model_files(vp=str(tmp_path / "vp.npy"), rho=str(tmp_path / "rho.bin"), dx=10.0,
                             nx=6, nz=4, name="mixed")
    assert np.array_equal(mf.rho, vp) and mf.name == "mixed" and mf.dz == F32(10.0)
    for bad in ("x.jld2", "x.h5", "x.txt"):
        with pytest.raises(ValueError):
            ML.load_model(str(tmp_path / bad))


def test_model_info_text():
    m = ML.velocity_model(np.full((3, 2), 1500.0), np.full((3, 2), 0.0), np.full((3, 2), 1000.0),
                          10.0, 5.0, name="tiny")
    t = ML.model_info(m)
    assert "Model: tiny" in t and "Grid:     3 × 2" in t
    assert "Spacing:  dx=10.0m, dz=5.0m" in t and "Size:     30.0m × 10.0m" in t
    assert "Vp:   1500.0 - 1500.0 m/s" in t
    with pytest.raises(ValueError):
        ML.velocity_model(np.zeros((3, 2)), np.zeros((2, 3)), np.zeros((3, 2)), 1.0, 1.0)
