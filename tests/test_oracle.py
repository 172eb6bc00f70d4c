"""Pins the CPU oracle (oracle/oracle.cpp) before results are checked against it.

- the reference's own output gather shot_1.bin (run.jl example);
- stencils: exact derivatives of polynomials (offsets, signs, coefficients);
- HABC: closed forms of single spikes (recurrence direction = program order);
- free surface rows, source injection, exact 2^k linearity, mirror symmetry,
  determinism, FP64 energy decay, FP32-vs-FP64 agreement.
"""
import os

import numpy as np
import pytest

import oracle as O
from fomo_julia_amd import elastic2d as E
from tests.helpers import Case, rel_l2, run_jl_case

F32 = np.float32
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "shot_1.bin")


def _oracle_from(d, flags=O.FMA, dtype=np.float32):
    M, H, p, src, rec = d["M"], d["H"], d["p"], d["src"], d["rec"]
    return O.OracleRun(nx=M.nx, nz=M.nz, pad=M.pad, nbc=H.nbc, fd_order=p.fd_order,
                       free_surface=M.is_free_surface, dt=p.dt, dtx=p.dtx, dtz=p.dtz,
                       q=(H.qx, H.qz, H.qt_x, H.qt_z, H.qxt), a=d["a"], rho_vx=M.rho_vx,
                       rho_vz=M.rho_vz, lam=M.lam, mu_txx=M.mu_txx, mu_txz=M.mu_txz,
                       w_vx=H.w_vx, w_vz=H.w_vz, w_tau=H.w_tau, src_i=src.i, src_j=src.j,
                       src_type=0, wavelet=src.wavelet, rec_i=rec.i, rec_j=rec.j, rec_type=0,
                       nt=d["nt"], dtype=dtype, flags=flags)


@pytest.mark.parametrize("flags", [O.FMA, 0, O.FMA | O.SNAPSHOT])
def test_oracle_reproduces_reference_shot_1(flags):
    d = run_jl_case()
    assert d["nt"] == 800 and (d["src"].i, d["src"].j) == (255, 60)
    gold = np.fromfile(GOLDEN, dtype="<f4").reshape(100, 800).T
    r = _oracle_from(d, flags)
    r.run(1, d["nt"])
    g = r.gather_nt_by_nrec()
    assert rel_l2(g, gold) < 5e-6
    # zero before the first arrival, as in the reference gather
    assert np.all(g[:2] == 0) and np.all(gold[:2] == 0)


def _poly_case(fd_order):
    c = Case(nx_in=40, nz_in=36, nbc=6, fd_order=fd_order, fs=False, nt=4)
    return c


@pytest.mark.parametrize("fd_order", [4, 8, 12])
def test_stencils_exact_on_polynomials(fd_order):
    c = _poly_case(fd_order)
    M = fd_order // 2
    nx, nz = c.M.nx, c.M.nz
    # exact (rational) coefficients so a degree-(2M-1) polynomial differentiates exactly
    from fomo_julia_amd.init import staggered_fd_coefficients_exact
    a = np.array([float(v) for v in staggered_fd_coefficients_exact(fd_order)])
    r = c.oracle(dtype=np.float64)
    r.a[:] = a
    x = np.arange(nx, dtype=np.float64)[:, None] * np.ones((1, nz))
    z = np.ones((nx, 1)) * np.arange(nz, dtype=np.float64)[None, :]
    deg = 2 * M - 1
    P = lambda t: ((t - 7.3) / 9.0) ** deg + 0.5 * ((t - 3.1) / 9.0) ** (deg - 1)
    dP = lambda t: deg / 9.0 * ((t - 7.3) / 9.0) ** (deg - 1) + 0.5 * (deg - 1) / 9.0 * ((t - 3.1) / 9.0) ** (deg - 2)
    # velocity stage: txx = P(x), txz = P(z) + P(x), tzz = P(z)
    r.field(2)[:] = P(x)
    r.field(4)[:] = P(z) + 2.0 * P(x)
    r.field(3)[:] = P(z)
    r.run(1, 1, O.STAGES["velocity"])
    sl = (slice(M, nx - M), slice(M, nz - M))
    bx = float(c.p.dtx) / r.med[0]
    bz = float(c.p.dtz) / r.med[0]
    want_vx = bx * dP(x - 0.5) + bz * dP(z - 0.5)
    np.testing.assert_allclose(r.field(0)[sl], want_vx[sl], rtol=1e-9, atol=1e-12)
    bx = float(c.p.dtx) / r.med[1]
    bz = float(c.p.dtz) / r.med[1]
    want_vz = bx * 2.0 * dP(x + 0.5) + bz * dP(z + 0.5)
    np.testing.assert_allclose(r.field(1)[sl], want_vz[sl], rtol=1e-9, atol=1e-12)
    # stress stage on fresh fields: vx = P(x) + 3P(z), vz = 2P(z) + P(x)
    r.fields[:] = 0
    r.field(0)[:] = P(x) + 3.0 * P(z)
    r.field(1)[:] = 2.0 * P(z) + P(x)
    r.run(1, 1, O.STAGES["stress"])
    lam, mu, mut = r.med[2], r.med[3], r.med[4]
    dvxdx = dP(x + 0.5)
    dvzdz = 2.0 * dP(z - 0.5)
    dvxdz = 3.0 * dP(z + 0.5)
    dvzdx = dP(x - 0.5)
    dtx, dtz = float(c.p.dtx), float(c.p.dtz)
    np.testing.assert_allclose(r.field(2)[sl], ((lam + 2 * mu) * dvxdx * dtx + lam * dvzdz * dtz)[sl], rtol=1e-9)
    np.testing.assert_allclose(r.field(3)[sl], (lam * dvxdx * dtx + (lam + 2 * mu) * dvzdz * dtz)[sl], rtol=1e-9)
    np.testing.assert_allclose(r.field(4)[sl], (mut * (dvxdz * dtz + dvzdx * dtx))[sl], rtol=1e-9)
    # outside the update region nothing changed
    assert np.all(r.field(2)[:M, :] == 0) and np.all(r.field(4)[:, nz - M:] == 0)


def test_habc_right_edge_is_a_forward_recurrence():
    c = Case(nx_in=30, nz_in=26, nbc=8, fd_order=4, fs=True, nt=2)
    nx, nz, n = c.M.nx, c.M.nz, 8
    r = c.oracle(dtype=np.float64)
    i0, j0 = nx - n, nz // 2  # first right-edge column (1-based)
    r.field(0)[i0 - 1, j0 - 1] = 1.0
    r.run(1, 1, O.STAGES["habc_v"])
    w = lambda i, j: float(c.H.w_vx[j - 1, i - 1])
    qx = float(c.H.qx)
    # f[i0] = w*1; f[i] = (1-w_i)(-qx f[i-1]) for i > i0 (program order along +x)
    want = w(i0, j0)
    assert r.field(0)[i0 - 1, j0 - 1] == pytest.approx(want, rel=1e-14)
    for i in range(i0 + 1, nx):
        want = (1 - w(i, j0)) * (-qx * want)
        assert r.field(0)[i - 1, j0 - 1] == pytest.approx(want, rel=1e-12)
    # the left edge is a Jacobi sweep: a spike at i affects only i-1
    r.fields[:] = 0
    r.field(0)[5 - 1, j0 - 1] = 1.0
    r.run(1, 1, O.STAGES["habc_v"])
    row = r.field(0)[:, j0 - 1]
    assert row[3] == pytest.approx((1 - w(4, j0)) * (-qx), rel=1e-13)
    assert row[4] == pytest.approx(w(5, j0), rel=1e-14)
    assert np.all(row[:3] == 0) and np.all(row[5:nx - n - 1] == 0)


def test_habc_bottom_edge_is_a_forward_recurrence_in_z():
    c = Case(nx_in=30, nz_in=26, nbc=8, fd_order=4, fs=True, nt=2)
    nx, nz, n = c.M.nx, c.M.nz, 8
    r = c.oracle(dtype=np.float64)
    i0, j0 = nx // 2, nz - n
    r.field(1)[i0 - 1, j0 - 1] = 1.0
    r.run(1, 1, O.STAGES["habc_v"])
    w = lambda i, j: float(c.H.w_vz[j - 1, i - 1])
    qz = float(c.H.qz)
    want = w(i0, j0)
    assert r.field(1)[i0 - 1, j0 - 1] == pytest.approx(want, rel=1e-14)
    for j in range(j0 + 1, nz):
        want = (1 - w(i0, j)) * (-qz * want)
        assert r.field(1)[i0 - 1, j - 1] == pytest.approx(want, rel=1e-12)


def test_habc_zero_q_is_pure_weighting():
    c = Case(nx_in=30, nz_in=22, nbc=7, fd_order=4, fs=False, nt=2)
    r = c.oracle(dtype=np.float64)
    r.desc.qx = r.desc.qz = r.desc.qt_x = r.desc.qt_z = r.desc.qxt = 0.0
    rng = np.random.default_rng(3)
    f0 = rng.standard_normal((c.M.nx, c.M.nz))
    r.field(2)[:] = f0
    r.run(1, 1, O.STAGES["habc_s"])
    w = c.H.w_tau.T.astype(np.float64)  # (nx, nz)
    nx, nz, n = c.M.nx, c.M.nz, 7
    reg = np.zeros((nx, nz), bool)
    reg[1:nx - 1, 1:nz - 1] = True
    reg[n + 1:nx - n - 1, n + 1:nz - n - 1] = False
    # without a free surface the left/right edge loops start at row 2
    # (boundary.jl:71,76,82), so the top corners are blended twice (edge, then corner)
    want = w * f0
    top = np.zeros((nx, nz), bool)
    top[1:n + 1, 1:n + 1] = True
    top[nx - n - 1:nx - 1, 1:n + 1] = True
    want[top] = (w * w * f0)[top]
    np.testing.assert_allclose(r.field(2)[reg], want[reg], rtol=1e-15)
    np.testing.assert_array_equal(r.field(2)[~reg], f0[~reg])


def test_free_surface_rows_and_source_injection():
    c = Case(nt=30)
    r = c.oracle()
    r.run(1, 30)
    pad = c.M.pad
    for f in (3, 4):
        assert np.all(r.field(f)[:, pad - 5:pad + 1] == 0)
    assert np.any(r.field(2)[:, pad - 5:pad + 1] != 0)
    r2 = c.oracle()
    r2.run(1, 1, O.STAGES["inject"])
    si, sj = c.src.i - 1, c.src.j - 1
    assert r2.field(2)[si, sj] == c.src.wavelet[0] and r2.field(3)[si, sj] == c.src.wavelet[0]
    assert np.count_nonzero(r2.fields) == 2


def test_exact_power_of_two_linearity_and_determinism():
    # Every operation of the step is linear and homogeneous in the wavefield,
    # so scaling the wavelet by 2^k scales every sample by exactly 2^k as long
    # as nothing is subnormal: exact in Float64 over this run.  In Float32 the
    # numerical wave's far leading edge goes subnormal and that rounding is
    # amplified when the wave arrives, so Float32 holds to ~1 ulp only.
    c = Case(nt=80)
    r1 = c.oracle(dtype=np.float64)
    r1.run(1, 80)
    assert np.abs(r1.fields[r1.fields != 0]).min() > 1e-300
    c.src.wavelet = c.src.wavelet * F32(8.0)
    r2 = c.oracle(dtype=np.float64)
    r2.run(1, 80)
    assert np.array_equal(r2.gather, r1.gather * 8.0)
    assert np.array_equal(r2.fields[:5], r1.fields[:5] * 8.0)
    s1 = c.oracle()
    s1.run(1, 80)
    c.src.wavelet = c.src.wavelet / F32(8.0)
    s0 = c.oracle()
    s0.run(1, 80)
    assert rel_l2(s1.gather, s0.gather * F32(8.0)) < 1e-6
    s2 = c.oracle()
    s2.run(1, 80)
    assert np.array_equal(s2.gather, s0.gather) and np.array_equal(s2.fields, s0.fields)


def test_mirror_symmetry_before_boundary_contact():
    # Homogeneous medium, odd nx, pressure source at the middle txx point
    # (x = i + 1/2): vz (also at i + 1/2) is mirror-symmetric, vx (at i) is
    # antisymmetric, as long as the wave has not reached the HABC strips
    # (vx points sit half a cell off the strips' index ranges).
    c = Case(nx_in=63, nz_in=40, nbc=10, fd_order=8, fs=True, layered=False, nt=40,
             src_xz=(F32(310.0), F32(30.0)))
    nx = c.M.nx
    assert nx % 2 == 1 and c.src.i - 1 == (nx - 1) // 2
    r = c.oracle(dtype=np.float64)
    r.run(1, 40)
    vz, vx = r.field(1), r.field(0)
    np.testing.assert_allclose(vz, vz[::-1], rtol=0, atol=1e-12 * np.abs(vz).max())
    np.testing.assert_allclose(vx[1:], -vx[1:][::-1], rtol=0, atol=1e-12 * np.abs(vx).max())


def test_fp64_energy_bounded_and_fp32_agrees():
    c = Case(nx_in=80, nz_in=60, nbc=15, nt=400)
    r64 = c.oracle(dtype=np.float64)
    energies = []
    for k in range(1, 401, 40):
        r64.run(k, 40)
        v = r64.fields[:2].astype(np.float64)
        energies.append(float((v ** 2).sum()))
    peak = int(np.argmax(energies))
    assert energies[-1] < energies[peak] and np.all(np.isfinite(energies))
    r32 = c.oracle()
    r32.run(1, 400)
    assert rel_l2(r32.gather, r64.gather) < 1e-4
    rsep = c.oracle(flags=0)
    rsep.run(1, 400)
    assert rel_l2(rsep.gather, r32.gather) < 1e-5


def _random_state(r, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    r.fields[:] = (scale * rng.standard_normal(r.fields.shape)).astype(r.fields.dtype)
    return r.fields.copy()


@pytest.mark.parametrize("src_type,comp,rho_idx", [("force_z", 1, 1), ("force_x", 0, 0)])
def test_force_injection_changes_one_cell_by_w_dt_over_rho(src_type, comp, rho_idx):
    """backend/cpu.jl:180-183: vz[si,sj] += wav*(dt/rho_vz[si,sj]) (force_z),
    vx[si,sj] += wav*(dt/rho_vx[si,sj]) (force_x), Float32 in that order.
    The layered, jittered medium makes rho_vx and rho_vz differ at the source
    cell and dx = 10 makes dt differ from dt/dx, so a wrong density array or a
    missing/extra 1/dx fails here."""
    c = Case(nx_in=40, nz_in=36, nbc=6, fd_order=4, fs=False, nt=8, src_type=src_type,
             src_xz=(F32(200.0), F32(170.0)))
    k = 3
    r = c.oracle()
    # a state of the increment's magnitude, so that the increment shows in Float32
    before = _random_state(r, 11, scale=1e-7)
    r.run(k, 1, O.STAGES["inject"])
    si, sj = c.src.i - 1, c.src.j - 1
    rho = (c.M.rho_vx, c.M.rho_vz)[rho_idx][si, sj]
    other = (c.M.rho_vx, c.M.rho_vz)[1 - rho_idx][si, sj]
    assert rho != other and c.p.dtx != c.p.dt
    w = c.src.wavelet[k - 1]
    assert w != 0
    want = F32(before[comp].T[si, sj] + F32(w * F32(c.p.dt / rho)))
    assert r.field(comp)[si, sj] == want
    changed = np.argwhere(r.fields != before)
    assert changed.tolist() == [[comp, sj, si]]


@pytest.mark.parametrize("src_type,comp,rho_idx", [("force_z", 1, 1), ("force_x", 0, 0)])
def test_force_source_adds_momentum_w_dt_per_step(src_type, comp, rho_idx):
    """Momentum balance of the staggered scheme, independent of how the source
    term is written: the stress-divergence stencil telescopes, so before the
    wave reaches the edges sum(rho_c * v_c) over the grid grows by exactly
    dt * w[k] per step of a force on component c (FP64 oracle).  A source
    scaled by the wrong density, by dt/dx, or applied to the other component
    breaks the balance."""
    nt = 24
    # the source sits on the last row above the layer interface, where rho_vz
    # (a 4-cell average across it) differs from rho_vx
    c = Case(nx_in=90, nz_in=90, nbc=10, fd_order=8, fs=False, nt=nt, src_type=src_type,
             src_xz=(F32(450.0), F32(440.0)), layered=True, f0=30.0)
    si, sj = c.src.i - 1, c.src.j - 1
    assert c.M.rho_vx[si, sj] != c.M.rho_vz[si, sj]
    r = c.oracle(dtype=np.float64)
    rho = np.asarray((c.M.rho_vx, c.M.rho_vz)[rho_idx], dtype=np.float64)
    w = c.src.wavelet.astype(np.float64)
    for k in range(1, nt + 1):
        r.run(k, 1)
        mom = float((rho * r.field(comp)).sum())
        want = float(c.p.dt) * float(w[:k].sum())
        assert mom == pytest.approx(want, rel=1e-9, abs=1e-9 * float(np.abs(w).sum()) * float(c.p.dt))
        # the other component gains no net momentum
        other = np.asarray((c.M.rho_vx, c.M.rho_vz)[1 - rho_idx], dtype=np.float64)
        assert abs(float((other * r.field(1 - comp)).sum())) < 1e-9 * max(abs(want), 1e-30) + 1e-18
    # the wave is still clear of the HABC strips (the balance would not hold
    # there): only the stencil's numerical tail reaches them
    n = c.nbc + 2
    peak = np.abs(r.fields[:5]).max()
    assert np.abs(r.fields[:5, :, :n]).max() < 1e-12 * peak
    assert np.abs(r.fields[:5, :n, :]).max() < 1e-12 * peak


def test_receivers_record_vz_vx_and_half_pressure():
    """kernels/source_receiver.jl:65-77 (CPU path): data[k, r] = vz, vx or
    (txx + tzz) * 0.5f0 at (i_r, j_r), Float32, after the step."""
    c = Case(nx_in=40, nz_in=36, nbc=6, fd_order=4, fs=False, nt=6)
    rng = np.random.default_rng(5)
    n = 24
    ri = rng.integers(1, c.M.nx + 1, n)
    rj = rng.integers(1, c.M.nz + 1, n)
    types = np.array([0, 1, 2] * (n // 3), dtype=np.int32)
    c.rec = E.Receivers(ri, rj, np.zeros((1, n), F32), type=[("vz", "vx", "p")[t] for t in types])
    r = c.oracle()
    st = _random_state(r, 9)
    k = 4
    r.run(k, 1, O.STAGES["record"])
    g = r.gather_nt_by_nrec()
    f = lambda comp, q: st[comp].T[ri[q] - 1, rj[q] - 1]
    for q in range(n):
        if types[q] == 0:
            want = f(1, q)
        elif types[q] == 1:
            want = f(0, q)
        else:
            want = F32(F32(f(2, q) + f(3, q)) * F32(0.5))
        assert g[k - 1, q] == want, q
    assert np.all(np.delete(g, k - 1, axis=0) == 0)
    # recording changes no field
    assert np.array_equal(r.fields, st)
    # p is the mean of txx and tzz, not their sum or vz: the three types differ here
    q = int(np.argmax(types == 2))
    assert g[k - 1, q] != f(1, q) and g[k - 1, q] != F32(f(2, q) + f(3, q))
