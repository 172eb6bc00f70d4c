"""Known-answer tests of the host-side setup (src/utils/init.jl restated in
fomo_julia_amd/init.py): closed forms, exact rationals and the index rules."""
from fractions import Fraction

import numpy as np
import pytest

from fomo_julia_amd import init as I
from fomo_julia_amd import elastic2d as E

F32 = np.float32


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_fd_table_matches_exact_rationals(order):
    exact = I.staggered_fd_coefficients_exact(order)
    tab = I.get_fd_coefficients(order)
    assert len(tab) == order // 2
    for e, t in zip(exact, tab):
        assert abs(float(t) - float(e)) <= 3.4e-10 + 6e-8 * abs(float(e))


def test_fd_order10_keeps_reference_table_value():
    # init.jl:18 tables a1 = 1.2115478515625 = 19850/16384; the exact value is 19845/16384.
    exact = I.staggered_fd_coefficients_exact(10)
    assert exact[0] == Fraction(19845, 16384)
    assert float(I.get_fd_coefficients(10)[0]) == 19850 / 16384
    s = sum((2 * l + 1) * float(c) for l, c in enumerate(I.get_fd_coefficients(10)))
    assert abs(s - 1.000305) < 2e-6


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10, 12, 14, 16])
def test_exact_coefficients_solve_taylor_system(order):
    # independent check: solve sum_m a_m (2m-1)^(2k-1) = delta_{k1} numerically
    M = order // 2
    A = np.array([[(2 * m - 1) ** (2 * k - 1) for m in range(1, M + 1)] for k in range(1, M + 1)],
                 dtype=np.float64)
    rhs = np.zeros(M)
    rhs[0] = 1.0
    sol = np.linalg.solve(A, rhs)
    exact = np.array([float(c) for c in I.staggered_fd_coefficients_exact(order)])
    np.testing.assert_allclose(exact, sol, rtol=1e-9, atol=1e-12)


def test_unsupported_order_raises():
    with pytest.raises(ValueError):
        I.get_fd_coefficients(7)
    with pytest.raises(ValueError):
        I.get_fd_coefficients(18)


def test_ricker_peak_and_symmetry():
    dt = F32(1e-3)
    f0 = F32(15.0)
    w = I.ricker_wavelet(f0, dt, 400)
    # t0 = 1/f0 = 66.67 steps: peak sample is the one nearest t0, value close to 1
    k = int(np.argmax(w))
    assert k == 67 and abs(w[k] - 1.0) < 2e-3
    # closed form in float64 at the reference's promotion chain
    for i in (0, 10, 66, 67, 120, 399):
        tau = np.float64(F32(i) * dt) - 1.0 / np.float64(f0)
        x = np.float64(F32(np.pi) * f0) * tau
        assert w[i] == F32((1 - 2 * x * x) * np.exp(-x * x))
    w2 = I.ricker_wavelet(f0, dt, 400, t0=1.2 / 15.0)
    assert int(np.argmax(w2)) == 80


def test_coord_to_index_ties_to_even():
    assert I.coord_to_index(F32(25.0), F32(10.0), 0) == 3   # 2.5 -> 2, +1
    assert I.coord_to_index(F32(35.0), F32(10.0), 0) == 5   # 3.5 -> 4, +1
    assert I.coord_to_index(F32(1995.0), F32(10.0), 54) == 255  # run.jl source x
    assert I.coord_to_index(F32(50.0), F32(10.0), 54) == 60
    # config 1: source (1000, 20) -> (155, 57); receivers x = 0..1990 at z = 10 -> i 55..254, j 56
    assert I.coord_to_index(F32(1000.0), F32(10.0), 54) == 155
    assert I.coord_to_index(F32(20.0), F32(10.0), 54) == 57
    i, j = I.receiver_indices(np.arange(200, dtype=F32) * F32(10), np.full(200, 10, F32),
                              F32(10), F32(10), 54)
    assert list(i) == list(range(55, 255)) and set(j) == {56}


def test_coord_to_index_brute_force_against_fractions():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x = F32(rng.uniform(0, 20000))
        dx = F32(rng.choice([10.0, 12.5, 5.0, 7.5]))
        q = Fraction(float(F32(x / dx)))  # Float32 quotient, as Julia's x/dx
        fl = q.numerator // q.denominator
        r = q - fl
        want = fl + 1 if r > Fraction(1, 2) else fl if r < Fraction(1, 2) else (fl if fl % 2 == 0 else fl + 1)
        assert I.coord_to_index(x, dx, 7) == want + 8


def test_run_jl_receiver_and_source_indices():
    x_rec = np.array([500 + k * 3000 / 99 for k in range(100)]).astype(F32)
    i, j = I.receiver_indices(x_rec, np.full(100, 100, F32), F32(10), F32(10), 54)
    assert i[0] == 105 and i[-1] == 405 and i[1] == 108 and set(j) == {65}


def test_habc_constants_config1():
    qx, qz, qt_x, qt_z, qxt = I.habc_coefficients(F32(3000), F32(1e-3), F32(10), F32(10))
    assert qx == F32(0.39860135) and qz == qx
    assert qt_x == F32(-0.58041966) and qxt == F32(-0.81818175)


def test_weight_ramp_orientation_and_values():
    nx, nz, nbc = 30, 20, 5
    w = I.habc_weight_matrix(nx, nz, nbc, 0.5)
    assert w.shape == (nz, nx)
    # (1-based) i=2, j=10: dist = 1 -> (1-0.5)/5 = 0.1
    assert w[9, 1] == F32(0.1)
    # i=15, j=2 (0-based row 1): dist = 1
    assert w[1, 14] == F32(0.1)
    assert w[10, 15] == 1.0 and w[0, 0] == 0.0
    r = I.habc_ramp(nbc, 0.75)
    assert r[0] == 0 and r[1] == F32(0.05) and r[-1] == 1.0


def test_staggered_params_closed_form():
    vp = np.array([[3000, 3000], [4000, 2000]], F32)
    vs = np.array([[1500, 0], [2000, 1000]], F32)
    rho = np.array([[2000, 2100], [2200, 2300]], F32)
    lam, mu, mu_txz, rvx, rvz = I._compute_staggered_params(vp, vs, rho)
    assert mu[0, 0] == F32(2000 * 1500 * 1500)
    assert lam[1, 0] == F32(F32(2200) * F32(4000 * 4000)) - F32(2) * mu[1, 0]
    assert rvz[0, 0] == F32(0.25) * (F32(2000) + F32(2200) + F32(2100) + F32(2300))
    assert rvz[1, 1] == rvz[0, 1] == rvz[0, 0]
    assert mu_txz[0, 0] == 0.0  # vs = 0 neighbour: harmonic mean -> 0
    assert mu_txz[1, 0] == F32(2) / (F32(1) / mu[1, 0] + F32(1) / mu[1, 1])
    assert np.array_equal(rvx, rho)


def test_pad_array_edge_replication():
    a = np.arange(6, dtype=F32).reshape(3, 2)
    p = I._pad_array(a, 2)
    assert p.shape == (7, 6)
    assert np.array_equal(p[2:5, 2:4], a)
    assert p[0, 0] == a[0, 0] and p[6, 5] == a[2, 1] and p[0, 5] == a[0, 1]


def test_julia_mean_run_jl_model():
    vp = np.full((400, 200), 3000, F32)
    vp[:, 99:] = 4000
    assert I.julia_mean_f32(vp) == F32(3505.0)


def test_init_medium_shapes_and_padding():
    M = E.init_medium(np.full((10, 8), 2000, F32), np.full((10, 8), 1000, F32),
                      np.full((10, 8), 1800, F32), 5.0, 5.0, 6, 4, None)
    assert (M.nx, M.nz, M.pad, M.M) == (10 + 16, 8 + 16, 8, 2)
    assert M.x_max == F32(45.0) and M.rho_vx.shape == (26, 24)


def test_simparams_and_backend_errors():
    p = E.SimParams(F32(1e-3), 5, F32(10), F32(12.5), 8)
    assert p.dtx == F32(F32(1e-3) / F32(10)) and p.dtz == F32(F32(1e-3) / F32(12.5)) and p.M == 4
    with pytest.raises(ValueError):
        E.backend("cpu")
    with pytest.raises(ValueError):
        E.MultiShotConfig([1.0], [1.0, 2.0], np.zeros(3, F32))
