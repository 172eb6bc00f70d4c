"""Host-side setup that feeds the time-stepping path.

Restates ``src/utils/init.jl`` (FD coefficient table, Ricker wavelet, padding,
staggered moduli/density, HABC coefficients and weight ramps, receiver indices)
and the index discretisation of ``src/simulation/shot_manager.jl:189-191``.

Julia's arithmetic is reproduced exactly, including its type promotion:
``Float32 op Float64 -> Float64`` (numpy with two numpy scalars does the same;
a bare Python float is first turned into ``np.float64`` so numpy's weak-scalar
rule never keeps a Julia Float64 expression in float32), ``round`` with ties to
even, and left-to-right evaluation without FMA contraction (Julia does not
contract ``a*b + c`` unless asked to).
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

F32 = np.float32
F64 = np.float64

# src/utils/init.jl:13-19 -- the reference's table, Float64 literals rounded to
# Float32 (``Float32[...]``).  Kept literally, including order 10's a1 =
# 1.2115478515625 (the exact staggered value is 19845/16384 = 1.21124267578125).
_FD_TABLE = {
    2: (1.0,),
    4: (1.125, -0.041666667),
    6: (1.171875, -0.065104167, 0.0046875),
    8: (1.1962890625, -0.079752604167, 0.0095703125, -0.000697544643),
    10: (1.2115478515625, -0.089721679687, 0.0138427734375, -0.00176565987723,
         0.0001186795166),
}


def staggered_fd_coefficients_exact(order: int) -> list[Fraction]:
    """Exact Taylor coefficients of the 2M-order staggered first derivative.

    a_m solves sum_m a_m (2m-1)^(2k-1) = delta_{k,1} for k = 1..M, i.e.
    the stencil sum_m a_m (f(x+(m-1/2)h) - f(x-(m-1/2)h)) / h is exact for
    polynomials of degree <= 2M.  Closed form (Lagrange on the odd nodes):
        a_m = (-1)^(m+1) prod_{n!=m} (2n-1)^2 / ((2m-1) prod_{n!=m} |(2m-1)^2-(2n-1)^2|)
    """
    if order < 2 or order % 2:
        raise ValueError(f"Unsupported FD order: {order}")
    M = order // 2
    out = []
    for m in range(1, M + 1):
        num = Fraction(1)
        den = Fraction(2 * m - 1)
        for n in range(1, M + 1):
            if n == m:
                continue
            num *= (2 * n - 1) ** 2
            den *= abs((2 * m - 1) ** 2 - (2 * n - 1) ** 2)
        sign = 1 if m % 2 == 1 else -1
        out.append(sign * num / den)
    return out


def get_fd_coefficients(order: int) -> np.ndarray:
    """``get_fd_coefficients`` (src/utils/init.jl:26-29).

    Orders 2..10 come from the reference table.  Order 12 (and 14, 16), which
    the reference rejects with "Unsupported FD order", come from the exact
    rationals above rounded to Float32.
    """
    if order in _FD_TABLE:
        return np.array(_FD_TABLE[order], dtype=F32)
    if order in (12, 14, 16):
        return np.array([float(c) for c in staggered_fd_coefficients_exact(order)], dtype=F32)
    raise ValueError(f"Unsupported FD order: {order}")


def _jl(x):
    """Map a scalar to the numpy type Julia would give it."""
    if isinstance(x, (np.floating,)):
        return x
    if isinstance(x, (int, np.integer)):
        return int(x)
    return F64(x)


def ricker_wavelet(f0, dt, nt: int, t0=None) -> np.ndarray:
    """``ricker_wavelet`` (src/utils/init.jl:40-49).

    Promotion chain of the reference: ``t0 = 1.0/f0`` is Float64;
    ``(i-1)*dt`` is in dt's type, ``- t0`` promotes to Float64;
    ``pi*f0`` is in f0's type (Float32(pi)*f0 for a Float32 f0), then ``*tau``
    in Float64; ``exp`` in Float64; the sample is rounded to Float32.
    ``t0`` (Float64) overrides the 1/f0 delay, e.g. 1.2/f0.
    """
    f0 = _jl(f0)
    dt = _jl(dt)
    t0 = F64(1.0) / F64(f0) if t0 is None else F64(t0)
    pf = (F32(np.pi) * f0) if isinstance(f0, F32) else (F64(np.pi) * F64(f0))
    out = np.zeros(nt, dtype=F32)
    dtype_dt = F32 if isinstance(dt, F32) else F64
    for i in range(1, nt + 1):
        tau = F64(dtype_dt(i - 1) * dt) - t0
        x = F64(pf) * tau
        arg = x * x
        out[i - 1] = F32((F64(1.0) - F64(2.0) * arg) * np.exp(-arg))
    return out


def _pad_array(data: np.ndarray, pad: int) -> np.ndarray:
    """``_pad_array`` (src/utils/init.jl:119-134): edge replication by ``pad``.

    Arrays are (nx, nz) in Julia's index order; memory order is irrelevant
    here because the padded arrays are handed over as (nx, nz) Fortran arrays.
    """
    nx, nz = data.shape
    res = np.zeros((nx + 2 * pad, nz + 2 * pad), dtype=F32)
    res[pad:pad + nx, pad:pad + nz] = data.astype(F32)
    for i in range(pad):
        res[i, :] = res[pad, :]
        res[-i - 1, :] = res[-pad - 1, :]
    for j in range(pad):
        res[:, j] = res[:, pad]
        res[:, -j - 1] = res[:, -pad - 1]
    return res


def _compute_staggered_params(vp, vs, rho, dtype=F32):
    """``_compute_staggered_params`` (src/utils/init.jl:136-158).

    Returns (lam, mu_txx, mu_txz, rho_vx, rho_vz).  ``dtype=np.float64``
    gives the same formulas evaluated in Float64 (the Float64 oracle run).
    """
    vp = vp.astype(dtype)
    vs = vs.astype(dtype)
    rho = rho.astype(dtype)
    two = dtype(2.0)
    with np.errstate(divide="ignore"):
        mu = rho * (vs * vs)
        lam = rho * (vp * vp) - two * mu
        rho_vx = rho.copy()
        rho_vz = np.zeros_like(rho)
        rho_vz[:-1, :-1] = dtype(0.25) * (((rho[:-1, :-1] + rho[1:, :-1]) + rho[:-1, 1:]) + rho[1:, 1:])
        rho_vz[-1, :] = rho_vz[-2, :]
        rho_vz[:, -1] = rho_vz[:, -2]
        one = dtype(1.0)
        mu_txz = np.zeros_like(rho)
        mu_txz[:, :-1] = two / (one / mu[:, :-1] + one / mu[:, 1:])
        mu_txz[:, -1] = mu_txz[:, -2]
    return lam, mu, mu_txz, rho_vx, rho_vz


def habc_coefficients(v_ref, dt, dx, dz):
    """Higdon extrapolation constants of ``init_habc`` (src/utils/init.jl:175-185).

    Returns Float32 (qx, qz, qt_x, qt_z, qxt).  ``v_ref*dt/dx`` is evaluated in
    the promoted type of the arguments, as Julia would.
    """
    v_ref, dt, dx, dz = _jl(v_ref), _jl(dt), _jl(dx), _jl(dz)
    rx = F32((v_ref * dt) / dx)
    rz = F32((v_ref * dt) / dz)
    b_p = F32(0.45)
    beta = F32(1.0)
    one = F32(1.0)
    qx = F32((b_p * (beta + rx) - rx) / ((beta + rx) * (one - b_p)))
    qz = F32((b_p * (beta + rz) - rz) / ((beta + rz) * (one - b_p)))
    qt_x = F32((b_p * (beta + rx) - beta) / ((beta + rx) * (one - b_p)))
    qt_z = F32((b_p * (beta + rz) - beta) / ((beta + rz) * (one - b_p)))
    qxt = F32(b_p / (b_p - F32(1.0)))
    return qx, qz, qt_x, qt_z, qxt


def habc_ramp(nbc: int, offset: float) -> np.ndarray:
    """Weight as a function of the integer distance to the outer edge.

    ``Float32(clamp((dist - off)/nbc, 0, 1))`` in Float64
    (src/utils/init.jl:188-195); entry ``nbc+1`` and beyond are 1.
    """
    d = np.arange(nbc + 2, dtype=F64)
    return np.clip((d - F64(offset)) / F64(nbc), 0.0, 1.0).astype(F32)


HABC_OFFSETS = {"vx": 0.0, "vz": 0.5, "tau": 0.75}


def habc_weight_matrix(nx: int, nz: int, nbc: int, offset: float) -> np.ndarray:
    """The reference's ``[nz, nx]`` weight matrix (``w[j, i]``), returned as a
    numpy array of shape (nz, nx) whose Fortran order matches Julia's memory."""
    i = np.arange(1, nx + 1)
    j = np.arange(1, nz + 1)
    dist = np.minimum(np.minimum(i[None, :] - 1, nx - i[None, :]),
                      np.minimum(j[:, None] - 1, nz - j[:, None]))
    return np.clip((dist.astype(F64) - F64(offset)) / F64(nbc), 0.0, 1.0).astype(F32)


def coord_to_index(x, dx, pad: int) -> int:
    """``_coord_to_index`` (src/simulation/shot_manager.jl:189-191):
    ``round(Int32, x/dx) + pad + 1`` with Float32 division, ties to even."""
    q = F32(x) / F32(dx)
    return int(np.rint(q)) + pad + 1


def receiver_indices(x, z, dx, dz, pad: int):
    """Index rule of ``setup_receivers`` (src/utils/init.jl:214-227):
    ``round(Int, x[r]/M.dx) + M.pad + 1``; the division runs in the promoted
    type of x's element type and Float32 dx (Float32 input -> Float32)."""
    x = np.asarray(x)
    z = np.asarray(z)
    dt_x = F32 if x.dtype == F32 else F64
    dt_z = F32 if z.dtype == F32 else F64
    qi = x.astype(dt_x) / dt_x(F32(dx))
    qj = z.astype(dt_z) / dt_z(F32(dz))
    return (np.rint(qi).astype(np.int64) + pad + 1,
            np.rint(qj).astype(np.int64) + pad + 1)


def julia_sum_f32(a: np.ndarray, blksize: int = 1024) -> np.float32:
    """Julia's pairwise ``sum`` of a Float32 array in column-major order
    (Base.mapreduce_impl: split at ``ifirst + (ilast-ifirst)>>1`` until a block
    has fewer than ``blksize`` elements, then a sequential leaf sum).

    Leaves are summed left to right; Julia's ``@simd`` leaf may reassociate, so
    the result equals Julia's exactly whenever leaf sums are exact (as for the
    integer-valued velocities of the reference's example model)."""
    flat = np.asarray(a, dtype=F32).ravel(order="F")

    def rec(lo, hi):  # inclusive bounds
        if lo == hi:
            return flat[lo]
        if hi - lo < blksize:
            v = F32(flat[lo] + flat[lo + 1])
            for t in range(lo + 2, hi + 1):
                v = F32(v + flat[t])
            return v
        mid = lo + ((hi - lo) >> 1)
        return F32(rec(lo, mid) + rec(mid + 1, hi))

    return rec(0, flat.size - 1)


def julia_mean_f32(a: np.ndarray) -> np.float32:
    """``Statistics.mean`` of a Float32 array: pairwise sum / length (Float32)."""
    return F32(julia_sum_f32(a) / F32(np.asarray(a).size))
