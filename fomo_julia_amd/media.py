"""Seeded synthetic media for the benchmark and test workloads.

All generators use numpy's PCG64 (``np.random.default_rng(seed)``), which is
portable across platforms, and return Float32 (nx, nz) arrays (vp, vs, rho)
in Julia index order.  vs and rho follow the reference loader's defaults
where a model has only vp: vs = vp / 1.73, rho = 310 vp^0.25 (Gardner)
(src/io/model_loader.jl:128-135).
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def _gardner(vp):
    return (F32(310.0) * np.power(vp.astype(F32), F32(0.25))).astype(F32)


def homogeneous(nx=200, nz=200, vp=3000.0, vs=1732.0, rho=2000.0):
    """Plain-method CPU case: constant medium."""
    return (np.full((nx, nz), vp, F32), np.full((nx, nz), vs, F32), np.full((nx, nz), rho, F32))


def seam_like(nx=4444, nz=3819, dx=12.5, dz=12.5, seed=2, water_depth=150.0, n_lenses=20):
    """SEAM-shaped model: a water layer (vp 1500, vs 0, rho 1000) over a vp
    gradient 1800 -> 4500 m/s with depth, `n_lenses` Gaussian velocity lenses
    (+-15 %, sigma 200-800 m) and a salt-like body (vp 4500, an ellipse with a
    seeded boundary perturbation).  vs = vp/1.73, Gardner rho below the water."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx, dtype=np.float64) * dx)[:, None]
    z = (np.arange(nz, dtype=np.float64) * dz)[None, :]
    zmax, xmax = (nz - 1) * dz, (nx - 1) * dx
    water = np.broadcast_to(z < water_depth, (nx, nz))
    frac = np.clip((z - water_depth) / max(zmax - water_depth, 1.0), 0.0, 1.0)
    vp = np.broadcast_to(1800.0 + (4500.0 - 1800.0) * frac, (nx, nz)).copy()
    for _ in range(n_lenses):
        cx = rng.uniform(0.0, xmax)
        cz = rng.uniform(water_depth, zmax)
        amp = rng.uniform(-0.15, 0.15)
        sig = rng.uniform(200.0, 800.0)
        gx = np.exp(-((x - cx) ** 2) / (2 * sig * sig))
        gz = np.exp(-((z - cz) ** 2) / (2 * sig * sig))
        vp *= 1.0 + amp * (gx * gz)
    # salt: ellipse with a low-order seeded boundary ripple
    sx, sz = rng.uniform(0.35, 0.65) * xmax, rng.uniform(0.35, 0.55) * zmax
    ax, az = rng.uniform(0.10, 0.18) * xmax, rng.uniform(0.08, 0.14) * zmax
    ph = rng.uniform(0, 2 * np.pi, 3)
    th = np.arctan2((z - sz) / az, (x - sx) / ax)
    rad = 1.0 + 0.12 * np.sin(3 * th + ph[0]) + 0.06 * np.sin(5 * th + ph[1]) + 0.04 * np.sin(7 * th + ph[2])
    salt = (((x - sx) / ax) ** 2 + ((z - sz) / az) ** 2) < rad ** 2
    vp[salt & ~water] = 4500.0
    vp[water] = 1500.0
    vp = vp.astype(F32)
    vs = (vp / F32(1.73)).astype(F32)
    rho = _gardner(vp)
    vs[water] = 0.0
    rho[water] = 1000.0
    return vp, vs, rho


def marmousi_like(nx=1700, nz=350, dx=10.0, dz=10.0, seed=3, n_layers=12, n_faults=3):
    """Marmousi-shaped model: `n_layers` seeded dipping layers with vp in
    1500-5500 m/s (increasing with depth on average), offset by `n_faults`
    normal faults; vs = vp/1.8, Gardner rho."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx, dtype=np.float64) * dx
    zmax = (nz - 1) * dz
    # layer interfaces: depth(x) = d0 + slope*x + fault throws
    d0 = np.sort(rng.uniform(0.02, 0.98, n_layers - 1)) * zmax
    slopes = rng.uniform(-0.12, 0.12, n_layers - 1)
    vels = np.sort(rng.uniform(1500.0, 5500.0, n_layers))
    vels[0] = 1500.0
    vels[-1] = 5500.0
    fx = np.sort(rng.uniform(0.15, 0.85, n_faults)) * x[-1]
    throws = rng.uniform(-0.06, 0.06, n_faults) * zmax
    shift = np.zeros(nx)
    for f, t in zip(fx, throws):
        shift += np.where(x > f, t, 0.0)
    z = np.arange(nz, dtype=np.float64) * dz
    vp = np.full((nx, nz), vels[0])
    for k in range(n_layers - 1):
        iface = d0[k] + slopes[k] * (x - x.mean()) + shift
        vp = np.where(z[None, :] >= iface[:, None], vels[k + 1], vp)
    vp = vp.astype(F32)
    vs = (vp / F32(1.8)).astype(F32)
    return vp, vs, _gardner(vp)


def layered_with_lenses(nx, nz, dx=10.0, dz=10.0, seed=5, n_layers=6, n_lenses=6,
                        water_depth=0.0):
    """Served-stream media: a few flat layers plus Gaussian lenses (as seam_like)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx, dtype=np.float64) * dx)[:, None]
    z = (np.arange(nz, dtype=np.float64) * dz)[None, :]
    zmax, xmax = (nz - 1) * dz, (nx - 1) * dx
    bounds = np.sort(rng.uniform(0.0, zmax, n_layers - 1))
    vels = np.sort(rng.uniform(1600.0, 4200.0, n_layers))
    vp = np.full((nx, nz), vels[0])
    for k, zb in enumerate(bounds):
        vp = np.where(np.broadcast_to(z >= zb, (nx, nz)), vels[k + 1], vp)
    for _ in range(n_lenses):
        cx, cz = rng.uniform(0, xmax), rng.uniform(0, zmax)
        amp, sig = rng.uniform(-0.15, 0.15), rng.uniform(0.03, 0.1) * max(xmax, zmax)
        vp = vp * (1.0 + amp * np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2 * sig * sig)))
    vp = vp.astype(F32)
    vs = (vp / F32(1.73)).astype(F32)
    rho = _gardner(vp)
    if water_depth > 0:
        w = np.broadcast_to(z < water_depth, (nx, nz))
        vp[w], vs[w], rho[w] = 1500.0, 0.0, 1000.0
    return vp, vs, rho


def cfl_dt(vp, dx, cfl=0.4) -> np.float32:
    """dt = CFL * dx / max(vp), as run.jl:52 (Float32)."""
    return F32(F32(cfl) * F32(dx) / F32(np.max(vp)))
