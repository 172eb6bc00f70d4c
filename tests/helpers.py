"""Shared test inputs: small media and shots built with the product's setup
functions (fomo_julia_amd.init / elastic2d), handed identically to the oracle
and to the MI355X library."""
from __future__ import annotations

import numpy as np

from fomo_julia_amd import init as I
from fomo_julia_amd import elastic2d as E

F32 = np.float32
REC_CODES = {"vz": 0, "vx": 1, "p": 2}
SRC_CODES = {"pressure": 0, "force_z": 1, "force_x": 2}


class Case:
    """One medium + HABC + params + source + receivers."""

    def __init__(self, nx_in=72, nz_in=56, nbc=12, fd_order=8, fs=True, dx=10.0, nt=120,
                 src_type="pressure", src_xz=None, rec_types=("vz", "vx", "p"), layered=True,
                 seed=0, f0=25.0, water=False, storage="f32"):
        rng = np.random.default_rng(seed)
        vp = np.full((nx_in, nz_in), 2500.0, F32)
        vs = np.full((nx_in, nz_in), 1400.0, F32)
        rho = np.full((nx_in, nz_in), 2000.0, F32)
        if layered:
            vp[:, nz_in // 2:] = 3200.0
            vs[:, nz_in // 2:] = 1850.0
            rho[:, nz_in // 2:] = 2300.0
            vp += rng.uniform(-50, 50, vp.shape).astype(F32)
        if water:
            vp[:, :3] = 1500.0
            vs[:, :3] = 0.0
            rho[:, :3] = 1000.0
        self.dx = F32(dx)
        self.dt = F32(0.4 * dx / float(vp.max()))
        self.nt = nt
        self.nbc = nbc
        self.fd_order = fd_order
        b = None
        self.M = E.init_medium(vp, vs, rho, self.dx, self.dx, nbc, fd_order, b, free_surface=fs,
                               storage=storage)
        self.H = E.init_habc(self.M.nx, self.M.nz, nbc, self.dt, self.dx, self.dx,
                             I.julia_mean_f32(vp), b)
        self.p = E.SimParams(self.dt, nt, self.dx, self.dx, fd_order)
        self.a = I.get_fd_coefficients(fd_order)
        wav = I.ricker_wavelet(F32(f0), self.dt, nt)
        if src_xz is None:
            src_xz = (F32(self.M.x_max / F32(2)), F32(3 * dx))
        si = I.coord_to_index(F32(src_xz[0]), self.M.dx, self.M.pad)
        sj = I.coord_to_index(F32(src_xz[1]), self.M.dz, self.M.pad)
        self.src = E.Source(si, sj, wav, src_type)
        xr = np.arange(0, nx_in, 3, dtype=F32) * self.dx
        types = [rec_types[k % len(rec_types)] for k in range(len(xr))]
        zr = np.full(len(xr), 2 * dx, F32)
        self.rec = E.setup_receivers(xr, zr, self.M, type=types)

    def oracle(self, dtype=np.float32, flags=1, threads=None, storage="f32", fscale=None):
        """The oracle on this case; ``storage`` emulates BF16/FP16 storage (FP32 run)."""
        import oracle as O
        M, H, p = self.M, self.H, self.p
        if dtype == np.float64:
            # the same discrete scheme evaluated in Float64 (coefficients, HABC
            # constants and weights are the Float32 values of the scheme)
            return O.OracleRun(
                nx=M.nx, nz=M.nz, pad=M.pad, nbc=H.nbc, fd_order=p.fd_order,
                free_surface=M.is_free_surface, dt=p.dt, dtx=p.dtx, dtz=p.dtz,
                q=(H.qx, H.qz, H.qt_x, H.qt_z, H.qxt), a=self.a, rho_vx=M.rho_vx, rho_vz=M.rho_vz,
                lam=M.lam, mu_txx=M.mu_txx, mu_txz=M.mu_txz, w_vx=H.w_vx, w_vz=H.w_vz,
                w_tau=H.w_tau, src_i=self.src.i, src_j=self.src.j,
                src_type=SRC_CODES[self.src.type], wavelet=self.src.wavelet, rec_i=self.rec.i,
                rec_j=self.rec.j, rec_type=self.rec.type_codes(), nt=self.nt, dtype=np.float64,
                flags=flags, threads=threads)
        return O.OracleRun(
            nx=M.nx, nz=M.nz, pad=M.pad, nbc=H.nbc, fd_order=p.fd_order,
            free_surface=M.is_free_surface, dt=p.dt, dtx=p.dtx, dtz=p.dtz,
            q=(H.qx, H.qz, H.qt_x, H.qt_z, H.qxt), a=self.a, rho_vx=M.rho_vx, rho_vz=M.rho_vz,
            lam=M.lam, mu_txx=M.mu_txx, mu_txz=M.mu_txz, w_vx=H.w_vx, w_vz=H.w_vz, w_tau=H.w_tau,
            src_i=self.src.i, src_j=self.src.j, src_type=SRC_CODES[self.src.type],
            wavelet=self.src.wavelet, rec_i=self.rec.i, rec_j=self.rec.j,
            rec_type=self.rec.type_codes(), nt=self.nt, dtype=dtype, flags=flags, threads=threads,
            storage=storage, fscale=fscale)


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


def run_jl_case():
    """The reference's run.jl synthetic example (run.jl:17-61,90-131)."""
    DX = F32(10.0)
    NX, NZ, NBC, FD = 400, 200, 50, 8
    vp = np.full((NX, NZ), 3000, F32)
    vs = np.full((NX, NZ), 1800, F32)
    rho = np.full((NX, NZ), 2200, F32)
    vp[:, NZ // 2 - 1:] = 4000  # vp[:, NZ÷2:end] (1-based)
    vs[:, NZ // 2 - 1:] = 2400
    dt = F32(F32(0.5) * DX / vp.max())
    nt = int(np.ceil(F32(1.0) / dt))
    M = E.init_medium(vp, vs, rho, DX, DX, NBC, FD, None, free_surface=True)
    H = E.init_habc(M.nx, M.nz, NBC, dt, DX, DX, I.julia_mean_f32(vp), None)
    p = E.SimParams(dt, nt, DX, DX, FD)
    a = I.get_fd_coefficients(FD)
    wav = I.ricker_wavelet(F32(15.0), dt, nt)
    x_rec = np.array([500 + k * 3000 / 99 for k in range(100)]).astype(F32)
    rec = E.setup_receivers(x_rec, np.full(100, 100, F32), M, type="vz")
    si = I.coord_to_index(F32(M.x_max / F32(2)), M.dx, M.pad)
    sj = I.coord_to_index(F32(50.0), M.dz, M.pad)
    src = E.Source(si, sj, wav, "pressure")
    return dict(M=M, H=H, p=p, a=a, src=src, rec=rec, nt=nt)


class C1Case(Case):
    """Homogeneous 200 x 200 plain-method configuration: dx = dz = 10 m,
    vp 3000, vs 1732, rho 2000, fd_order 8, nbc 50, free surface, dt 1e-3 s,
    nt 1000; Ricker 15 Hz delayed 1.2/f0 as a force on vz at (1000 m, 20 m);
    200 receivers on every column at z = 10 m recording vx, vz and pressure."""

    def __init__(self, nt=1000, storage="f32", rec_types=("vx", "vz", "p")):
        nx_in = nz_in = 200
        vp = np.full((nx_in, nz_in), 3000.0, F32)
        vs = np.full((nx_in, nz_in), 1732.0, F32)
        rho = np.full((nx_in, nz_in), 2000.0, F32)
        self.dx = F32(10.0)
        self.dt = F32(1e-3)
        self.nt = nt
        self.nbc = 50
        self.fd_order = 8
        self.M = E.init_medium(vp, vs, rho, self.dx, self.dx, 50, 8, None, free_surface=True,
                               storage=storage)
        self.H = E.init_habc(self.M.nx, self.M.nz, 50, self.dt, self.dx, self.dx,
                             I.julia_mean_f32(vp), None)
        self.p = E.SimParams(self.dt, nt, self.dx, self.dx, 8)
        self.a = I.get_fd_coefficients(8)
        f0 = F32(15.0)
        wav = I.ricker_wavelet(f0, self.dt, nt, t0=1.2 / 15.0)
        si = I.coord_to_index(F32(1000.0), self.M.dx, self.M.pad)
        sj = I.coord_to_index(F32(20.0), self.M.dz, self.M.pad)
        self.src = E.Source(si, sj, wav, "force_z")
        xr = np.arange(nx_in, dtype=F32) * self.dx
        x_all = np.concatenate([xr] * len(rec_types))
        types = [t for t in rec_types for _ in range(nx_in)]
        self.rec = E.setup_receivers(x_all, np.full(len(x_all), 10.0, F32), self.M, type=types)
        self.rec_types = tuple(rec_types)

    def split(self, gather):
        """{type: [nt x 200]} from a gather over all receiver types."""
        g = np.asarray(gather)
        return {t: g[:, k * 200:(k + 1) * 200] for k, t in enumerate(self.rec_types)}
