"""ORACLE -- test infrastructure only (tests/, smoke(), bench.py cpu_baseline).

ctypes wrapper around oracle/oracle.cpp, the CPU restatement of the
reference's time-stepping path (src/kernels/*.jl, src/simulation/time_stepper.jl).
The product package never imports this module.

Parity status per function (what pins it) is listed in oracle.cpp's header and
exercised by tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

FMA = 1
SNAPSHOT = 2
STAGE_ALL = 255
STAGES = {"backup": 1, "inject": 2, "velocity": 4, "habc_v": 8, "stress": 16,
          "habc_s": 32, "free_surface": 64, "record": 128}


class OracleDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("nz", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("nbc", ctypes.c_int32), ("fd_order", ctypes.c_int32),
                ("free_surface", ctypes.c_int32),
                ("dt", ctypes.c_double), ("dtx", ctypes.c_double), ("dtz", ctypes.c_double),
                ("qx", ctypes.c_double), ("qz", ctypes.c_double), ("qt_x", ctypes.c_double),
                ("qt_z", ctypes.c_double), ("qxt", ctypes.c_double),
                ("flags", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("storage", ctypes.c_int32), ("fscale", ctypes.c_float * 5)]

# "bfp16" (block floating point fields, FP32 coefficients) and "bfp16c" (the
# same with FP16 coefficients) are a design study of a 16-bit field format:
# the oracle emulates them, no kernel implements them.
STORAGE = {"f32": 0, "bf16": 1, "f16": 2, "f16c": 3, "bfp16": 4, "bfp16c": 5, "f24": 6}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
        for name, ct in (("oracle_run_f32", ctypes.c_float), ("oracle_run_f64", ctypes.c_double)):
            fn = getattr(_lib, name)
            P = ctypes.POINTER(ct)
            PP = ctypes.POINTER(P)
            IP = ctypes.POINTER(ctypes.c_int)
            fn.argtypes = [ctypes.POINTER(OracleDesc), P, PP, PP, P, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, P, ctypes.c_int, IP, IP, IP, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, ctypes.c_int, P, ctypes.c_int]
            fn.restype = ctypes.c_int
        _lib.oracle_max_threads.restype = ctypes.c_int
        _lib.oracle_round_storage.restype = ctypes.c_float
        _lib.oracle_round_storage.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float]
        _lib.oracle_round_block.restype = None
        _lib.oracle_round_block.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    return _lib


def round_storage(storage: str, x: float, scale: float = 1.0) -> float:
    """x rounded to the storage format (at the FP16 scale), as the oracle rounds fields."""
    return lib().oracle_round_storage(STORAGE[storage], float(x), float(scale))


def round_block(x) -> np.ndarray:
    """One block floating-point segment (storage "bfp16") as the oracle rounds it."""
    y = np.ascontiguousarray(np.asarray(x, dtype=np.float32)).copy()
    lib().oracle_round_block(y.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(y))
    return y


def max_threads() -> int:
    return lib().oracle_max_threads()


class OracleRun:
    """State of one oracle run: 10 fields (vx vz txx tzz txz + *_old) and the gather.

    Arguments mirror the reference's objects: ``medium`` arrays are (nx, nz)
    numpy arrays (any order), weights are the (nz, nx) ``HABCConfig`` matrices.
    ``dtype`` np.float32 is the reference's FP32 path; np.float64 evaluates the
    same algorithm in double precision (inputs converted from the given
    arrays, which the caller may have computed in Float64).  ``storage``
    "bf16" / "f16" / "f16c" (FP32 only) emulates reduced-precision storage of
    the coefficients and fields; ``fscale`` are the FP16 field store scales.
    "bfp16" emulates the block floating-point field format of the design study
    (FP32 coefficients); "f24" the library's F24 storage (FP32 coefficients,
    fields rounded to a 16-bit significand).
    """

    def __init__(self, *, nx, nz, pad, nbc, fd_order, free_surface, dt, dtx, dtz,
                 q, a, rho_vx, rho_vz, lam, mu_txx, mu_txz, w_vx, w_vz, w_tau,
                 src_i, src_j, src_type, wavelet, rec_i, rec_j, rec_type, nt,
                 dtype=np.float32, flags=FMA, threads=None, storage="f32", fscale=None):
        self.dtype = np.dtype(dtype)
        self.nx, self.nz, self.nt = int(nx), int(nz), int(nt)
        self.desc = OracleDesc(int(nx), int(nz), int(pad), int(nbc), int(fd_order),
                               int(bool(free_surface)), float(dt), float(dtx), float(dtz),
                               *[float(v) for v in q], int(flags),
                               int(threads or max_threads()), STORAGE[storage],
                               (ctypes.c_float * 5)(*(fscale if fscale is not None else [1.0] * 5)))
        if storage != "f32" and self.dtype != np.float32:
            raise ValueError("storage emulation runs on the FP32 oracle only")
        cv = lambda x: np.asfortranarray(np.asarray(x, dtype=self.dtype))
        self.a = np.ascontiguousarray(np.asarray(a, dtype=self.dtype))
        self.med = [cv(x) for x in (rho_vx, rho_vz, lam, mu_txx, mu_txz)]
        self.w = [cv(x) for x in (w_vx, w_vz, w_tau)]
        for x in self.med:
            assert x.shape == (nx, nz)
        for x in self.w:
            assert x.shape == (nz, nx)
        self.fields = np.zeros((10, nz, nx), dtype=self.dtype)  # C order: [c][j][i]
        self.src = (int(src_i), int(src_j), int(src_type))
        self.wavelet = np.ascontiguousarray(np.asarray(wavelet, dtype=self.dtype))
        self.rec_i = np.ascontiguousarray(np.asarray(rec_i, dtype=np.int32))
        self.rec_j = np.ascontiguousarray(np.asarray(rec_j, dtype=np.int32))
        rt = np.broadcast_to(np.asarray(rec_type, dtype=np.int32), self.rec_i.shape)
        self.rec_type = np.ascontiguousarray(rt)
        self.gather = np.zeros((len(self.rec_i), self.nt), dtype=self.dtype)  # [r][k] == [nt x n_rec] col-major

    def field(self, c: int) -> np.ndarray:
        """Field c as an (nx, nz) view (Julia index order)."""
        return self.fields[c].T

    def run(self, k_first: int, n_steps: int, stage_mask: int = STAGE_ALL) -> None:
        L = lib()
        is32 = self.dtype == np.float32
        ct = ctypes.c_float if is32 else ctypes.c_double
        P = ctypes.POINTER(ct)
        ptr = lambda x: x.ctypes.data_as(P)
        med = (P * 5)(*[ptr(x) for x in self.med])
        w = (P * 3)(*[ptr(x) for x in self.w])
        IP = ctypes.POINTER(ctypes.c_int)
        fn = L.oracle_run_f32 if is32 else L.oracle_run_f64
        rc = fn(ctypes.byref(self.desc), ptr(self.a), med, w, ptr(self.fields),
                self.src[0], self.src[1], self.src[2], ptr(self.wavelet), len(self.wavelet),
                self.rec_i.ctypes.data_as(IP), self.rec_j.ctypes.data_as(IP),
                self.rec_type.ctypes.data_as(IP), len(self.rec_i), self.nt,
                int(k_first), int(n_steps), ptr(self.gather), int(stage_mask))
        if rc != 0:
            raise ValueError(f"oracle rejected the run (code {rc})")

    def gather_nt_by_nrec(self) -> np.ndarray:
        """The gather as the reference's [nt x n_rec] matrix."""
        return self.gather.T
