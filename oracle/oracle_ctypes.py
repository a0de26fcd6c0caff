"""ctypes binding of the CPU oracle (oracle/libribo_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product.  It mirrors the product
World API (ribossome_amd.World) so a test can run the same world on both and
compare buffers.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libribo_oracle.so")
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P, U32, I, SZ = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t
        sig = {
            "oracle_create": (P, [P, U32, I]),
            "oracle_destroy": (None, [P]),
            "oracle_invalidate_light": (None, [P]),
            "oracle_set_params": (None, [P, P]),
            "oracle_get_params": (None, [P, P]),
            "oracle_set_settings": (I, [P, P]),
            "oracle_upload": (I, [P, I, P, SZ]),
            "oracle_download": (I, [P, I, P, SZ]),
            "oracle_download_f64": (I, [P, I, P, SZ]),
            "oracle_encode_steps": (I, [P, U32]),
            "oracle_ledger": (None, [P, P]),
            "oracle_chem_ops": (I, [P, P, U32, P]),
            "oracle_run_kernel": (I, [P, ctypes.c_char_p]),
            "oracle_mg_residual": (ctypes.c_double, [P]),
            "oracle_mg_stats": (None, [P, ctypes.POINTER(U32), ctypes.POINTER(ctypes.c_double)]),
            "oracle_solver_residuals": (None, [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(U32)]),
            "oracle_set_threads": (None, [I]),
            "oracle_get_threads": (I, []),
            "oracle_pcg4d": (None, [P, P, SZ]),
            "oracle_hash": (U32, [U32]),
            "oracle_expf": (ctypes.c_float, [ctypes.c_float]),
            "oracle_sinf": (ctypes.c_float, [ctypes.c_float]),
            "oracle_mg_solve": (I, [U32, U32, P, P, P, U32, I]),
            "oracle_mg_solve_f64": (I, [U32, U32, P, P, P, U32]),
            "oracle_mg_restrict": (I, [U32, P, P, P, P, P]),
            "oracle_mg_prolong": (I, [U32, P, P, P]),
            "oracle_bl_transfer": (I, [U32, P, P, P, P, P, P]),
            "oracle_coarse_inverse_error": (ctypes.c_double, [U32, P, ctypes.c_float]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


_F32 = {2, 3, 4, 5, 6, 7, 8, 9, 11, 15}
_I32 = {14}


def _dtype(which: int):
    return np.float32 if which in _F32 else np.int32 if which in _I32 else np.uint32


class OracleWorld:
    """CPU restatement of one world; fp64=True runs the fluid in binary64."""

    def __init__(self, cfg, seed: int = 1, fp64: bool = False):
        import ribossome_amd as ra  # buffer sizes and struct layouts are shared

        self.ra = ra
        self.cfg = cfg
        c = cfg.c()
        self._h = lib().oracle_create(ctypes.byref(c), seed, 1 if fp64 else 0)
        if not self._h:
            raise RuntimeError("oracle_create failed")

    def close(self):
        if self._h:
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = self.ra.SimParams()
        lib().oracle_get_params(self._h, ctypes.byref(p))
        return p

    @params.setter
    def params(self, p):
        lib().oracle_set_params(self._h, ctypes.byref(p))

    def set_param(self, **kw):
        p = self.params
        for k, v in kw.items():
            setattr(p, k, v)
        self.params = p

    def set_settings(self, s):
        rc = lib().oracle_set_settings(self._h, ctypes.byref(s))
        if rc != 0:
            raise RuntimeError(f"oracle_set_settings: {rc}")

    def upload(self, which: int, data: np.ndarray):
        a, dt = np.asarray(data), _dtype(which)
        if a.dtype != dt and a.dtype.kind in "iu" and np.dtype(dt).kind in "iu" and a.dtype.itemsize == 4:
            a = a.view(dt)  # 32-bit words: signed and unsigned are the same bits
        a = np.ascontiguousarray(a, dtype=dt)
        rc = lib().oracle_upload(self._h, which, _ptr(a), a.nbytes)
        if rc != 0:
            raise RuntimeError(f"oracle_upload {which}: {rc}")

    def download(self, which: int) -> np.ndarray:
        n = self.ra.buffer_bytes(self.cfg, which) // 4
        a = np.empty(n, dtype=_dtype(which))
        rc = lib().oracle_download(self._h, which, _ptr(a), a.nbytes)
        if rc != 0:
            raise RuntimeError(f"oracle_download {which}: {rc}")
        return a

    def download_f64(self, which: int) -> np.ndarray:
        n = self.ra.buffer_bytes(self.cfg, which) // 4
        a = np.empty(n, dtype=np.float64)
        rc = lib().oracle_download_f64(self._h, which, _ptr(a), n)
        if rc != 0:
            raise RuntimeError(f"oracle_download_f64 {which}: {rc}")
        return a

    def encode_steps(self, n: int) -> int:
        return lib().oracle_encode_steps(self._h, n)

    def invalidate_light(self):
        lib().oracle_invalidate_light(self._h)

    def run(self, n: int):
        done = 0
        while done < n:
            done += self.encode_steps(min(64, n - done))

    def chem_ops(self, ops) -> np.ndarray:
        """Exchanges (cell, channel, kind, count) rows, applied in order;
        returns the monomers each moved."""
        a = np.ascontiguousarray(ops, dtype=np.uint32).reshape(-1, 4)
        done = np.zeros(len(a), dtype=np.uint32)
        rc = lib().oracle_chem_ops(self._h, _ptr(a), len(a), _ptr(done))
        if rc != 0:
            raise RuntimeError(f"oracle_chem_ops: {rc}")
        return done

    def ledger(self):
        out = (ctypes.c_uint32 * 12)()
        lib().oracle_ledger(self._h, out)
        return self.ra.Ledger.from_words(out)

    def run_kernel(self, name: str):
        if lib().oracle_run_kernel(self._h, name.encode()) != 0:
            raise RuntimeError(f"unknown oracle kernel {name}")

    def mg_stats(self):
        c, r = ctypes.c_uint32(), ctypes.c_double()
        lib().oracle_mg_stats(self._h, ctypes.byref(c), ctypes.byref(r))
        return int(c.value), float(r.value)

    def solver_residuals(self) -> dict:
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint32()
        lib().oracle_solver_residuals(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {"last": float(a.value), "last_rounded": float(b.value), "restarts": int(c.value)}

    def mg_residual(self) -> float:
        return float(lib().oracle_mg_residual(self._h))


def pcg4d(v: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(v, dtype=np.uint32).reshape(-1, 4)
    out = np.empty_like(v)
    lib().oracle_pcg4d(_ptr(v), _ptr(out), v.shape[0])
    return out


def expf(x: float) -> float:
    return float(lib().oracle_expf(x))


def sinf(x: float) -> float:
    return float(lib().oracle_sinf(x))


def mg_solve(n: int, min_size: int, fluid: np.ndarray, rhs: np.ndarray, p0: np.ndarray, cycles: int, fp64=False):
    if fp64:
        f = np.ascontiguousarray(fluid, dtype=np.uint32)
        r = np.ascontiguousarray(rhs, dtype=np.float64)
        p = np.array(p0, dtype=np.float64, copy=True)
        lib().oracle_mg_solve_f64(n, min_size, _ptr(f), _ptr(r), _ptr(p), cycles)
        return p
    f = np.ascontiguousarray(fluid, dtype=np.uint32)
    r = np.ascontiguousarray(rhs, dtype=np.float32)
    p = np.array(p0, dtype=np.float32, copy=True)
    lib().oracle_mg_solve(n, min_size, _ptr(f), _ptr(r), _ptr(p), cycles, 0)
    return p


def mg_restrict(n: int, fluid: np.ndarray, rhs: np.ndarray, p: np.ndarray):
    """One mg_restrict of an n x n level: (coarse rhs, coarse fluid mask)."""
    f = np.ascontiguousarray(fluid, dtype=np.uint32)
    r = np.ascontiguousarray(rhs, dtype=np.float32)
    q = np.ascontiguousarray(p, dtype=np.float32)
    rc = np.zeros((n // 2) * (n // 2), np.float32)
    fc = np.zeros((n // 2) * (n // 2), np.uint32)
    if lib().oracle_mg_restrict(n, _ptr(f), _ptr(r), _ptr(q), _ptr(rc), _ptr(fc)) != 0:
        raise ValueError("n must be even")
    return rc, fc


def mg_prolong(n: int, fluid: np.ndarray, p: np.ndarray, p_c: np.ndarray) -> np.ndarray:
    """One mg_prolong of an n x n level: p + the parent's coarse value on fluid cells."""
    f = np.ascontiguousarray(fluid, dtype=np.uint32)
    q = np.array(p, dtype=np.float32, copy=True)
    c = np.ascontiguousarray(p_c, dtype=np.float32)
    if lib().oracle_mg_prolong(n, _ptr(f), _ptr(q), _ptr(c)) != 0:
        raise ValueError("n must be even")
    return q


def bl_transfer(n: int, fluid: np.ndarray, fluid_c: np.ndarray, c: np.ndarray, r: np.ndarray):
    """The tolerance mode's bilinear transfer in binary64: (P c, R r)."""
    f = np.ascontiguousarray(fluid, dtype=np.uint32)
    fc = np.ascontiguousarray(fluid_c, dtype=np.uint32)
    ci = np.ascontiguousarray(c, dtype=np.float64)
    ri = np.ascontiguousarray(r, dtype=np.float64)
    po = np.zeros(n * n)
    rc = np.zeros((n // 2) * (n // 2))
    if lib().oracle_bl_transfer(n, _ptr(f), _ptr(fc), _ptr(ci), _ptr(po), _ptr(ri), _ptr(rc)) != 0:
        raise ValueError("n must be even")
    return po, rc


def coarse_inverse_error(n: int, fluid: np.ndarray, eps: float) -> float:
    f = np.ascontiguousarray(fluid, dtype=np.uint32)
    return float(lib().oracle_coarse_inverse_error(n, _ptr(f), eps))
