"""ctypes binding of oracle/_ref/libnodals_ref.so (the reference's host setup
code compiled from the reference tree by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: the checker side of setup parity tests.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional

import numpy as np

REF_LIB = Path(__file__).resolve().parent / "_ref" / "libnodals_ref.so"


def available() -> bool:
    return REF_LIB.exists()


class RefCase:
    def __init__(self, mesh: str, re: float = 20.0, bulk: float = 1.0, wall: str = "patch_0_0",
                 inlet: str = "patch_2_0", outlet: str = "patch_1_0", alpha_u: float = 0.5, dg_inlet: bool = False):
        L = C.CDLL(str(REF_LIB))
        for fn in (L.ref_open, L.ref_open_dg):
            fn.restype = C.c_void_p
            fn.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_char_p, C.c_char_p, C.c_char_p, C.c_double]
        L.ref_close.argtypes = [C.c_void_p]
        L.ref_size.restype = C.c_longlong
        L.ref_size.argtypes = [C.c_void_p, C.c_char_p]
        L.ref_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.ref_build_sa.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        L.ref_schur_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_assemble.argtypes = [C.c_void_p] + [C.c_void_p] * 3 + [C.c_double] + [C.c_void_p] * 4
        L.ref_pressure_drop.restype = C.c_double
        L.ref_pressure_drop.argtypes = [C.c_void_p, C.c_void_p]
        L.ref_last_error.restype = C.c_char_p
        self.L = L
        opener = L.ref_open_dg if dg_inlet else L.ref_open
        self.h = opener(str(mesh).encode(), re, bulk, wall.encode(), inlet.encode(), outlet.encode(), alpha_u)
        if not self.h:
            raise RuntimeError("reference setup failed: " + L.ref_last_error().decode())
        self.nv = self.get("mom_row", np.int64).size - 1
        self.nc = self.get("cell_det", np.float64).size

    def close(self):
        if self.h:
            self.L.ref_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get(self, name: str, dtype) -> np.ndarray:
        n = self.L.ref_size(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        out = np.empty(n // np.dtype(dtype).itemsize, dtype=dtype)
        self.L.ref_get(self.h, name.encode(), out.ctypes.data)
        return out

    def build_sa(self, terminal=1000, power_its=16, safety=1.5, low_fraction=0.05, damping=4.0 / 3.0):
        if self.L.ref_build_sa(self.h, terminal, power_its, safety, low_fraction, damping) != 0:
            raise RuntimeError(self.L.ref_last_error().decode())

    def schur_apply(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self.L.ref_schur_apply(self.h, x.ctypes.data, y.ctypes.data)
        return y

    def assemble(self, U: np.ndarray, alpha_u: float):
        U = np.ascontiguousarray(U, dtype=np.float64)
        nnz = int(self.get("mom_row", np.int64)[-1])
        aval = np.empty(nnz)
        conv = np.empty(3 * self.nv)
        rau = np.empty(self.nv)
        delta = np.empty(self.nv)
        rc = self.L.ref_assemble(self.h, U[0].ctypes.data, U[1].ctypes.data, U[2].ctypes.data, alpha_u,
                                 aval.ctypes.data, conv.ctypes.data, rau.ctypes.data, delta.ctypes.data)
        if rc != 0:
            raise RuntimeError(self.L.ref_last_error().decode())
        return aval, conv.reshape(3, self.nv), rau, delta

    @classmethod
    def blank(cls) -> "RefCase":
        """Handle without a mesh, for ref_sa_from_csr."""
        self = cls.__new__(cls)
        L = C.CDLL(str(REF_LIB))
        L.ref_new.restype = C.c_void_p
        L.ref_close.argtypes = [C.c_void_p]
        L.ref_size.restype = C.c_longlong
        L.ref_size.argtypes = [C.c_void_p, C.c_char_p]
        L.ref_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.ref_sa_from_csr.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                      C.c_double, C.c_double]
        L.ref_last_error.restype = C.c_char_p
        self.L = L
        self.h = L.ref_new()
        return self

    def sa_from_csr(self, row, col, val, terminal=1000, power_its=16, safety=1.5, damping=4.0 / 3.0):
        row = np.ascontiguousarray(row, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        rc = self.L.ref_sa_from_csr(self.h, row.size - 1, row.ctypes.data, col.ctypes.data, val.ctypes.data,
                                    terminal, power_its, safety, damping)
        if rc != 0:
            raise RuntimeError(self.L.ref_last_error().decode())

    def pressure_drop(self, p: np.ndarray) -> float:
        p = np.ascontiguousarray(p, dtype=np.float64)
        return float(self.L.ref_pressure_drop(self.h, p.ctypes.data))


REF_SST_LIB = Path(__file__).resolve().parent / "_ref" / "libnodals_ref_sst.so"


def port_available() -> bool:
    return PORT_SST_LIB.exists()


def sst_available() -> bool:
    return REF_SST_LIB.exists()


class RefSst:
    """The reference's portable SST-2003 and wall-law cores
    (modules/physics/sst2003m_host.cpp, modules/wall/portable/wall_host.cpp),
    compiled from the reference tree, looped over arrays."""

    def __init__(self):
        L = C.CDLL(str(REF_SST_LIB))
        vp = C.c_void_p
        L.ref_sst_local.argtypes = [C.c_longlong, C.c_int, vp, vp, vp, C.c_double, C.c_double, vp, vp, vp, vp, vp,
                                    C.c_double, C.c_double, C.c_int, vp, vp]
        L.ref_spalding.argtypes = [C.c_longlong, vp, vp, C.c_double, vp, vp]
        L.ref_profile5.argtypes = [C.c_longlong, vp, vp, vp, C.c_double, vp, vp]
        L.ref_of_auto.argtypes = [C.c_longlong, vp, vp, vp]
        L.ref_fit_tangential.argtypes = [C.c_longlong, vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, vp, vp]
        L.ref_fit_values.argtypes = [C.c_longlong, vp, vp, C.c_int, C.c_double, vp, vp]
        self.L = L

    @staticmethod
    def _p(a):
        return C.c_void_p(a.ctypes.data) if a is not None else None

    def sst_local(self, gu, k, omega, grad_k, grad_omega, nu, radius, dist=None, pipe_xy=None, k_floor=1e-8,
                  omega_floor=1e-3, deviatoric=False):
        n = int(np.asarray(k).size)
        f = lambda a, s=None: None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(s or -1)
        gu, k, omega, gk, gw = f(gu, (n, 9)), f(k), f(omega), f(grad_k, (n, 3)), f(grad_omega, (n, 3))
        pipe = pipe_xy is not None
        x = y = d = None
        if pipe:
            xy = f(pipe_xy, (n, 2))
            x, y = np.ascontiguousarray(xy[:, 0]), np.ascontiguousarray(xy[:, 1])
        else:
            d = f(dist)
        out = np.empty((n, 14))
        fl = np.empty((n, 2), dtype=np.int32)
        self.L.ref_sst_local(n, 1 if pipe else 0, self._p(x), self._p(y), self._p(d), radius, nu, self._p(gu),
                             self._p(k), self._p(omega), self._p(gk), self._p(gw), k_floor, omega_floor,
                             1 if deviatoric else 0, self._p(out), self._p(fl))
        return out, fl

    def spalding(self, slip, y, nu):
        n = int(np.asarray(slip).size)
        s, yy = np.ascontiguousarray(slip, dtype=np.float64), np.ascontiguousarray(y, dtype=np.float64)
        out = np.empty((n, 3))
        ok = np.empty(n, dtype=np.int32)
        self.L.ref_spalding(n, self._p(s), self._p(yy), nu, self._p(out), self._p(ok))
        return out, ok

    def profile5(self, U, Us, y, nu):
        n = int(np.asarray(U).size // 5)
        a = [np.ascontiguousarray(v, dtype=np.float64).reshape(n, 5) for v in (U, Us, y)]
        out = np.empty((n, 6))
        valid = np.empty(n, dtype=np.int32)
        self.L.ref_profile5(n, *(self._p(v) for v in a), nu, self._p(out), self._p(valid))
        return out, valid

    def fit_values(self, U, y, nu, count):
        """Spalding fit over the active points (active::fit_values): U
        speeds and y (n,5); out (n,2) = u_tau y+, valid."""
        n = int(np.asarray(y).size // 5)
        a = [np.ascontiguousarray(v, dtype=np.float64).reshape(n, 5) for v in (U, y)]
        out = np.empty((n, 2))
        valid = np.empty(n, dtype=np.int32)
        self.L.ref_fit_values(n, *(self._p(v) for v in a), int(count), nu, self._p(out), self._p(valid))
        return out, valid

    def fit_tangential(self, samples, y, nrm, pic, nu, count, zero_limit=True):
        """Tangential Spalding fit per face (vector3::fit_active): samples
        (n,5,3), y (n,5), nrm and pic (n,3); out (n,3) = beta u_tau y+, valid."""
        n = int(np.asarray(y).size // 5)
        a = [np.ascontiguousarray(samples, dtype=np.float64).reshape(n, 5, 3),
             np.ascontiguousarray(y, dtype=np.float64).reshape(n, 5),
             np.ascontiguousarray(nrm, dtype=np.float64).reshape(n, 3),
             np.ascontiguousarray(pic, dtype=np.float64).reshape(n, 3)]
        out = np.empty((n, 3))
        valid = np.empty(n, dtype=np.int32)
        self.L.ref_fit_tangential(n, *(self._p(v) for v in a), nu, int(count), 1 if zero_limit else 0,
                                  self._p(out), self._p(valid))
        return out, valid

    def of_auto(self, inp, shared):
        n = int(np.asarray(shared).size)
        i = np.ascontiguousarray(inp, dtype=np.float64).reshape(n, 15)
        sh = np.ascontiguousarray(shared, dtype=np.int32)
        out = np.empty((n, 8))
        self.L.ref_of_auto(n, self._p(i), self._p(sh), self._p(out))
        return out


PORT_SST_LIB = Path(__file__).resolve().parent / "_ref" / "libnodals_sst_port.so"


class PortSst(RefSst):
    """Host build of the product's pointwise SST functions with the RefSst
    interface (distance form only for the local state)."""

    def __init__(self):
        L = C.CDLL(str(PORT_SST_LIB))
        vp = C.c_void_p
        L.port_sst_local.argtypes = [C.c_longlong, vp, C.c_double, C.c_double, vp, vp, vp, vp, vp, C.c_double,
                                     C.c_double, C.c_int, vp, vp]
        for f in ("port_spalding",):
            getattr(L, f).argtypes = [C.c_longlong, vp, vp, C.c_double, vp, vp]
        L.port_profile5.argtypes = [C.c_longlong, vp, vp, vp, C.c_double, vp, vp]
        L.port_of_auto.argtypes = [C.c_longlong, vp, vp, vp]
        self.L = L
        L.ref_spalding = L.port_spalding
        L.ref_profile5 = L.port_profile5
        L.ref_of_auto = L.port_of_auto
        L.port_fit_tangential.argtypes = [C.c_longlong, vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, vp, vp]
        L.ref_fit_tangential = L.port_fit_tangential

    def sst_local(self, gu, k, omega, grad_k, grad_omega, nu, radius, dist=None, pipe_xy=None, k_floor=1e-8,
                  omega_floor=1e-3, deviatoric=False):
        n = int(np.asarray(k).size)
        f = lambda a, s=None: np.ascontiguousarray(a, dtype=np.float64).reshape(s or -1)
        out = np.empty((n, 14))
        fl = np.empty((n, 2), dtype=np.int32)
        d = f(dist)
        arrs = [f(gu, (n, 9)), f(k), f(omega), f(grad_k, (n, 3)), f(grad_omega, (n, 3))]
        self.L.port_sst_local(n, self._p(d), radius, nu, *(self._p(a) for a in arrs), k_floor, omega_floor,
                              1 if deviatoric else 0, self._p(out), self._p(fl))
        return out, fl
