"""ctypes binding of libnodals_hip.so (include/nodals_hip.h).

The library is built in-tree by ``nodals_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Dict, Optional, Sequence

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libnodals_hip.so"


class NodalsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[status {code}] {msg}")
        self.code = code


class SolveResult(C.Structure):
    _fields_ = [("its", C.c_int32), ("ok", C.c_int32), ("rel", C.c_double), ("r0", C.c_double)]

    def as_dict(self) -> Dict[str, float]:
        return {"its": int(self.its), "ok": bool(self.ok), "rel": float(self.rel), "r0": float(self.r0)}


class RunResult(C.Structure):
    _fields_ = [("outer", C.c_int32), ("converged", C.c_int32), ("pass_", C.c_int32), ("reserved", C.c_int32),
                ("final_rel_cont", C.c_double), ("avg_pressure_its", C.c_double), ("pressure_drop", C.c_double),
                ("exact_pressure_drop", C.c_double), ("loop_ms", C.c_double), ("assembly_ms", C.c_double),
                ("momentum_ms", C.c_double), ("continuity_ms", C.c_double), ("pressure_ms", C.c_double),
                ("pupdate_ms", C.c_double), ("setup_ms", C.c_double), ("pressure_its", C.c_int64),
                ("warm_starts", C.c_int32), ("fallbacks", C.c_int32), ("recorded", C.c_int32),
                ("reserved2", C.c_int32), ("turb_ms", C.c_double), ("final_du", C.c_double),
                ("final_dk", C.c_double), ("final_dw", C.c_double), ("final_cont", C.c_double)]

    def as_dict(self) -> Dict[str, float]:
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


_lib: Optional[C.CDLL] = None

EXPORTS = [
    "nodals_last_error", "nodals_version",
    "nodals_h8_create", "nodals_h8_create_host", "nodals_h8_destroy", "nodals_h8_info", "nodals_h8_assemble", "nodals_h8_momentum_work",
    "nodals_h8_momentum_audit", "nodals_h8_continuity", "nodals_h8_pressure_refresh", "nodals_h8_pressure_solve",
    "nodals_h8_pressure_update", "nodals_h8_momentum_solve", "nodals_h8_pcg_amg", "nodals_h8_scalar_solve",
    "nodals_h8_run", "nodals_h8_run_window", "nodals_h8_write_vtu", "nodals_h8_get", "nodals_h8_set", "nodals_h8_fine_spmv",
    "nodals_h8_precond",
    "nodals_h8_time_kernel", "nodals_h8_probe",
    "nodals_sys_create", "nodals_sys_create_host", "nodals_sys_get", "nodals_sys_destroy", "nodals_sys_info",
    "nodals_sys_spmv", "nodals_sys_precond",
    "nodals_sys_solve", "nodals_sys_solve_timed", "nodals_sys_set_rhs", "nodals_sys_time_spmv",
    "nodals_sst_local", "nodals_wall_spalding", "nodals_wall_profile5", "nodals_wall_omega_target",
    "nodals_wall_profile5_noalloc", "nodals_wall_fit_tangential", "nodals_hex_element_blocks", "nodals_ray_trace",
    "nodals_gnn_create_for_h8", "nodals_gnn_destroy", "nodals_gnn_info", "nodals_gnn_get_params",
    "nodals_gnn_set_params", "nodals_gnn_reserve", "nodals_gnn_sample", "nodals_gnn_forward", "nodals_gnn_gradient",
    "nodals_gnn_train", "nodals_gnn_evaluate", "nodals_h8_attach_gnn", "nodals_h8_sst_volume",
    "nodals_h8_turb_init", "nodals_h8_turb_step", "nodals_h8_turb_couple",
    "nodals_h8_wall_init", "nodals_h8_wall_assemble", "nodals_h8_wall_momentum", "nodals_h8_wall_schur",
    "nodals_h8_wall_correct",
]


def lib() -> C.CDLL:
    """Load libnodals_hip.so (built in-tree).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m nodals_amd.build` (no CPU fallback exists)")
    L = C.CDLL(str(LIB_PATH))
    L.nodals_last_error.restype = C.c_char_p
    L.nodals_version.restype = C.c_char_p
    vp = C.c_void_p
    L.nodals_h8_create.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(vp)]
    L.nodals_h8_create_host.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(vp)]
    L.nodals_h8_destroy.argtypes = [vp]
    L.nodals_h8_destroy.restype = None
    L.nodals_h8_info.argtypes = [vp, C.POINTER(C.c_int64)]
    L.nodals_h8_assemble.argtypes = [vp, C.c_double, C.c_double]
    L.nodals_h8_momentum_work.argtypes = [vp, C.c_char_p, C.c_double, C.c_int32]
    L.nodals_h8_momentum_audit.argtypes = [vp, C.POINTER(C.c_double)]
    L.nodals_h8_continuity.argtypes = [vp, C.POINTER(C.c_double)]
    L.nodals_h8_pressure_refresh.argtypes = [vp]
    L.nodals_h8_pressure_solve.argtypes = [vp, C.c_char_p, C.c_double, C.c_double, C.c_int32, C.POINTER(SolveResult)]
    L.nodals_h8_pressure_update.argtypes = [vp, C.c_double]
    L.nodals_h8_momentum_solve.argtypes = [vp, C.c_char_p, C.c_int32, C.c_double, C.POINTER(SolveResult)]
    L.nodals_h8_pcg_amg.argtypes = [vp, vp, vp, C.c_double, C.c_double, C.c_int32, C.POINTER(SolveResult)]
    L.nodals_h8_scalar_solve.argtypes = [vp, C.c_int32, C.POINTER(SolveResult)]
    L.nodals_h8_run.argtypes = [vp, C.POINTER(RunResult)]
    L.nodals_h8_run_window.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(RunResult)]
    L.nodals_h8_write_vtu.argtypes = [vp, C.c_char_p]
    L.nodals_h8_get.argtypes = [vp, C.c_char_p, vp, C.c_int64]
    L.nodals_h8_set.argtypes = [vp, C.c_char_p, vp, C.c_int64]
    L.nodals_h8_fine_spmv.argtypes = [vp, vp, vp]
    L.nodals_h8_precond.argtypes = [vp, vp, vp]
    L.nodals_h8_time_kernel.argtypes = [vp, C.c_char_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.nodals_h8_probe.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.nodals_sst_local.argtypes = [C.c_int64, vp, vp, C.c_double, C.c_double, vp, vp, vp, vp, vp, C.c_double,
                                   C.c_double, C.c_int32, vp, vp]
    L.nodals_wall_spalding.argtypes = [C.c_int64, vp, vp, C.c_double, vp, vp]
    L.nodals_wall_profile5.argtypes = [C.c_int64, vp, vp, vp, C.c_double, vp, vp]
    L.nodals_wall_omega_target.argtypes = [C.c_int64, vp, vp, vp]
    L.nodals_wall_profile5_noalloc.argtypes = [C.c_int64, vp, vp, vp, C.c_double, vp, vp]
    L.nodals_wall_fit_tangential.argtypes = [C.c_int64, vp, vp, vp, vp, C.c_double, C.c_int32, C.c_int32, vp, vp]
    L.nodals_hex_element_blocks.argtypes = [C.c_int64, vp, vp, C.c_double, vp, vp, vp]
    L.nodals_ray_trace.argtypes = [C.c_int64, vp, C.c_int64, vp, C.c_int64, vp, vp, vp, C.c_int32, vp, C.c_int32,
                                   C.c_double, vp, vp, vp, vp, vp]
    L.nodals_gnn_create_for_h8.argtypes = [vp, C.c_uint32, C.POINTER(vp)]
    L.nodals_gnn_destroy.argtypes = [vp]
    L.nodals_gnn_info.argtypes = [vp, vp]
    L.nodals_gnn_get_params.argtypes = [vp, vp, C.c_int64]
    L.nodals_gnn_set_params.argtypes = [vp, vp, C.c_int64]
    L.nodals_gnn_reserve.argtypes = [vp, C.c_int32]
    L.nodals_gnn_sample.argtypes = [vp, C.c_int32, vp, vp]
    L.nodals_gnn_forward.argtypes = [vp, vp, vp]
    L.nodals_gnn_gradient.argtypes = [vp, vp, vp, vp, C.POINTER(C.c_double)]
    L.nodals_gnn_train.argtypes = [vp, C.c_int32, C.c_float, C.c_uint32, C.POINTER(C.c_double)]
    L.nodals_gnn_evaluate.argtypes = [vp, C.POINTER(C.c_double)]
    L.nodals_h8_attach_gnn.argtypes = [vp, vp, C.c_int32, C.c_int32]
    L.nodals_h8_sst_volume.argtypes = [vp, C.c_int32, vp, vp, vp, vp, C.c_double, C.c_double, C.c_int32, vp, vp]
    L.nodals_h8_turb_init.argtypes = [vp, C.c_int, vp]
    L.nodals_h8_turb_step.argtypes = [vp, vp]
    L.nodals_h8_turb_couple.argtypes = [vp, C.c_int32]
    L.nodals_h8_wall_init.argtypes = [vp, C.c_int, vp]
    L.nodals_h8_wall_assemble.argtypes = [vp]
    L.nodals_h8_wall_momentum.argtypes = [vp]
    L.nodals_h8_wall_schur.argtypes = [vp, vp, vp]
    L.nodals_h8_wall_correct.argtypes = [vp, vp]
    L.nodals_sys_create.argtypes = [C.c_int32, vp, vp, vp, C.c_char_p, C.POINTER(vp)]
    L.nodals_sys_create_host.argtypes = [C.c_int32, vp, vp, vp, C.c_char_p, C.POINTER(vp)]
    L.nodals_sys_get.argtypes = [vp, C.c_char_p, vp, C.c_int64]
    L.nodals_sys_destroy.argtypes = [vp]
    L.nodals_sys_destroy.restype = None
    L.nodals_sys_info.argtypes = [vp, C.POINTER(C.c_int64)]
    L.nodals_sys_spmv.argtypes = [vp, vp, vp]
    L.nodals_sys_precond.argtypes = [vp, vp, vp]
    L.nodals_sys_solve.argtypes = [vp, vp, vp, C.c_char_p, C.c_double, C.c_int32, C.c_int32, C.POINTER(SolveResult)]
    L.nodals_sys_solve_timed.argtypes = [vp, C.c_char_p, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                         C.POINTER(C.c_double), C.POINTER(SolveResult)]
    L.nodals_sys_set_rhs.argtypes = [vp, vp]
    L.nodals_sys_time_spmv.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise NodalsError(rc, lib().nodals_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


_DTYPES = {
    "mom_row": np.int64, "mom_col": np.int32, "color": np.int32, "g2free": np.int32, "fine_row": np.int32,
    "fine_col": np.int32, "cell_ref": np.int32, "cell_slot": np.uint8, "pin_cell": np.int32, "fine_lanes": np.int32,
    "mesh_tets": np.int32, "mesh_opp": np.int32, "mesh_faces": np.int32, "mesh_owner": np.int32,
    "mesh_neighbour": np.int32, "mesh_face_patch": np.int32, "mesh_sizes": np.int64, "fixed_slot": np.int32,
}


class TurbStepResult(C.Structure):
    _fields_ = [("k_clips", C.c_int32), ("omega_clips", C.c_int32), ("wall_invalid", C.c_int32),
                ("reserved", C.c_int32), ("ms", C.c_double)]

    def as_dict(self) -> Dict[str, float]:
        return {"k_clips": self.k_clips, "omega_clips": self.omega_clips, "wall_invalid": self.wall_invalid,
                "ms": self.ms}


class H8:
    """H8 laminar SIMPLE solver context on the GPU (argv as the H8 binary)."""

    def __init__(self, argv: Sequence[str], device: bool = True):
        L = lib()
        args = ["nodals_gpu_h8"] + [str(a) for a in argv]
        arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
        h = C.c_void_p()
        create = L.nodals_h8_create if device else L.nodals_h8_create_host
        _check(create(len(args), arr, C.byref(h)))
        self._h = h
        info = (C.c_int64 * 8)()
        _check(L.nodals_h8_info(self._h, info))
        self.info = {k: int(v) for k, v in zip(
            ["cells", "nv", "mom_nnz", "colors", "fine_nnz", "amg_levels", "terminal", "nfixed"], info)}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().nodals_h8_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # solver calls -------------------------------------------------------
    def assemble(self, alpha_u: float, rau_scale: float = 1.0) -> None:
        _check(lib().nodals_h8_assemble(self._h, alpha_u, rau_scale))

    def momentum_work(self, mode: str = "fgs1", omega: float = 1.0, outer: int = 1) -> None:
        _check(lib().nodals_h8_momentum_work(self._h, mode.encode(), omega, outer))

    def momentum_audit(self) -> np.ndarray:
        r = (C.c_double * 3)()
        _check(lib().nodals_h8_momentum_audit(self._h, r))
        return np.array(r[:])

    def continuity(self) -> float:
        n = C.c_double()
        _check(lib().nodals_h8_continuity(self._h, C.byref(n)))
        return n.value

    def pressure_refresh(self) -> None:
        _check(lib().nodals_h8_pressure_refresh(self._h))

    def pressure_solve(self, solver: str, rtol: float, atol: float, maxit: int) -> Dict[str, float]:
        r = SolveResult()
        _check(lib().nodals_h8_pressure_solve(self._h, solver.encode(), rtol, atol, maxit, C.byref(r)))
        return r.as_dict()

    def pressure_update(self, alpha_p: float) -> None:
        _check(lib().nodals_h8_pressure_update(self._h, alpha_p))

    def momentum_solve(self, mode: str, sweeps: int, omega: float) -> Dict[str, float]:
        r = SolveResult()
        _check(lib().nodals_h8_momentum_solve(self._h, mode.encode(), int(sweeps), float(omega), C.byref(r)))
        return r.as_dict()

    def pcg_amg(self, rhs, rtol: float, atol: float, maxit: int):
        """(x, result) of PCG-AMG on the live pressure operator."""
        b = _f64(rhs)
        x = np.empty_like(b)
        r = SolveResult()
        _check(lib().nodals_h8_pcg_amg(self._h, _ptr(b), _ptr(x), rtol, atol, int(maxit), C.byref(r)))
        return x, r.as_dict()

    def scalar_solve(self, equation: int) -> Dict[str, float]:
        r = SolveResult()
        _check(lib().nodals_h8_scalar_solve(self._h, int(equation), C.byref(r)))
        return r.as_dict()

    def run(self) -> Dict[str, float]:
        r = RunResult()
        _check(lib().nodals_h8_run(self._h, C.byref(r)))
        return r.as_dict()

    def run_window(self, first_it: int, count: int) -> Dict[str, float]:
        """Wall-function path: `count` iterations numbered from first_it,
        continuing from the current state."""
        r = RunResult()
        _check(lib().nodals_h8_run_window(self._h, int(first_it), int(count), C.byref(r)))
        return r.as_dict()

    def write_vtu(self, path) -> None:
        _check(lib().nodals_h8_write_vtu(self._h, str(path).encode()))

    def fine_spmv(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(lib().nodals_h8_fine_spmv(self._h, _ptr(x), _ptr(y)))
        return y

    def precond(self, r: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        _check(lib().nodals_h8_precond(self._h, _ptr(r), _ptr(z)))
        return z

    def time_kernel(self, name: str, reps: int):
        ms, by = C.c_double(), C.c_double()
        _check(lib().nodals_h8_time_kernel(self._h, name.encode(), reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def sst_volume(self, equation: int, k_free, k_fixed, omega_free, omega_fixed, k_floor=1e-8,
                   omega_floor=1e-3, deviatoric=False):
        """k-omega SST volume terms of every cell: (mat[nc,8,8], rhs[nc,8])."""
        nc = self.info["cells"]
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (k_free, k_fixed, omega_free, omega_fixed)]
        mat = np.empty((nc, 8, 8))
        rhs = np.empty((nc, 8))
        _check(lib().nodals_h8_sst_volume(self._h, equation, *(_ptr(v) for v in a), k_floor, omega_floor,
                                          1 if deviatoric else 0, _ptr(mat), _ptr(rhs)))
        return mat, rhs

    def turb_init(self, args: Sequence[str] = ()) -> None:
        """Set up the k-omega transport step (reference SST flags, e.g.
        ["--sst-alpha-k", "0.3"]); k and omega start at their inlet values."""
        a = [str(x).encode() for x in args]
        arr = (C.c_char_p * max(len(a), 1))(*a) if a else None
        _check(lib().nodals_h8_turb_init(self._h, len(a), arr))
        st = self.get("turb_inlet_state", 7)
        self.turb = {"k_in": st[0], "omega_in": st[1], "inlet_faces": int(st[2]), "wall_faces": int(st[3]),
                     "ns": int(st[4]), "nvert": int(st[5]), "colors": int(st[6])}

    def turb_couple(self, on: bool = True) -> None:
        """Coupled SST mode of run(): eddy viscosity in the momentum
        diffusion and one k-omega step after every pressure update."""
        _check(lib().nodals_h8_turb_couple(self._h, 1 if on else 0))

    def turb_step(self) -> Dict[str, float]:
        r = TurbStepResult()
        _check(lib().nodals_h8_turb_step(self._h, C.byref(r)))
        return r.as_dict()

    def wall_init(self, args: Sequence[str] = ()) -> None:
        """Wall-function momentum path (duct created with --wall-model
        spalding, after turb_init); args are its flags, e.g. ["--m3-sweeps",
        "20"].  Afterwards run() is the wall-function SIMPLE loop."""
        a = [str(x).encode() for x in args]
        arr = (C.c_char_p * max(len(a), 1))(*a) if a else None
        _check(lib().nodals_h8_wall_init(self._h, len(a), arr))

    def wall_assemble(self) -> None:
        _check(lib().nodals_h8_wall_assemble(self._h))

    def wall_momentum(self) -> None:
        _check(lib().nodals_h8_wall_momentum(self._h))

    def wall_schur(self, x) -> np.ndarray:
        nc = self.info["cells"]
        x = _f64(x, (nc,))
        y = np.empty(nc)
        _check(lib().nodals_h8_wall_schur(self._h, _ptr(x), _ptr(y)))
        return y

    def wall_correct(self, dp) -> None:
        nc = self.info["cells"]
        d = _f64(dp, (nc,))
        _check(lib().nodals_h8_wall_correct(self._h, _ptr(d)))

    def gnn_graph(self):
        """(off, nbr) of the face-adjacent cell graph the network runs on."""
        nc = self.info["cells"]
        off = self.get("cell_nbr_off", nc + 1, np.int32)
        return off, self.get("cell_nbr", int(off[-1]), np.int32)

    def attach_gnn(self, gnn, warm_start=True, record_every: int = 0) -> None:
        """warm_start: True or "guess" (the prediction, at its residual-optimal
        multiple, as the initial guess), "direction" (as an added first FGMRES
        direction), False (off)."""
        mode = {True: 1, "guess": 1, "direction": 2, False: 0, None: 0}[warm_start]
        _check(lib().nodals_h8_attach_gnn(self._h, gnn._h if gnn is not None else None, mode, record_every))

    def probe_start(self, kind: int = 1):
        """kind 1: fine Schur SpMV launches; kind 2: momentum Gauss-Seidel
        colour launches."""
        _check(lib().nodals_h8_probe(self._h, int(kind), None, None))

    def probe_stop(self):
        """(launches, summed ms) of the probed launches since probe_start."""
        n, ms = C.c_int64(), C.c_double()
        _check(lib().nodals_h8_probe(self._h, 0, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    # host access --------------------------------------------------------
    def get(self, name: str, count: int, dtype=None) -> np.ndarray:
        dt = dtype or _DTYPES.get(name, np.float64)
        out = np.empty(count, dtype=dt)
        _check(lib().nodals_h8_get(self._h, name.encode(), _ptr(out), count))
        return out

    def set(self, name: str, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(lib().nodals_h8_set(self._h, name.encode(), _ptr(v), v.size))


class System:
    """Explicit FP64 SPD CSR system with a device SA hierarchy."""

    def __init__(self, row: np.ndarray, col: np.ndarray, val: np.ndarray, options: str = "", device: bool = True):
        self._row = np.ascontiguousarray(row, dtype=np.int64)
        self._col = np.ascontiguousarray(col, dtype=np.int32)
        self._val = np.ascontiguousarray(val, dtype=np.float64)
        n = self._row.size - 1
        h = C.c_void_p()
        create = lib().nodals_sys_create if device else lib().nodals_sys_create_host
        _check(create(n, _ptr(self._row), _ptr(self._col), _ptr(self._val), options.encode(), C.byref(h)))
        self._h = h
        self.n = n
        info = (C.c_int64 * 16)()
        _check(lib().nodals_sys_info(self._h, info))
        self.levels = int(info[2])
        self.terminal = int(info[3])
        self.level_rows = [int(info[4 + l]) for l in range(min(self.levels, 12))]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().nodals_sys_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def spmv(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(lib().nodals_sys_spmv(self._h, _ptr(x), _ptr(y)))
        return y

    def get(self, name: str, count: int, dtype=np.float64) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        _check(lib().nodals_sys_get(self._h, name.encode(), _ptr(out), count))
        return out

    def hierarchy(self) -> dict:
        """Host SA hierarchy as scipy CSR matrices (oracle V-cycle input)."""
        import scipy.sparse as sp

        H = {"P": [], "A": []}
        nfine = self.n
        for l in range(self.levels):
            pr = self.get(f"P{l}_row", nfine + 1, np.int64)
            pc = self.get(f"P{l}_col", int(pr[-1]), np.int32)
            pv = self.get(f"P{l}_val", int(pr[-1]))
            nc = self.level_rows[l]
            ar = self.get(f"A{l}_row", nc + 1, np.int64)
            ac = self.get(f"A{l}_col", int(ar[-1]), np.int32)
            av = self.get(f"A{l}_val", int(ar[-1]))
            H["P"].append(sp.csr_matrix((pv, pc, pr), shape=(nfine, nc)))
            H["A"].append(sp.csr_matrix((av, ac, ar), shape=(nc, nc)))
            nfine = nc
        H["Tinv"] = self.get("terminal_inv", nfine * nfine).reshape(nfine, nfine)
        H["fine_lambda"] = float(self.get("fine_lambda", 1)[0])
        H["level_lambda"] = self.get("level_lambda", self.levels)
        return H

    def precond(self, r: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        _check(lib().nodals_sys_precond(self._h, _ptr(r), _ptr(z)))
        return z

    def solve(self, b: np.ndarray, solver: str = "fgmres", rtol: float = 1e-8, restart: int = 30, maxit: int = 500):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        r = SolveResult()
        _check(lib().nodals_sys_solve(self._h, _ptr(b), _ptr(x), solver.encode(), rtol, restart, maxit, C.byref(r)))
        return x, r.as_dict()

    def set_rhs(self, b: np.ndarray) -> None:
        b = np.ascontiguousarray(b, dtype=np.float64)
        _check(lib().nodals_sys_set_rhs(self._h, _ptr(b)))

    def solve_timed(self, solver: str = "fgmres", rtol: float = 1e-8, restart: int = 30, maxit: int = 500,
                    reps: int = 1):
        ms = C.c_double()
        r = SolveResult()
        _check(lib().nodals_sys_solve_timed(self._h, solver.encode(), rtol, restart, maxit, reps, C.byref(ms),
                                            C.byref(r)))
        return ms.value, r.as_dict()

    def time_spmv(self, level: int, reps: int):
        ms, by = C.c_double(), C.c_double()
        _check(lib().nodals_sys_time_spmv(self._h, level, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value


# ---------------------------------------------------------------- SST / wall
def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


SST_FIELDS = ("nut", "F1", "F2", "sigmaK", "sigmaW", "beta", "gamma", "strain", "pk", "srcOmega", "reactK",
              "reactOmega", "omegaProduction", "omegaCross")


def sst_local(gu, k, omega, grad_k, grad_omega, nu, radius, dist=None, pipe_xy=None, k_floor=1e-8,
              omega_floor=1e-3, deviatoric=False):
    """SST-2003 local state at n points on the GPU: (out[n,14], floored[n,2])."""
    n = int(np.asarray(k).size)
    gu, k, omega = _f64(gu, (n, 9)), _f64(k), _f64(omega)
    gk, gw = _f64(grad_k, (n, 3)), _f64(grad_omega, (n, 3))
    d = None if dist is None else _f64(dist)
    xy = None if pipe_xy is None else _f64(pipe_xy, (n, 2))
    out = np.empty((n, 14))
    fl = np.empty((n, 2), dtype=np.int32)
    _check(lib().nodals_sst_local(n, None if d is None else _ptr(d), None if xy is None else _ptr(xy), radius, nu,
                                  _ptr(gu), _ptr(k), _ptr(omega), _ptr(gk), _ptr(gw), k_floor, omega_floor,
                                  1 if deviatoric else 0, _ptr(out), _ptr(fl)))
    return out, fl


def wall_spalding(slip, y, nu):
    n = int(np.asarray(slip).size)
    s, yy = _f64(slip), _f64(y)
    out = np.empty((n, 3))
    ok = np.empty(n, dtype=np.int32)
    _check(lib().nodals_wall_spalding(n, _ptr(s), _ptr(yy), nu, _ptr(out), _ptr(ok)))
    return out, ok


def wall_profile5(U, Us, y, nu):
    n = int(np.asarray(U).size // 5)
    U, Us, y = _f64(U, (n, 5)), _f64(Us, (n, 5)), _f64(y, (n, 5))
    out = np.empty((n, 6))
    valid = np.empty(n, dtype=np.int32)
    _check(lib().nodals_wall_profile5(n, _ptr(U), _ptr(Us), _ptr(y), nu, _ptr(out), _ptr(valid)))
    return out, valid


def wall_profile5_noalloc(U, Us, y, nu):
    """wall_profile5 through the allocation-free entry point (n <= 65536)."""
    n = int(np.asarray(U).size // 5)
    U, Us, y = _f64(U, (n, 5)), _f64(Us, (n, 5)), _f64(y, (n, 5))
    out = np.empty((n, 6))
    valid = np.empty(n, dtype=np.int32)
    _check(lib().nodals_wall_profile5_noalloc(n, _ptr(U), _ptr(Us), _ptr(y), nu, _ptr(out), _ptr(valid)))
    return out, valid


def wall_fit_tangential(samples, y, nrm, pic, nu, count, zero_limit=True):
    """Tangential Spalding fit per face on the GPU: samples (n,5,3), y (n,5),
    nrm and pic (n,3); returns out (n,3) = beta u_tau y+ and valid (n,)."""
    n = int(np.asarray(y).size // 5)
    s, yy, nn, pp = _f64(samples, (n, 5, 3)), _f64(y, (n, 5)), _f64(nrm, (n, 3)), _f64(pic, (n, 3))
    out = np.empty((n, 3))
    valid = np.empty(n, dtype=np.int32)
    _check(lib().nodals_wall_fit_tangential(n, _ptr(s), _ptr(yy), _ptr(nn), _ptr(pp), nu, int(count),
                                            1 if zero_limit else 0, _ptr(out), _ptr(valid)))
    return out, valid


def hex_element_blocks(coords, u, nu):
    """Q1+BF2 hexahedral diffusion and convection blocks on the GPU:
    coords (n,8,3), u (n,14,3) -> (diff (n,14,14), conv (n,14,14), bad)."""
    n = int(np.asarray(coords).size // 24)
    X, U = _f64(coords, (n, 8, 3)), _f64(u, (n, 14, 3))
    D = np.empty((n, 14, 14))
    Cv = np.empty((n, 14, 14))
    bad = C.c_int32(0)
    _check(lib().nodals_hex_element_blocks(n, _ptr(X), _ptr(U), float(nu), _ptr(D), _ptr(Cv), C.byref(bad)))
    return D, Cv, int(bad.value)


def ray_trace(points, tets, start, origin, direction, dist, max_crossings=4096, tol=1e-9):
    """Host-only wall-normal segment traversal (no GPU): rays r from
    origin[r] in cell start[r] along direction[r], sampled at dist[r][:].
    Returns dict(cell (nr,nd), bary (nr,nd,4), crossings (nr,nd), status
    (nr,), certified (nr,))."""
    P = _f64(points, (-1, 3))
    T = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
    st = np.ascontiguousarray(start, dtype=np.int32).ravel()
    nr = st.size
    O, D = _f64(origin, (nr, 3)), _f64(direction, (nr, 3))
    dist = np.asarray(dist, dtype=np.float64)
    nd = dist.size // max(nr, 1) if nr else max(dist.size, 1)
    dist = _f64(dist, (nr, nd)) if nr else np.zeros((0, nd))
    cell = np.empty((nr, nd), np.int32)
    bary = np.empty((nr, nd, 4))
    cross = np.empty((nr, nd), np.int32)
    status = np.empty(nr, np.int32)
    cert = np.empty(nr)
    _check(lib().nodals_ray_trace(P.shape[0], _ptr(P), T.shape[0], _ptr(T), nr, _ptr(st), _ptr(O), _ptr(D), nd,
                                  _ptr(dist), int(max_crossings), float(tol), _ptr(cell), _ptr(bary), _ptr(cross),
                                  _ptr(status), _ptr(cert)))
    return {"cell": cell, "bary": bary, "crossings": cross, "status": status, "certified": cert}


def wall_omega_target(inp, shared):
    n = int(np.asarray(shared).size)
    inp = _f64(inp, (n, 15))
    sh = np.ascontiguousarray(shared, dtype=np.int32)
    out = np.empty((n, 8))
    _check(lib().nodals_wall_omega_target(n, _ptr(inp), _ptr(sh), _ptr(out)))
    return out


# ---------------------------------------------------------------- learned warm start
class Gnn:
    """Message-passing pressure-correction network on an H8 context's cell graph."""

    IN = 16

    def __init__(self, h8: "H8", seed: int = 1):
        self._h = C.c_void_p()
        _check(lib().nodals_gnn_create_for_h8(h8._h, seed, C.byref(self._h)))
        info = np.zeros(4, dtype=np.int64)
        _check(lib().nodals_gnn_info(self._h, _ptr(info)))
        self.n, self.rows, self.nparams = int(info[0]), int(info[1]), int(info[2])

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().nodals_gnn_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def samples(self) -> int:
        info = np.zeros(4, dtype=np.int64)
        _check(lib().nodals_gnn_info(self._h, _ptr(info)))
        return int(info[3])

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        _check(lib().nodals_gnn_get_params(self._h, _ptr(out), self.nparams))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().nodals_gnn_set_params(self._h, _ptr(p), p.size))

    def reserve(self, capacity: int) -> None:
        _check(lib().nodals_gnn_reserve(self._h, capacity))

    def sample(self, i: int):
        x = np.empty((self.n, self.IN), dtype=np.float32)
        t = np.empty(self.n, dtype=np.float32)
        _check(lib().nodals_gnn_sample(self._h, i, _ptr(x), _ptr(t)))
        return x, t

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.n, self.IN)
        y = np.empty(self.n, dtype=np.float32)
        _check(lib().nodals_gnn_forward(self._h, _ptr(x), _ptr(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.n, self.IN)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = C.c_double()
        _check(lib().nodals_gnn_gradient(self._h, _ptr(x), _ptr(t), _ptr(g), C.byref(loss)))
        return g, loss.value

    def train(self, epochs: int, lr: float = 1e-3, seed: int = 1) -> float:
        loss = C.c_double()
        _check(lib().nodals_gnn_train(self._h, epochs, lr, seed, C.byref(loss)))
        return loss.value

    def evaluate(self) -> float:
        loss = C.c_double()
        _check(lib().nodals_gnn_evaluate(self._h, C.byref(loss)))
        return loss.value

