"""CPU oracle (test infrastructure only).

ctypes bindings for oracle/build/liboracle.so — the FP64 restatement of the
reference's CPU path in dg_oracle.cc — and for the reference's own kernel
files built as CPU libraries under oracle/_ref/ (one per physics
configuration, built from /root/reference by oracle/Makefile when the sources
are present).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this package, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_SRC = "/root/reference/code/src"

_c_dp = ctypes.POINTER(ctypes.c_double)
_c_ip = ctypes.POINTER(ctypes.c_int)


class OrCfg(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("D", "N_F", "physics", "flux", "N_s", "N_G", "M_s", "M_G", "N_N", "N_E",
                                           "M_T", "N_s1D")] + [("gx", ctypes.c_double), ("gy", ctypes.c_double)]


_TAB_F = ["phi", "dphi", "phi_w", "dphi_w", "psi", "psi_w", "J", "invJac", "JF", "normals", "Minv"]
_TAB_I = ["map", "invmap", "boundaryMap", "neighbors"]
_TAB_F2 = ["Lag2Mono", "Mono2Lag", "Lag2MonoX", "MonoX2MonoY", "MonoY2Lag", "Lag2MonoY", "MonoX2Lag"]


class OrTables(ctypes.Structure):
    _fields_ = ([(n, _c_dp) for n in _TAB_F] + [(n, _c_ip) for n in _TAB_I] + [("rflctiveIdx", ctypes.c_int)] +
                [(n, _c_dp) for n in _TAB_F2])


_lib = None


def build(quiet: bool = True) -> None:
    """Compile the oracle (and the reference libraries when the sources exist)."""
    r = subprocess.run(["make", "-C", HERE, "-j8"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        cp, tp = ctypes.POINTER(OrCfg), ctypes.POINTER(OrTables)
        vp = ctypes.c_void_p
        L.or_rhs.argtypes = [cp, tp, vp, vp]
        L.or_dt_from_cfl.argtypes = [cp, vp, ctypes.c_double]
        L.or_dt_from_cfl.restype = ctypes.c_double
        L.or_limit_hr.argtypes = [cp, tp, vp]
        L.or_sensing.argtypes = [cp, vp, vp, vp, vp, vp, vp]
        L.or_limit_hri.argtypes = [cp, tp, vp, vp, vp, vp]
        L.or_limit_flagged.argtypes = [cp, tp, vp, vp]
        L.or_rk4.argtypes = [cp, tp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                             vp, vp, vp, vp, ctypes.c_int]
        L.or_rk4_out.argtypes = [cp, tp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, vp, ctypes.c_int,
                                 ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, vp]
        L.or_dgemm.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_double, vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_double, vp, ctypes.c_int]
        for n in ("or_mapToFace", "or_evaluate_sf", "or_evaluate_q", "or_redistribute_sf", "or_redistribute_q",
                  "or_mapToElement", "or_addSFQ", "or_findUPA", "or_calc_sensors", "or_copy_detected", "or_hri1D",
                  "or_rflctive", "or_solve"):
            getattr(L, n).restype = None
        L.or_solve.argtypes = [cp, ctypes.c_double, vp, vp, vp]
        L.or_hrl1D.argtypes = [ctypes.c_int] * 5 + [vp, ctypes.c_int, vp, vp]
        L.or_format.argtypes = [cp, vp, vp, ctypes.c_int]
        L.or_format.restype = None
        L.or_limit_m2.argtypes = [cp, tp, vp]
        L.or_limit_m2.restype = None
        L.or_limit_pr.argtypes = [cp, tp, vp]
        L.or_limit_pr.restype = None
        L.or_limit_pri.argtypes = [cp, tp, vp, vp, vp, vp]
        L.or_limit_pri.restype = None
        L.or_limit_m2i.argtypes = [cp, tp, vp, vp, vp, vp]
        L.or_limit_m2i.restype = None
        L.or_m2i1D.argtypes = [cp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp]
        L.or_m2i1D.restype = None
        L.or_limit_p0i.argtypes = [cp, tp, vp, vp, vp, vp]
        L.or_limit_p0i.restype = None
        L.or_p0i.argtypes = [cp, vp, ctypes.c_double, vp, vp, vp]
        L.or_p0i.restype = None
        i = ctypes.c_int
        L.or_stridedcopy.argtypes = [i, i, i, i, i, i, vp, vp]
        L.or_reconstruct_energy.argtypes = [i, i, i, vp, vp, vp, vp]
        L.or_internal_energy.argtypes = [i, i, i, vp, vp, vp, vp]
        L.or_pressure.argtypes = [cp, vp, vp]
        L.or_kinetic_energy.argtypes = [i, i, vp, vp, vp, vp]
        L.or_average_cell_p0.argtypes = [cp, vp]
        L.or_change_basis.argtypes = [cp, i, i, vp, vp, vp]
        L.or_cons_prim.argtypes = [cp, vp, i]
        L.or_format_sensor.argtypes = [i, i, vp, vp]
        for n in ("or_stridedcopy", "or_reconstruct_energy", "or_internal_energy", "or_pressure", "or_kinetic_energy",
                  "or_average_cell_p0", "or_change_basis", "or_cons_prim", "or_format_sensor"):
            getattr(L, n).restype = None
        L.or_conservation.argtypes = [cp, vp, vp, vp, vp, vp]
        L.or_conservation.restype = None
        L.or_ref_area.argtypes = [cp]
        L.or_ref_area.restype = ctypes.c_double
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.isdir(REF_DIR) and any(f.endswith(".so") for f in os.listdir(REF_DIR))


def ref_name(dim: int, physics: int, flux: int) -> str:
    return f"{dim}d_{('sf', 'mf', 'pa', 'st')[physics]}_{'roe' if flux else 'rus'}"


_refs: dict = {}


def ref_lib(dim: int, physics: int, flux: int):
    """The reference's launchers built as CPU code for this configuration, or None."""
    name = ref_name(dim, physics, flux)
    if name not in _refs:
        path = os.path.join(REF_DIR, f"libdgref_{name}.so")
        _refs[name] = ctypes.CDLL(path, mode=os.RTLD_LOCAL) if os.path.exists(path) else None
    return _refs[name]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Problem:
    """Oracle view of one problem: configuration + tables (kept alive here)."""

    def __init__(self, tables: dict, physics: int, flux: int):
        self.t = tables
        D = tables["D"]
        self.nf = 2 + D + (0, 1, 2, 2)[physics]
        c = OrCfg()
        c.D, c.N_F, c.physics, c.flux = D, self.nf, physics, flux
        for n in ("N_s", "N_G", "M_s", "M_G", "N_N", "N_E", "M_T", "N_s1D"):
            setattr(c, n, int(tables[n]))
        c.gx = c.gy = 0.0
        self.cfg = c
        self.physics, self.flux, self.D = physics, flux, D
        self._keep = {}
        tb = OrTables()
        for n in _TAB_F + _TAB_F2:
            a = np.ascontiguousarray(tables.get(n, np.zeros(1)), dtype=np.float64)
            if a.size == 0:
                a = np.zeros(1)
            self._keep[n] = a
            setattr(tb, n, a.ctypes.data_as(_c_dp))
        for n in _TAB_I:
            a = np.ascontiguousarray(tables.get(n, np.zeros(1, np.int32)), dtype=np.int32)
            if a.size == 0:
                a = np.zeros(1, np.int32)
            self._keep[n] = a
            setattr(tb, n, a.ctypes.data_as(_c_ip))
        tb.rflctiveIdx = int(tables["rflctiveIdx"])
        self.tab = tb

    @property
    def state_size(self) -> int:
        return self.cfg.N_E * self.nf * self.cfg.N_s

    def rhs(self, U: np.ndarray) -> np.ndarray:
        U = np.ascontiguousarray(U, dtype=np.float64)
        f = np.zeros_like(U)
        lib().or_rhs(ctypes.byref(self.cfg), ctypes.byref(self.tab), _dptr(U), _dptr(f))
        return f

    def dt_from_cfl(self, U: np.ndarray, cfl: float) -> float:
        U = np.ascontiguousarray(U, dtype=np.float64)
        return lib().or_dt_from_cfl(ctypes.byref(self.cfg), _dptr(U), cfl)

    def solve(self, dt: float, f: np.ndarray) -> np.ndarray:
        du = np.zeros_like(f)
        lib().or_solve(ctypes.byref(self.cfg), dt, _dptr(self._keep["Minv"]), _dptr(np.ascontiguousarray(f)), _dptr(du))
        return du

    def limit_m2(self, U: np.ndarray) -> np.ndarray:
        """Limiting::M2limiting (two-fluid)."""
        U = np.array(U, dtype=np.float64, copy=True)
        lib().or_limit_m2(ctypes.byref(self.cfg), ctypes.byref(self.tab), _dptr(U))
        return U

    def limit_pr(self, U: np.ndarray) -> np.ndarray:
        """Limiting::PRlimiting (two-fluid)."""
        U = np.array(U, dtype=np.float64, copy=True)
        lib().or_limit_pr(ctypes.byref(self.cfg), ctypes.byref(self.tab), _dptr(U))
        return U

    def limit_pri(self, U: np.ndarray, thresholds):
        """Limiting::PRIlimiting (two-fluid); returns (U, sensors)."""
        U = np.array(U, dtype=np.float64, copy=True)
        on, th = self._thresh(thresholds)
        sens = np.zeros(self.cfg.N_E, dtype=np.int32)
        lib().or_limit_pri(ctypes.byref(self.cfg), ctypes.byref(self.tab), _dptr(on), _dptr(th), _dptr(U), _dptr(sens))
        return U, sens

    def limit_m2i(self, U: np.ndarray, thresholds):
        """Limiting::M2Ilimiting (two-fluid); returns (U, sensors)."""
        U = np.array(U, dtype=np.float64, copy=True)
        on, th = self._thresh(thresholds)
        sens = np.zeros(self.cfg.N_E, dtype=np.int32)
        lib().or_limit_m2i(ctypes.byref(self.cfg), ctypes.byref(self.tab), _dptr(on), _dptr(th), _dptr(U), _dptr(sens))
        return U, sens

    def limit_p0i(self, U: np.ndarray, thresholds):
        """Limiting::P0Ilimiting (two-fluid sensor); returns (U, sensors)."""
        U = np.array(U, dtype=np.float64, copy=True)
        on, th = self._thresh(thresholds)
        sens = np.zeros(self.cfg.N_E, dtype=np.int32)
        lib().or_limit_p0i(ctypes.byref(self.cfg), ctypes.byref(self.tab), _dptr(on), _dptr(th), _dptr(U), _dptr(sens))
        return U, sens

    def conservation(self, U: np.ndarray) -> np.ndarray:
        """DG_SOLVER::conservation: the integral of each field over the mesh."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        w = np.ascontiguousarray(self.t["weight"], dtype=np.float64)
        out = np.zeros(self.nf)
        lib().or_conservation(ctypes.byref(self.cfg), _dptr(self._keep["phi"]), _dptr(self._keep["J"]), _dptr(w),
                              _dptr(U), _dptr(out))
        return out

    def format(self, U: np.ndarray) -> np.ndarray:
        """PRINTER formatter: the printed primitive fields of U, node by node."""
        U = np.array(U, dtype=np.float64, copy=True)
        out = np.zeros_like(U)
        lib().or_format(ctypes.byref(self.cfg), _dptr(U), _dptr(out), 0)
        return out

    def unformat(self, out: np.ndarray) -> np.ndarray:
        """Inverse formatter (restart): the conserved state from printed fields."""
        out = np.array(out, dtype=np.float64, copy=True)
        U = np.zeros_like(out)
        lib().or_format(ctypes.byref(self.cfg), _dptr(U), _dptr(out), 1)
        return U

    def limit_hr(self, U: np.ndarray) -> np.ndarray:
        U = np.array(U, dtype=np.float64, copy=True)
        lib().or_limit_hr(ctypes.byref(self.cfg), ctypes.byref(self.tab), _dptr(U))
        return U

    @staticmethod
    def _thresh(thresholds):
        on = np.array([1 if k < len(thresholds) else 0 for k in range(3)], dtype=np.int32)
        th = np.array(list(thresholds) + [0.0] * (3 - len(thresholds)), dtype=np.float64)
        return on, th

    def sensing(self, U: np.ndarray, thresholds):
        U = np.ascontiguousarray(U, dtype=np.float64)
        on, th = self._thresh(thresholds)
        sens = np.zeros(self.cfg.N_E, dtype=np.int32)
        uavg = np.zeros(self.cfg.N_E * self.nf)
        lib().or_sensing(ctypes.byref(self.cfg), _dptr(self._keep["neighbors"]), _dptr(U), _dptr(on), _dptr(th),
                         _dptr(sens), _dptr(uavg))
        return sens, uavg

    def limit_hri(self, U: np.ndarray, thresholds):
        U = np.array(U, dtype=np.float64, copy=True)
        on, th = self._thresh(thresholds)
        sens = np.zeros(self.cfg.N_E, dtype=np.int32)
        lib().or_limit_hri(ctypes.byref(self.cfg), ctypes.byref(self.tab), _dptr(on), _dptr(th), _dptr(U), _dptr(sens))
        return U, sens

    def limit_flagged(self, U: np.ndarray, flags):
        """HR sweeps of the sensor-gated limiter on the given flags."""
        U = np.array(U, dtype=np.float64, copy=True)
        f = np.ascontiguousarray(flags, dtype=np.int32)
        lib().or_limit_flagged(ctypes.byref(self.cfg), ctypes.byref(self.tab), f.ctypes.data_as(ctypes.c_void_p),
                               _dptr(U))
        return U

    def rk4_out(self, U: np.ndarray, cfl: float, out_times, limiter: int = 0, thresholds=(), nsteps: int = 0,
                max_hist: int = 100000):
        """CFL-mode RK4 with the reference's output times (the last one is the
        final time); returns (U, steps, T, dt history, step count at each output)."""
        U = np.array(U, dtype=np.float64, copy=True)
        on, th = self._thresh(thresholds)
        outs = np.ascontiguousarray(out_times, dtype=np.float64)
        tout = ctypes.c_double()
        hist = np.zeros(max_hist)
        osteps = np.full(outs.size, -1, dtype=np.int32)
        n = lib().or_rk4_out(ctypes.byref(self.cfg), ctypes.byref(self.tab), _dptr(U), 1, cfl, nsteps, _dptr(outs),
                             outs.size, limiter, _dptr(on), _dptr(th), ctypes.byref(tout), _dptr(hist), max_hist,
                             _dptr(osteps))
        return U, n, tout.value, hist[:max(n, 0)], osteps

    def rk4(self, U: np.ndarray, mode: int, dt_or_cfl: float, nsteps: int = 0, tf: float = 0.0, limiter: int = 0,
            thresholds=(), max_hist: int = 100000):
        U = np.array(U, dtype=np.float64, copy=True)
        on, th = self._thresh(thresholds)
        tout = ctypes.c_double()
        hist = np.zeros(max_hist)
        n = lib().or_rk4(ctypes.byref(self.cfg), ctypes.byref(self.tab), _dptr(U), mode, dt_or_cfl, nsteps, tf,
                         limiter, _dptr(on), _dptr(th), ctypes.byref(tout), _dptr(hist), max_hist)
        return U, n, tout.value, hist[:max(n, 0)]
