"""dg_amd — MI355X-native DG right-hand side and RK4 engine (Python bindings).

Thin ctypes layer over the C ABI in include/dgx.h (libdgx.so, built in-tree by
``python -m dg_amd.build``).  Device buffers are passed as raw pointers;
tests and the benchmark use torch tensors for device memory and streams.

The library must be present: nothing here falls back to a CPU path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdgx.so")

SINGLEFLUID, MULTIFLUID, PASSIVE, STIFFENED = 0, 1, 2, 3


def nfields(dim: int, physics: int) -> int:
    """N_F: rho, momenta, E, then alpha (two-fluid), rho*phic and phinc
    (passive) or alpha and beta (stiffened)."""
    return 2 + dim + (0, 1, 2, 2)[physics]
RUSANOV, ROE = 0, 1
PERIODIC, FARFIELD, REFLECTIVE = 1, 2, 3
# limiter codes of the reference (main.cc:167-175) and the learned indicator (no reference code)
LIMIT_NONE, LIMIT_HR, LIMIT_M2, LIMIT_HRI, LIMIT_M2I, LIMIT_PR, LIMIT_PRI, LIMIT_P0I, LIMIT_LEARNED = range(9)
PATH_AUTO, PATH_GENERAL, PATH_TILED = 0, 1, 2


class Config(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("order", ctypes.c_int), ("physics", ctypes.c_int), ("flux", ctypes.c_int)]

    @property
    def nfields(self) -> int:
        return nfields(self.dim, self.physics)


_lib = None
# dgx_output_fn: (user, count, step, time, U, stream) -> status
OUTPUT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                             ctypes.c_void_p, ctypes.c_void_p)


def lib() -> ctypes.CDLL:
    """Load libdgx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        path = os.environ.get("DGX_LIBRARY", LIB_PATH)  # alternative builds (tuning variants)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python -m dg_amd.build`")
        L = ctypes.CDLL(path)
        vp, ip, dp, i = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_double, ctypes.c_int
        L.dgx_last_error.restype = ctypes.c_char_p
        L.dgx_mesh_build.argtypes = [i, i, i, i, dp, dp, dp, dp, ip, i, ctypes.POINTER(vp)]
        L.dgx_mesh_free.argtypes = [vp]
        L.dgx_mesh_tileable.argtypes = [vp]
        L.dgx_mesh_sizes.argtypes = [vp, ip]
        L.dgx_mesh_dx.argtypes = [vp]
        L.dgx_mesh_dx.restype = dp
        L.dgx_mesh_f64.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)]
        L.dgx_mesh_f64.restype = ctypes.POINTER(ctypes.c_double)
        L.dgx_mesh_i32.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)]
        L.dgx_mesh_i32.restype = ctypes.POINTER(ctypes.c_int)
        L.dgx_solver_create.argtypes = [ctypes.POINTER(Config), vp, ctypes.POINTER(vp)]
        L.dgx_solver_free.argtypes = [vp]
        L.dgx_solver_state_size.argtypes = [vp]
        L.dgx_solver_state_size.restype = ctypes.c_long
        L.dgx_rhs.argtypes = [vp, vp, vp, vp]
        L.dgx_set_sensor_thresholds.argtypes = [vp, i, ctypes.POINTER(ctypes.c_double)]
        L.dgx_limit.argtypes = [vp, i, vp, vp]
        L.dgx_mesh_read_gmsh.argtypes = [ctypes.c_char_p, i, vp]
        L.dgx_format.argtypes = [vp, vp, vp, i, vp]
        L.dgx_pos_write.argtypes = [vp, i, vp, ctypes.c_char_p, i, ctypes.c_double]
        L.dgx_pos_read.argtypes = [vp, i, ctypes.c_char_p, i, vp, vp]
        L.dgx_mesh_write_gmsh.argtypes = [vp, ctypes.c_char_p]
        L.dgx_limit_into.argtypes = [vp, i, vp, vp, vp]
        L.dgx_sensing.argtypes = [vp, vp, vp, vp, vp]
        L.dgx_dt_from_cfl.argtypes = [vp, vp, dp, vp, vp]
        L.dgx_rk4_step.argtypes = [vp, vp, vp, i, vp]
        L.dgx_rk4_integrate.argtypes = [vp, vp, i, dp, i, dp, i, ip, ctypes.POINTER(dp), vp]
        L.dgx_rk4_integrate_out.argtypes = [vp, vp, dp, ctypes.POINTER(dp), i, i, OUTPUT_FN, vp, ip,
                                            ctypes.POINTER(dp), vp]
        L.dgx_conservation.argtypes = [vp, vp, vp, vp]
        L.dgx_set_config.argtypes = [ctypes.POINTER(Config)]
        L.dgx_batch_create.argtypes = [ctypes.POINTER(Config), i, ip, ip, dp, dp, ip, ctypes.POINTER(vp)]
        L.dgx_batch_free.argtypes = [vp]
        L.dgx_batch_state_offset.argtypes = [vp, i]
        L.dgx_batch_state_offset.restype = ctypes.c_long
        L.dgx_batch_block_sizes.argtypes = [vp, ip]
        L.dgx_batch_set_block_sizes.argtypes = [vp, ip]
        L.dgx_batch_set_min_chunk.argtypes = [vp, i]
        L.dgx_batch_autotune.argtypes = [vp, vp, dp, i, vp]
        L.dgx_batched_rk4.argtypes = [vp, vp, dp, i, i, i, vp, vp, vp]
        L.dgx_mlp_param_count.restype = i
        L.dgx_mlp_create.argtypes = [ctypes.c_ulonglong, ctypes.POINTER(vp)]
        L.dgx_mlp_free.argtypes = [vp]
        L.dgx_mlp_set_params.argtypes = [vp, vp]
        L.dgx_mlp_get_params.argtypes = [vp, vp]
        L.dgx_mlp_forward.argtypes = [vp, vp, ctypes.c_long, vp, vp]
        L.dgx_mlp_gradient.argtypes = [vp, vp, vp, ctypes.c_long, ctypes.c_float, vp, vp]
        L.dgx_mlp_train.argtypes = [vp, vp, vp, ctypes.c_long, i, i, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_ulonglong, vp, vp]
        L.dgx_solver_set_indicator.argtypes = [vp, vp, ctypes.c_float]
        L.dgx_solver_set_fused_means.argtypes = [vp, i]
        L.dgx_solver_set_device_time.argtypes = [vp, i]
        L.dgx_solver_set_shock_guard.argtypes = [vp, i]
        L.dgx_mlp_features.argtypes = [vp, vp, vp, vp]
        L.dgx_indicator_flags.argtypes = [vp, vp, vp, vp, vp]
        L.dgx_solver_set_path.argtypes = [vp, i]
        L.dgx_solver_set_precision.argtypes = [vp, i]
        L.dgx_solver_path.argtypes = [vp]
        L.dgx_solver_hr_one_launch.argtypes = [vp]
        L.dgx_profile_enable.argtypes = [vp, i]
        L.dgx_profile_read.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(dp), ctypes.POINTER(ctypes.c_long)]
        _lib = L
    return _lib


class DgxError(RuntimeError):
    pass


def _check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().dgx_last_error().decode(errors="replace")
        raise DgxError(f"{what} failed ({status}): {msg}")


SIZE_NAMES = ["D", "order", "N_s", "N_G", "M_s", "M_G", "N_N", "N_E", "M_T", "M_B", "N_F", "N_s1D", "nx", "ny",
              "rflctiveIdx"]
F64_TABLES = ["phi", "dphi", "phi_w", "dphi_w", "psi", "psi_w", "weight", "points", "pointsF", "weightF",
              "ref_nodes", "xyz", "J", "invJac", "JF", "normals", "Minv", "Lag2Mono", "Mono2Lag", "Lag2MonoX",
              "MonoX2MonoY", "MonoY2Lag", "Lag2MonoY", "MonoX2Lag"]
I32_TABLES = ["map", "invmap", "boundaryMap", "neighbors", "face_elem", "face_closure", "face_tag", "closures",
              "elem_id"]


class Mesh:
    """Cartesian mesh + DG tables in the reference layouts (host copies)."""

    def __init__(self, dim: int, order: int, nx: int, ny: int = 1, x0: float = 0.0, x1: float = 1.0,
                 y0: float = 0.0, y1: float = 1.0, bc=(FARFIELD, FARFIELD, FARFIELD, FARFIELD), nfields: int = 4):
        L = lib()
        h = ctypes.c_void_p()
        bcs = (ctypes.c_int * 4)(*(list(bc) + [FARFIELD] * (4 - len(bc))))
        _check(L.dgx_mesh_build(dim, order, nx, ny, x0, x1, y0, y1, bcs, nfields, ctypes.byref(h)), "dgx_mesh_build")
        self._adopt(h)
        self.bc = tuple(bc)
        self.extent = (x0, x1, y0, y1)

    def _adopt(self, h) -> None:
        L = lib()
        self._h = h
        z = (ctypes.c_int * 16)()
        L.dgx_mesh_sizes(h, z)
        self.sizes = {n: z[k] for k, n in enumerate(SIZE_NAMES)}
        self.dx = L.dgx_mesh_dx(h)

    @classmethod
    def from_gmsh(cls, path: str, nfields: int = 4) -> "Mesh":
        """Mesh from a Gmsh 2.2 ASCII file (the reference's mesh input)."""
        h = ctypes.c_void_p()
        _check(lib().dgx_mesh_read_gmsh(str(path).encode(), nfields, ctypes.byref(h)), "dgx_mesh_read_gmsh")
        m = cls.__new__(cls)
        m._adopt(h)
        m.bc = ()
        xyz = m.array("xyz").reshape(-1, m.sizes["D"])
        lo, hi = xyz.min(axis=0), xyz.max(axis=0)
        m.extent = (lo[0], hi[0], lo[1], hi[1]) if m.sizes["D"] == 2 else (lo[0], hi[0], 0.0, 1.0)
        return m

    def write_pos(self, physics: int, out, prefix: str, step: int, time: float) -> None:
        """Printed fields (host, [(e*N_F+f)*N_s+i]) as .pos files "<prefix><fname><step>.pos"."""
        out = np.ascontiguousarray(out, dtype=np.float64)
        _check(lib().dgx_pos_write(self._h, physics, out.ctypes.data, str(prefix).encode(), step, time),
               "dgx_pos_write")

    def read_pos(self, physics: int, prefix: str, step: int):
        """(printed fields, time) read from .pos files."""
        nf = nfields(self.sizes["D"], physics)
        out = np.zeros(self.sizes["N_E"] * nf * self.sizes["N_s"])
        t = ctypes.c_double()
        _check(lib().dgx_pos_read(self._h, physics, str(prefix).encode(), step, ctypes.byref(t), out.ctypes.data),
               "dgx_pos_read")
        return out, t.value

    def write_gmsh(self, path: str) -> None:
        """Write the mesh as Gmsh 2.2 ASCII (boundary faces as tagged elements)."""
        _check(lib().dgx_mesh_write_gmsh(self._h, str(path).encode()), "dgx_mesh_write_gmsh")

    @property
    def tileable(self) -> bool:
        """True when solvers on this mesh run the fused tile kernel by default."""
        return bool(lib().dgx_mesh_tileable(self._h))

    def __getattr__(self, name):
        if name.startswith("_") or name not in F64_TABLES + I32_TABLES:
            raise AttributeError(name)
        return self.array(name)

    def array(self, name: str) -> np.ndarray:
        L = lib()
        n = ctypes.c_long()
        if name in F64_TABLES:
            p = L.dgx_mesh_f64(self._h, name.encode(), ctypes.byref(n))
            if not p or n.value == 0:
                return np.zeros(0)
            return np.ctypeslib.as_array(p, shape=(n.value,)).copy()
        p = L.dgx_mesh_i32(self._h, name.encode(), ctypes.byref(n))
        if not p or n.value == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    def tables(self) -> dict:
        d = {n: self.array(n) for n in F64_TABLES + I32_TABLES}
        d.update(self.sizes)
        d["Dx"] = self.dx
        return d

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().dgx_mesh_free(self._h)
                self._h = None
        except Exception:
            pass


def _ptr(x) -> int:
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _state_ptr(x, n: int, what: str) -> int:
    """Device pointer of a state array of n float64 values; a tensor of
    another size, dtype or layout is refused before anything is launched."""
    if x is None or isinstance(x, int):
        return _ptr(x)
    import torch

    if x.dtype != torch.float64 or x.numel() != n or not x.is_contiguous():
        raise ValueError(f"{what}: expected a contiguous float64 state of {n} values, got {x.dtype} "
                         f"{tuple(x.shape)}")
    return x.data_ptr()


def _stream(stream):
    if stream is None:
        return None
    return getattr(stream, "cuda_stream", stream)


class Solver:
    """Fused engine on one GPU (DG_SOLVER + RK + Limiting + SENSOR)."""

    def __init__(self, config: Config, mesh: Mesh):
        self.config = config
        self.mesh = mesh
        h = ctypes.c_void_p()
        _check(lib().dgx_solver_create(ctypes.byref(config), mesh._h, ctypes.byref(h)), "dgx_solver_create")
        self._h = h
        self.state_size = lib().dgx_solver_state_size(h)

    def _u(self, U) -> int:
        return _state_ptr(U, self.state_size, "Solver")

    def set_sensor_thresholds(self, thresholds) -> None:
        th = (ctypes.c_double * 3)(*(list(thresholds) + [0.0] * (3 - len(thresholds))))
        _check(lib().dgx_set_sensor_thresholds(self._h, len(thresholds), th), "dgx_set_sensor_thresholds")

    def rhs(self, U, f_out, stream=None) -> None:
        _check(lib().dgx_rhs(self._h, self._u(U), self._u(f_out), _stream(stream)), "dgx_rhs")

    def limit(self, limiter: int, U, stream=None) -> None:
        _check(lib().dgx_limit(self._h, limiter, self._u(U), _stream(stream)), "dgx_limit")

    def limit_into(self, limiter: int, In, Out, stream=None) -> None:
        """Out = limited In (In unchanged); 2D hrl on tile meshes runs both sweeps in one launch."""
        _check(lib().dgx_limit_into(self._h, limiter, self._u(In), self._u(Out), _stream(stream)), "dgx_limit_into")

    def format(self, U, out, inverse: bool = False, stream=None) -> None:
        """PRINTER formatter on the device: out = printed fields of U, or U from out when inverse."""
        _check(lib().dgx_format(self._h, self._u(U), _ptr(out), 1 if inverse else 0, _stream(stream)), "dgx_format")

    def checkpoint(self, U, prefix: str, step: int, time: float) -> None:
        """Write U as the reference's .pos files (printed fields, one file per field)."""
        import torch

        out = torch.empty_like(U)
        self.format(U, out)
        torch.cuda.synchronize()
        self.mesh.write_pos(self.config.physics, out.cpu().numpy(), prefix, step, time)

    def restart(self, prefix: str, step: int, U) -> float:
        """Read .pos files into U (device state) through the inverse formatter; returns the time."""
        import torch

        out_h, time = self.mesh.read_pos(self.config.physics, prefix, step)
        out = torch.from_numpy(out_h).to(U.device)
        self.format(U, out, inverse=True)
        torch.cuda.synchronize()
        return time

    def sensing(self, U, sensors, Uavg=None, stream=None) -> None:
        _check(lib().dgx_sensing(self._h, self._u(U), _ptr(sensors), _ptr(Uavg), _stream(stream)), "dgx_sensing")

    def dt_from_cfl(self, U, cfl: float, dt_dev, stream=None) -> None:
        _check(lib().dgx_dt_from_cfl(self._h, self._u(U), cfl, _ptr(dt_dev), _stream(stream)), "dgx_dt_from_cfl")

    def rk4_step(self, U, dt_dev, limiter: int = 0, stream=None) -> None:
        _check(lib().dgx_rk4_step(self._h, self._u(U), _ptr(dt_dev), limiter, _stream(stream)), "dgx_rk4_step")

    def rk4_integrate(self, U, mode: int, dt_or_cfl: float, nsteps: int = 0, tf: float = 0.0, limiter: int = 0,
                      stream=None):
        steps = ctypes.c_int()
        tfin = ctypes.c_double()
        _check(lib().dgx_rk4_integrate(self._h, self._u(U), mode, dt_or_cfl, nsteps, tf, limiter, ctypes.byref(steps),
                                       ctypes.byref(tfin), _stream(stream)), "dgx_rk4_integrate")
        return steps.value, tfin.value

    def rk4_integrate_out(self, U, cfl: float, out_times, limiter: int = 0, on_output=None, stream=None):
        """CFL-mode RK4 with the reference's output times (the last is the final
        time): Dt is clamped to each of them; on_output(count, step, time) is
        called after each output step (the stream synchronised; U holds the
        state).  Returns (steps, final time)."""
        outs = (ctypes.c_double * len(out_times))(*out_times)

        def cb(user, count, step, time, Uptr, st):
            try:
                r = on_output(count, step, time) if on_output else 0
                return int(r or 0)
            except Exception:
                return -1

        fn = OUTPUT_FN(cb)
        steps = ctypes.c_int()
        tfin = ctypes.c_double()
        _check(lib().dgx_rk4_integrate_out(self._h, self._u(U), cfl, outs, len(out_times), limiter, fn, None,
                                           ctypes.byref(steps), ctypes.byref(tfin), _stream(stream)),
               "dgx_rk4_integrate_out")
        return steps.value, tfin.value

    def conservation(self, U, totals, stream=None) -> None:
        """DG_SOLVER::conservation: the integral of each field, into totals (device float64[N_F])."""
        _check(lib().dgx_conservation(self._h, self._u(U), _ptr(totals), _stream(stream)), "dgx_conservation")

    def set_path(self, path: int) -> None:
        """PATH_AUTO, PATH_GENERAL (face + element kernels) or PATH_TILED (one fused kernel per stage)."""
        _check(lib().dgx_solver_set_path(self._h, path), "dgx_solver_set_path")

    def set_precision(self, precision: int) -> None:
        """0: FP64; 1: FP32-MFMA map/integrate contractions in the tile kernel (2D p=2);
        2: the same contractions as split-BF16 products with FP32 accumulation."""
        _check(lib().dgx_solver_set_precision(self._h, precision), "dgx_solver_set_precision")

    @property
    def path(self) -> int:
        """Kernel path in use: PATH_GENERAL or PATH_TILED."""
        return lib().dgx_solver_path(self._h)

    @property
    def hr_one_launch(self) -> bool:
        """True when the 2D global HR limiter runs as one launch."""
        return bool(lib().dgx_solver_hr_one_launch(self._h))

    def set_indicator(self, mlp, threshold: float) -> None:
        """Attach a learned indicator (Indicator) for limiter 8; None detaches."""
        self._mlp = mlp
        _check(lib().dgx_solver_set_indicator(self._h, mlp._h if mlp else None, threshold),
               "dgx_solver_set_indicator")

    def set_fused_means(self, on: bool) -> None:
        """With limiter 8 on the tile kernel, let the RK stages write the cell
        means the indicator reads (default) or leave them to its own pass."""
        _check(lib().dgx_solver_set_fused_means(self._h, 1 if on else 0), "dgx_solver_set_fused_means")

    def set_device_time(self, on: bool) -> None:
        """CFL runs of rk4_integrate with limiter 0 or 1: time control on the
        device, one read-back per chunk of steps (default), or the per-step
        host loop."""
        _check(lib().dgx_solver_set_device_time(self._h, 1 if on else 0), "dgx_solver_set_device_time")

    def set_shock_guard(self, on: bool) -> None:
        """With the sensor thresholds set, let the indicator also flag every
        pair of cells the sensor's shock test marks (default), or leave the
        flags to the network alone."""
        _check(lib().dgx_solver_set_shock_guard(self._h, 1 if on else 0), "dgx_solver_set_shock_guard")

    def features(self, U, X, stream=None) -> None:
        """Indicator features of U into X (float32 device array N_E x 64)."""
        _check(lib().dgx_mlp_features(self._h, self._u(U), _ptr(X), _stream(stream)), "dgx_mlp_features")

    def indicator_flags(self, U, flags, z=None, stream=None) -> None:
        _check(lib().dgx_indicator_flags(self._h, self._u(U), _ptr(flags), _ptr(z), _stream(stream)),
               "dgx_indicator_flags")

    def profile(self, enable: bool = True) -> None:
        _check(lib().dgx_profile_enable(self._h, 1 if enable else 0), "dgx_profile_enable")

    def profile_read(self, kernel: str):
        """(total milliseconds, launches) recorded for kernel since the last read."""
        ms = ctypes.c_double()
        n = ctypes.c_long()
        _check(lib().dgx_profile_read(self._h, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)), "dgx_profile_read")
        return ms.value, n.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().dgx_solver_free(self._h)
                self._h = None
        except Exception:
            pass


class Batch:
    """Ensemble of independent 1D problems advanced together (dgx_batch_*).

    Problem k has n_elem[k] elements of order orders[k] on [x0, x1] with the
    boundary tags bc; its state occupies U[offset(k):offset(k+1)] of the
    packed state, in the usual per-problem layout."""

    def __init__(self, config: Config, n_elem, orders, x0: float = -1.0, x1: float = 1.0,
                 bc=(FARFIELD, FARFIELD)):
        n = len(n_elem)
        ne = (ctypes.c_int * n)(*n_elem)
        po = (ctypes.c_int * n)(*orders)
        bcs = (ctypes.c_int * 2)(*bc)
        h = ctypes.c_void_p()
        _check(lib().dgx_batch_create(ctypes.byref(config), n, ne, po, x0, x1, bcs, ctypes.byref(h)),
               "dgx_batch_create")
        self._h = h
        self.n = n
        self.config = config
        self.offsets = [lib().dgx_batch_state_offset(h, k) for k in range(n + 1)]
        self.state_size = self.offsets[-1]

    def pack(self, states) -> np.ndarray:
        return np.concatenate([np.asarray(s, dtype=np.float64).ravel() for s in states])

    def unpack(self, U):
        U = np.asarray(U)
        return [U[self.offsets[k]:self.offsets[k + 1]] for k in range(self.n)]

    @property
    def block_sizes(self):
        """[workgroup size, workgroups of a launch, problems cut into chunks]."""
        z = (ctypes.c_int * 3)()
        _check(lib().dgx_batch_block_sizes(self._h, z), "dgx_batch_block_sizes")
        return list(z)

    def set_block_sizes(self, sizes) -> None:
        """sizes = [workgroup size (a multiple of 64 in [64, 1024])]."""
        sizes = list(sizes) + [0] * (3 - len(sizes))
        z = (ctypes.c_int * 3)(*sizes)
        _check(lib().dgx_batch_set_block_sizes(self._h, z), "dgx_batch_set_block_sizes")

    def set_min_chunk(self, n: int) -> None:
        """Elements per chunk of a cut problem at least (0: no problem is cut)."""
        _check(lib().dgx_batch_set_min_chunk(self._h, n), "dgx_batch_set_min_chunk")

    def autotune(self, U, cfl: float, limiter: int = LIMIT_HR, stream=None) -> None:
        _check(lib().dgx_batch_autotune(self._h, _ptr(U), cfl, limiter, _stream(stream)), "dgx_batch_autotune")

    def rk4(self, U, cfl: float, nsteps: int, T, status, limiter: int = LIMIT_HR, limit_initial: bool = True,
            stream=None) -> None:
        """nsteps RK4 steps of every problem with Dt = cfl*Dx_k/(2p_k+1)/max(|u|+a); T (float64[n]) and
        status (int32[n]) are device arrays updated per problem."""
        _check(lib().dgx_batched_rk4(self._h, _ptr(U), cfl, nsteps, limiter, 1 if limit_initial else 0, _ptr(T),
                                     _ptr(status), _stream(stream)), "dgx_batched_rk4")

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().dgx_batch_free(self._h)
                self._h = None
        except Exception:
            pass


class Indicator:
    """Learned troubled-cell indicator (dgx_mlp_*): 49 -> 64 -> 64 -> 1 MLP."""

    N_IN, N_PAD, N_HIDDEN = 49, 64, 64

    def __init__(self, seed: int = 0):
        h = ctypes.c_void_p()
        _check(lib().dgx_mlp_create(seed, ctypes.byref(h)), "dgx_mlp_create")
        self._h = h
        self.n_params = lib().dgx_mlp_param_count()

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        _check(lib().dgx_mlp_get_params(self._h, out.ctypes.data), "dgx_mlp_get_params")
        return out

    @params.setter
    def params(self, value) -> None:
        v = np.ascontiguousarray(value, dtype=np.float32)
        assert v.size == self.n_params
        _check(lib().dgx_mlp_set_params(self._h, v.ctypes.data), "dgx_mlp_set_params")

    def forward(self, X, z, n=None, stream=None) -> None:
        n = X.shape[0] if n is None else n
        _check(lib().dgx_mlp_forward(self._h, _ptr(X), n, _ptr(z), _stream(stream)), "dgx_mlp_forward")

    def gradient(self, X, y, grad, pos_weight: float = 1.0, stream=None) -> None:
        """grad: float32 device array of n_params + 1 (the last entry is the loss)."""
        _check(lib().dgx_mlp_gradient(self._h, _ptr(X), _ptr(y), X.shape[0], pos_weight, _ptr(grad),
                                      _stream(stream)), "dgx_mlp_gradient")

    def train(self, X, y, epochs: int, batch: int = 1024, lr: float = 1e-3, pos_weight: float = 1.0,
              seed: int = 0, stream=None) -> np.ndarray:
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        _check(lib().dgx_mlp_train(self._h, _ptr(X), _ptr(y), X.shape[0], epochs, batch, lr, pos_weight, seed,
                                   losses.ctypes.data, _stream(stream)), "dgx_mlp_train")
        return losses[:epochs]

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().dgx_mlp_free(self._h)
                self._h = None
        except Exception:
            pass
