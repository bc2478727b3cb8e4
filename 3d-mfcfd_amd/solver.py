"""Python host mirror of the reference's hot-path interface over the C ABI.

``Solver`` owns one libmfcfd_hip.so context and exposes the reference's
entry points under their names (fpi_solver.f90, compute_conserved_vector.f90,
q_variables.f90, q_derivatives.f90, timestep_delt.f90, flux_residual.f90,
implicit_forward_and_reverse_sweeps.f90, state_update.f90).  Arrays cross
the boundary in the reference's layouts (see include/mfcfd_hip.h); the
device copies stay in HBM between calls.

There is no CPU fallback: constructing a Solver without the built library
or without a GPU raises.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfcfd_hip.so")

EXPLICIT, IMPLICIT, IMPLICIT_BACKUP = 0, 1, 2
# iteration forms by name: the explicit form, the canonical implicit pairing
# (implicit_forward_and_reverse_sweeps.f90 + state_update.f90) and the
# alternative one (implicit_forward_and_reverse_sweeps_backup.f90 +
# state_update_implicit.f90)
MODES = {"explicit": EXPLICIT, "implicit": IMPLICIT, "implicit_backup": IMPLICIT_BACKUP}
FP64, MIXED = 0, 1
NOALLOC_MAX_ITERS = 4096   # iterations of one mfcfd_run_noalloc call
ARR = {"prim": 0, "U": 1, "q": 2, "qm": 3, "dq": 4, "delt": 5, "flux_res": 6, "delUp": 7,
       "delUn": 8}
ARR_SHAPE = {"prim": (5,), "U": (5,), "q": (5,), "qm": (5, 2), "dq": (5, 3), "delt": (),
             "flux_res": (5,), "delUp": (5,), "delUn": (5,)}
SPLIT_NAMES = ["xpos", "xneg", "ypos", "yneg", "zpos", "zneg"]
SPLIT_SLACK = 32

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class MfcfdError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [("mach", C.c_double), ("aoa", C.c_double), ("cfl", C.c_double),
                ("vl_const", C.c_double), ("power", C.c_double),
                ("inner_iterations", C.c_int), ("mode", C.c_int), ("precision", C.c_int),
                ("max_iters", C.c_int), ("flux_order", C.c_int), ("sweep_schedule", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("flux_interior_block", C.c_int), ("flux_interior_waves", C.c_int),
                ("flux_wall_block", C.c_int), ("flux_wall_waves", C.c_int),
                ("qderiv_block", C.c_int), ("qderiv_waves", C.c_int), ("sweep_div", C.c_int),
                ("flux_interior_lds", C.c_int), ("flux_wall_lds", C.c_int), ("qderiv_lds", C.c_int),
                ("flux_interior_ms", C.c_double), ("flux_wall_ms", C.c_double),
                ("qderiv_ms", C.c_double), ("sweeps_ms", C.c_double),
                ("flux_interior_default_ms", C.c_double), ("flux_wall_default_ms", C.c_double),
                ("qderiv_default_ms", C.c_double), ("sweeps_default_ms", C.c_double)]

    def as_dict(self):
        return {nm: getattr(self, nm) for nm, _ in self._fields_}


class Geometry(C.Structure):
    _fields_ = [("n", C.c_int), ("x", _dp), ("y", _dp), ("z", _dp), ("tan1", _dp),
                ("tan2", _dp), ("nor", _dp), ("min_dist", _dp), ("status", _ip),
                ("nbhs", _ip), ("conn", _ip), ("split_nbhs", _ip * 6), ("split_conn", _ip * 6),
                ("alias", _ip), ("point_with_alias", _ip)]


class Residual(C.Structure):
    _fields_ = [("res_old", C.c_double), ("res_new", C.c_double), ("residue", C.c_double),
                ("max_res", C.c_double), ("sum_res_sqr", C.c_double),
                ("max_res_point", C.c_int), ("iteration", C.c_int)]


_lib = None

EXPORTS = ["mfcfd_create", "mfcfd_destroy", "mfcfd_last_error", "mfcfd_abi_version",
           "mfcfd_set_params", "mfcfd_bind_geometry", "mfcfd_bind_geometry_arrays",
           "mfcfd_generate_split_stencils",
           "mfcfd_aliasing", "mfcfd_upload", "mfcfd_download",
           "mfcfd_compute_conserved_vector", "mfcfd_eval_q_variables",
           "mfcfd_eval_q_derivatives", "mfcfd_timestep_delt", "mfcfd_eval_flux_residual",
           "mfcfd_implicit_forward_backward_sweeps", "mfcfd_state_update", "mfcfd_fpi_solver",
           "mfcfd_run", "mfcfd_get_residual", "mfcfd_synchronize", "mfcfd_sweep_info",
           "mfcfd_time_phases", "mfcfd_point_classes", "mfcfd_point_class_map",
           "mfcfd_kernel_timing", "mfcfd_kernel_time", "mfcfd_bind_batch", "mfcfd_batch_info",
           "mfcfd_run_graph", "mfcfd_autotune", "mfcfd_set_tuning", "mfcfd_run_noalloc"]
PHASES = ["conserved_q", "qderiv_timestep", "flux_interior", "flux_boundary", "sweeps",
          "state_update"]


def lib():
    """Load libmfcfd_hip.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MfcfdError(f"{LIB_PATH} is not built; run 3d-mfcfd_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.mfcfd_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.mfcfd_destroy.argtypes = [vp]
        L.mfcfd_last_error.restype = C.c_char_p
        L.mfcfd_set_params.argtypes = [vp, C.POINTER(Params)]
        L.mfcfd_bind_geometry.argtypes = [vp, C.POINTER(Geometry)]
        L.mfcfd_generate_split_stencils.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                                                    _ip, _ip, _ip * 6, _ip * 6]
        L.mfcfd_aliasing.argtypes = [C.c_int, _ip, _ip, _ip, _ip]
        L.mfcfd_upload.argtypes = [vp, C.c_int, _dp]
        L.mfcfd_download.argtypes = [vp, C.c_int, _dp]
        for nm in ["mfcfd_compute_conserved_vector", "mfcfd_eval_q_variables",
                   "mfcfd_eval_q_derivatives", "mfcfd_timestep_delt", "mfcfd_eval_flux_residual",
                   "mfcfd_implicit_forward_backward_sweeps", "mfcfd_state_update",
                   "mfcfd_synchronize"]:
            getattr(L, nm).argtypes = [vp]
        L.mfcfd_fpi_solver.argtypes = [vp, C.c_int, _dp, _dp]
        L.mfcfd_run.argtypes = [vp, C.c_int, C.c_int, _dp, _dp]
        L.mfcfd_run_noalloc.argtypes = [vp, C.c_int, C.c_int, _dp, _dp]
        L.mfcfd_get_residual.argtypes = [vp, C.POINTER(Residual)]
        L.mfcfd_sweep_info.argtypes = [vp, _ip, _ip, _ip]
        L.mfcfd_time_phases.argtypes = [vp, C.c_int, C.c_int, _dp]
        L.mfcfd_point_classes.argtypes = [vp, _ip, _ip, _ip]
        L.mfcfd_point_class_map.argtypes = [vp, _ip]
        L.mfcfd_kernel_timing.argtypes = [vp, C.c_int]
        L.mfcfd_kernel_time.argtypes = [vp, _dp, C.POINTER(C.c_longlong)]
        L.mfcfd_bind_batch.argtypes = [vp, C.c_int, C.POINTER(Geometry), _dp, _dp]
        L.mfcfd_batch_info.argtypes = [vp, _ip, _ip]
        L.mfcfd_run_graph.argtypes = [vp, C.c_int, C.c_int, _dp, _dp]
        L.mfcfd_autotune.argtypes = [vp, C.c_int, C.POINTER(Tuning)]
        L.mfcfd_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MfcfdError(f"mfcfd error {rc}: {lib().mfcfd_last_error().decode()}")


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def split_stencils(cloud):
    """generate_split_stencils through the library's host implementation.

    Returns (counts (6,N) int32, lists (6, 22N+32) int32) in xpos..zneg
    order, the lists in the reference's raw (22,N) addressing.
    """
    n = cloud.n
    cnt = np.zeros((6, n), dtype=np.int32)
    lists = np.zeros((6, 22 * n + SPLIT_SLACK), dtype=np.int32)
    arrs = _cloud_arrays(cloud)
    cn = (_ip * 6)(*[_i(cnt[d]) for d in range(6)])
    ls = (_ip * 6)(*[_i(lists[d]) for d in range(6)])
    _check(lib().mfcfd_generate_split_stencils(
        n, _d(arrs["x"]), _d(arrs["y"]), _d(arrs["z"]), _d(arrs["tan1"]), _d(arrs["tan2"]),
        _d(arrs["nor"]), _i(arrs["status"]), _i(arrs["nbhs"]), _i(arrs["conn"]), cn, ls))
    return cnt, lists


def alias_order(cloud):
    """aliasing() through the library's host implementation."""
    n = cloud.n
    alias = np.zeros(n, dtype=np.int32)
    pwa = np.zeros(n, dtype=np.int32)
    arrs = _cloud_arrays(cloud)
    _check(lib().mfcfd_aliasing(n, _i(arrs["nbhs"]), _i(arrs["conn"]), _i(alias), _i(pwa)))
    return alias, pwa


def _cloud_arrays(cloud):
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    i = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    return {"x": f(cloud.x), "y": f(cloud.y), "z": f(cloud.z), "tan1": f(cloud.tan1),
            "tan2": f(cloud.tan2), "nor": f(cloud.nor), "min_dist": f(cloud.min_dist),
            "status": i(cloud.status), "nbhs": i(cloud.nbhs), "conn": i(cloud.conn)}


def _geometry(cloud, split=None, alias=None):
    """mfcfd_geometry of one cloud and the arrays it points into."""
    a = _cloud_arrays(cloud)
    cnt, lists = split if split is not None else split_stencils(cloud)
    al, pwa = alias if alias is not None else alias_order(cloud)
    keep = [a, np.ascontiguousarray(cnt, dtype=np.int32), np.ascontiguousarray(lists, dtype=np.int32),
            np.ascontiguousarray(al, dtype=np.int32), np.ascontiguousarray(pwa, dtype=np.int32)]
    _, cnt, lists, al, pwa = keep
    g = Geometry()
    g.n = cloud.n
    for nm in ("x", "y", "z", "tan1", "tan2", "nor", "min_dist"):
        setattr(g, nm, _d(a[nm]))
    for nm in ("status", "nbhs", "conn"):
        setattr(g, nm, _i(a[nm]))
    g.split_nbhs = (_ip * 6)(*[_i(cnt[d]) for d in range(6)])
    g.split_conn = (_ip * 6)(*[_i(lists[d]) for d in range(6)])
    g.alias = _i(al)
    g.point_with_alias = _i(pwa)
    return g, keep


class Solver:
    """One device context for one point cloud.

    Mirrors the reference's hot-path module procedures; array names and
    shapes follow data_structure.f90 (row-per-point views of the (m,N)
    column-major arrays).
    """

    def __init__(self, cloud, mach, aoa, cfl, mode="explicit", precision="fp64",
                 vl_const=2.0, power=0.0, inner_iterations=0, max_iters=1000, device=0,
                 split=None, alias=None, flux_order="fused", sweep_schedule="async"):
        self._create(cloud.n, mach, aoa, cfl, mode, precision, vl_const, power, inner_iterations,
                     max_iters, device, flux_order, sweep_schedule)
        self.cloud = cloud
        g, self._keep = _geometry(cloud, split, alias)
        _check(lib().mfcfd_bind_geometry(self._ctx, C.byref(g)))

    def _create(self, n, mach, aoa, cfl, mode, precision, vl_const, power, inner_iterations,
                max_iters, device, flux_order, sweep_schedule):
        L = lib()
        self.n = n
        self._ctx = C.c_void_p()
        _check(L.mfcfd_create(device, C.byref(self._ctx)))
        self.params = Params(mach, aoa, cfl, vl_const, power, inner_iterations,
                             MODES[mode],
                             MIXED if precision == "mixed" else FP64, max_iters,
                             1 if flux_order == "lists" else 0,
                             1 if sweep_schedule == "levels" else 0)
        _check(L.mfcfd_set_params(self._ctx, C.byref(self.params)))

    def close(self):
        if self._ctx:
            lib().mfcfd_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state transfer --------------------------------------------------------
    def upload(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape((self.n,) + ARR_SHAPE[name])
        _check(lib().mfcfd_upload(self._ctx, ARR[name], _d(a)))

    def download(self, name):
        a = np.empty((self.n,) + ARR_SHAPE[name], dtype=np.float64)
        _check(lib().mfcfd_download(self._ctx, ARR[name], _d(a)))
        return a

    # entry points ----------------------------------------------------------
    def compute_conserved_vector(self):
        _check(lib().mfcfd_compute_conserved_vector(self._ctx))

    def eval_q_variables(self):
        _check(lib().mfcfd_eval_q_variables(self._ctx))

    def eval_q_derivatives(self):
        _check(lib().mfcfd_eval_q_derivatives(self._ctx))

    def timestep_delt(self):
        _check(lib().mfcfd_timestep_delt(self._ctx))

    def eval_flux_residual(self):
        _check(lib().mfcfd_eval_flux_residual(self._ctx))

    def implicit_forward_backward_sweeps(self):
        _check(lib().mfcfd_implicit_forward_backward_sweeps(self._ctx))

    def state_update(self):
        _check(lib().mfcfd_state_update(self._ctx))

    def fpi_solver(self, t):
        rn = np.zeros(1)
        rs = np.zeros(1)
        _check(lib().mfcfd_fpi_solver(self._ctx, t, _d(rn), _d(rs)))
        return float(rn[0]), float(rs[0])

    def run(self, t0, n_iters):
        rn = np.zeros(max(n_iters, 1))
        rs = np.zeros(max(n_iters, 1))
        _check(lib().mfcfd_run(self._ctx, t0, n_iters, _d(rn), _d(rs)))
        return rn[:n_iters], rs[:n_iters]

    def run_noalloc(self, t0, n_iters):
        """mfcfd_run without allocation (no counterpart in the reference):
        at most NOALLOC_MAX_ITERS iterations, and the implicit form's sweep
        plan must already be built (sweep_info, fpi_solver or run)."""
        rn = np.zeros(max(n_iters, 1))
        rs = np.zeros(max(n_iters, 1))
        _check(lib().mfcfd_run_noalloc(self._ctx, t0, n_iters, _d(rn), _d(rs)))
        return rn[:n_iters], rs[:n_iters]

    def residual(self):
        r = Residual()
        _check(lib().mfcfd_get_residual(self._ctx, C.byref(r)))
        return {f: getattr(r, f) for f, _ in Residual._fields_}

    def synchronize(self):
        _check(lib().mfcfd_synchronize(self._ctx))

    def time_phases(self, t0, n_iters):
        """Mean milliseconds per iteration of each phase (HIP events on the
        context's stream); advances the state by n_iters iterations."""
        out = np.zeros(len(PHASES))
        _check(lib().mfcfd_time_phases(self._ctx, t0, n_iters, _d(out)))
        return dict(zip(PHASES, out.tolist()))

    def point_classes(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().mfcfd_point_classes(self._ctx, C.byref(a), C.byref(b), C.byref(c)))
        return {"fused_interior": a.value, "fused_general": b.value, "listwise": c.value}

    def point_class_map(self):
        """Per host point: 0 fused interior, 1 fused general (wall, outer, spill-altered
        interior), 2 list-by-list."""
        out = np.zeros(self.n, dtype=np.int32)
        _check(lib().mfcfd_point_class_map(self._ctx, out.ctypes.data_as(_ip)))
        return out

    def kernel_timing(self, enable=True, kernel="flux"):
        """Record HIP events around every launch of one kernel: "flux" (the
        fused interior flux kernel) or "sweeps" (each sync-free sweep kernel
        launch, forward and backward)."""
        sel = {"flux": 1, "sweeps": 2}[kernel] if enable else 0
        _check(lib().mfcfd_kernel_timing(self._ctx, sel))

    def kernel_time(self):
        """(summed ms, launches) of the fused interior flux kernel since
        kernel_timing() was called."""
        ms, k = C.c_double(), C.c_longlong()
        _check(lib().mfcfd_kernel_time(self._ctx, C.byref(ms), C.byref(k)))
        return ms.value, k.value

    def run_graph(self, t0, n_iters):
        """run() with the iteration captured as a HIP graph; returns the
        res_new history and the per-cloud residual history (n_iters, clouds)."""
        rn = np.zeros(n_iters)
        ni = self.batch_info()[0]
        inst = np.zeros((n_iters, ni))
        _check(lib().mfcfd_run_graph(self._ctx, t0, n_iters, _d(rn), _d(inst)))
        return rn, inst

    def autotune(self, repeats=3):
        """Time the compiled launch-shape variants on this cloud and keep the
        fastest (mfcfd_autotune); returns the choices and times."""
        t = Tuning()
        _check(lib().mfcfd_autotune(self._ctx, repeats, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, tuning):
        t = Tuning(**{k: v for k, v in tuning.items() if k in dict(Tuning._fields_)})
        _check(lib().mfcfd_set_tuning(self._ctx, C.byref(t)))

    def batch_info(self):
        k = C.c_int()
        _check(lib().mfcfd_batch_info(self._ctx, C.byref(k), None))
        counts = np.zeros(k.value, dtype=np.int32)
        _check(lib().mfcfd_batch_info(self._ctx, C.byref(k), counts.ctypes.data_as(_ip)))
        return k.value, counts

    def sweep_info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().mfcfd_sweep_info(self._ctx, C.byref(a), C.byref(b), C.byref(c)))
        return {"levels_fwd": a.value, "levels_bwd": b.value, "swept": c.value}


class BatchSolver(Solver):
    """Several independent clouds in one context (mfcfd_bind_batch): every
    launch covers all of them; each keeps its own Mach and angle of attack.
    Arrays of upload/download hold the clouds one after the other."""

    def __init__(self, clouds, machs, aoas, cfl, mode="explicit", precision="fp64", vl_const=2.0,
                 power=0.0, max_iters=1000, device=0, flux_order="fused", sweep_schedule="async"):
        n = int(sum(c.n for c in clouds))
        self._create(n, float(machs[0]), float(aoas[0]), cfl, mode, precision, vl_const, power, 0,
                     max_iters, device, flux_order, sweep_schedule)
        self.clouds = list(clouds)
        self.offsets = np.concatenate([[0], np.cumsum([c.n for c in clouds])]).astype(np.int64)
        geoms = (Geometry * len(clouds))()
        self._keep = []
        for b, cl in enumerate(clouds):
            g, keep = _geometry(cl)
            geoms[b] = g
            self._keep.append(keep)
        self._geoms = geoms
        m = np.ascontiguousarray(machs, dtype=np.float64)
        a = np.ascontiguousarray(aoas, dtype=np.float64)
        _check(lib().mfcfd_bind_batch(self._ctx, len(clouds), geoms, _d(m), _d(a)))

    def split(self, arr):
        """Per-cloud views of a concatenated (N, ...) array."""
        return [arr[self.offsets[b]:self.offsets[b + 1]] for b in range(len(self.clouds))]

