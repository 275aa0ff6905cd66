"""Python front end of libshesolve.so (the MI355X SHE linear solver).

The library is the product; this module only binds its C ABI
(include/shesolve.h) with ctypes for tests, benchmarks and scripting.  It
never falls back to a CPU path: if the library is missing it raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBPATH = os.path.join(PKG, "libshesolve.so")

PRECOND_NONE, PRECOND_ILU0, PRECOND_BLOCK_ILU0, PRECOND_MC_ILU0, PRECOND_ILUK, PRECOND_BMC_ILU0 = 0, 1, 3, 4, 5, 6
PRECOND_BLOCK_ILUK = 7  # ILU(k) of the per-energy slabs (k = fill_level)
FP64, FP32_IR = 0, 1

# symbols declared in include/shesolve.h
EXPORTS = (
    "shesolve_default_params", "shesolve_last_error", "shesolve_create", "shesolve_destroy",
    "shesolve_set_params", "shesolve_set_matrix", "shesolve_set_matrix_device", "shesolve_update_values",
    "shesolve_update_values_device", "shesolve_solve", "shesolve_solve_device", "shesolve_spmv",
    "shesolve_spmv_f32", "shesolve_precond_apply", "shesolve_factor_nnz", "shesolve_get_factors",
    "shesolve_levels", "shesolve_set_profiling", "shesolve_kernel_stats", "shesolve_time_spmv", "shesolve_spmv_bytes",
    "shesolve_get_ordering", "shesolve_stream_config", "shesolve_set_reference_solution", "shesolve_ir_history",
    "shesolve_batch_create", "shesolve_batch_destroy", "shesolve_batch_set_params", "shesolve_batch_set_systems",
    "shesolve_batch_set_systems_device", "shesolve_batch_update_values", "shesolve_batch_update_values_device",
    "shesolve_batch_solve", "shesolve_batch_solve_device", "shesolve_batch_options", "shesolve_batch_info",
    "shesolve_batch_solver", "shesolve_batch_set_fill_levels",
)


class Params(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iters", C.c_int), ("precond", C.c_int), ("precision", C.c_int),
                ("inner_tol", C.c_double), ("max_ir_steps", C.c_int), ("lanes_per_row", C.c_int),
                ("fill_level", C.c_int), ("drop_tol", C.c_double), ("autotune", C.c_int)]


class Report(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("ir_steps", C.c_int), ("rel_residual", C.c_double),
                ("setup_ms", C.c_double), ("solve_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class IRStep(C.Structure):
    """One record of the refinement history (shesolve_ir_step)."""
    _fields_ = [("step", C.c_int32), ("inner_iterations", C.c_int32), ("kind", C.c_int32), ("reserved", C.c_int32),
                ("rel_residual", C.c_double), ("rel_error", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class SheSolveError(RuntimeError):
    pass


_lib = None
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def lib():
    """Loads libshesolve.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise SheSolveError(f"{LIBPATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIBPATH)
    L.shesolve_default_params.argtypes = [C.POINTER(Params)]
    L.shesolve_last_error.restype = C.c_char_p
    L.shesolve_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(Params)]
    L.shesolve_destroy.argtypes = [C.c_void_p]
    L.shesolve_set_params.argtypes = [C.c_void_p, C.POINTER(Params)]
    L.shesolve_set_matrix.argtypes = [C.c_void_p, C.c_int64, C.c_int64, _i64p, _i32p, _f64p, C.c_int, C.c_void_p]
    L.shesolve_set_matrix_device.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_int, C.c_void_p]
    L.shesolve_update_values.argtypes = [C.c_void_p, _f64p]
    L.shesolve_update_values_device.argtypes = [C.c_void_p, C.c_void_p]
    L.shesolve_solve.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int, C.POINTER(Report)]
    L.shesolve_solve_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(Report)]
    L.shesolve_spmv.argtypes = [C.c_void_p, _f64p, _f64p]
    L.shesolve_spmv_f32.argtypes = [C.c_void_p, _f32p, _f32p]
    L.shesolve_precond_apply.argtypes = [C.c_void_p, _f64p, _f64p]
    L.shesolve_factor_nnz.restype = C.c_int64
    L.shesolve_factor_nnz.argtypes = [C.c_void_p]
    L.shesolve_get_factors.argtypes = [C.c_void_p, _i64p, _i32p, _f64p]
    L.shesolve_levels.argtypes = [C.c_void_p, _i64p]
    L.shesolve_get_ordering.argtypes = [C.c_void_p, _i32p]
    L.shesolve_stream_config.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _f64p]
    L.shesolve_set_profiling.argtypes = [C.c_void_p, C.c_int]
    L.shesolve_kernel_stats.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.shesolve_time_spmv.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.shesolve_spmv_bytes.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]
    L.shesolve_set_reference_solution.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int]
    L.shesolve_ir_history.argtypes = [C.c_void_p, C.c_int32, C.POINTER(IRStep), C.POINTER(C.c_int32)]
    VP = C.POINTER(C.c_void_p)
    L.shesolve_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(Params)]
    L.shesolve_batch_destroy.argtypes = [C.c_void_p]
    L.shesolve_batch_set_params.argtypes = [C.c_void_p, C.POINTER(Params)]
    L.shesolve_batch_set_systems.argtypes = [C.c_void_p, C.c_int, _i64p, VP, VP, VP, C.c_void_p, VP]
    L.shesolve_batch_set_systems_device.argtypes = [C.c_void_p, C.c_int, _i64p, VP, VP, VP, C.c_void_p, VP]
    L.shesolve_batch_update_values.argtypes = [C.c_void_p, VP]
    L.shesolve_batch_update_values_device.argtypes = [C.c_void_p, VP]
    L.shesolve_batch_solve.argtypes = [C.c_void_p, VP, VP, C.c_int, C.POINTER(Report)]
    L.shesolve_batch_solve_device.argtypes = [C.c_void_p, VP, VP, C.c_int, C.POINTER(Report)]
    L.shesolve_batch_options.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.shesolve_batch_set_fill_levels.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.shesolve_batch_info.argtypes = [C.c_void_p, _i64p]
    L.shesolve_batch_solver.argtypes = [C.c_void_p]
    L.shesolve_batch_solver.restype = C.c_void_p
    for name in EXPORTS:
        if name not in ("shesolve_last_error", "shesolve_factor_nnz", "shesolve_default_params",
                        "shesolve_batch_solver"):
            getattr(L, name).restype = C.c_int
    L.shesolve_default_params.restype = None
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        raise SheSolveError(f"{what} failed ({rc}): {lib().shesolve_last_error().decode()}")


def default_params(**kw):
    p = Params()
    lib().shesolve_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Solver:
    """One solver handle on one HIP device."""

    def __init__(self, device=0, **params):
        self.params = default_params(**params)
        h = C.c_void_p()
        _check(lib().shesolve_create(C.byref(h), device, C.byref(self.params)), "shesolve_create")
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().shesolve_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        _check(lib().shesolve_set_params(self._h, C.byref(self.params)), "shesolve_set_params")

    def set_matrix(self, row_ptr, col, val, block_ranges=None):
        row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        n = len(row_ptr) - 1
        rr = None
        nb = 0
        if block_ranges is not None:
            rr = np.ascontiguousarray(np.asarray(block_ranges, np.int64).reshape(-1))
            nb = len(rr) // 2
        _check(lib().shesolve_set_matrix(self._h, n, int(row_ptr[-1]), row_ptr, col, val, nb,
                                         rr.ctypes.data if rr is not None else None), "shesolve_set_matrix")
        self.n = n

    def set_matrix_device(self, n, nnz, d_row_ptr, d_col, d_val, block_ranges=None):
        rr = None
        nb = 0
        if block_ranges is not None:
            rr = np.ascontiguousarray(np.asarray(block_ranges, np.int64).reshape(-1))
            nb = len(rr) // 2
        _check(lib().shesolve_set_matrix_device(self._h, n, nnz, d_row_ptr, d_col, d_val, nb,
                                                rr.ctypes.data if rr is not None else None),
               "shesolve_set_matrix_device")
        self.n = n

    def update_values(self, val):
        _check(lib().shesolve_update_values(self._h, np.ascontiguousarray(val, np.float64)), "shesolve_update_values")

    def update_values_device(self, d_val):
        _check(lib().shesolve_update_values_device(self._h, d_val), "shesolve_update_values_device")

    def solve(self, b, x0=None):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64, copy=True)
        rep = Report()
        _check(lib().shesolve_solve(self._h, b, x, 0 if x0 is None else 1, C.byref(rep)), "shesolve_solve")
        return x, rep.as_dict()

    def solve_device(self, d_b, d_x, warm_start=False):
        rep = Report()
        _check(lib().shesolve_solve_device(self._h, d_b, d_x, 1 if warm_start else 0, C.byref(rep)),
               "shesolve_solve_device")
        return rep.as_dict()

    def spmv(self, x):
        y = np.zeros(self.n)
        _check(lib().shesolve_spmv(self._h, np.ascontiguousarray(x, np.float64), y), "shesolve_spmv")
        return y

    def spmv_f32(self, x):
        y = np.zeros(self.n, np.float32)
        _check(lib().shesolve_spmv_f32(self._h, np.ascontiguousarray(x, np.float32), y), "shesolve_spmv_f32")
        return y

    def precond_apply(self, r):
        z = np.zeros(self.n)
        _check(lib().shesolve_precond_apply(self._h, np.ascontiguousarray(r, np.float64), z), "shesolve_precond_apply")
        return z

    def factors(self):
        nnz = int(lib().shesolve_factor_nnz(self._h))
        rp = np.zeros(self.n + 1, np.int64)
        ci = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1))
        _check(lib().shesolve_get_factors(self._h, rp, ci, v), "shesolve_get_factors")
        return rp, ci[:nnz], v[:nnz]

    def stream_config(self):
        it, xc = C.c_int32(), C.c_int32()
        t = np.zeros(9)
        _check(lib().shesolve_stream_config(self._h, C.byref(it), C.byref(xc), t), "shesolve_stream_config")
        return {"items": it.value, "xcd": xc.value, "tune_us": [round(v, 1) for v in t]}

    def set_reference_solution(self, x_ref=None, device_ptr=None):
        """Reference solution for the refinement history: a host array, or the
        address of n FP64 values in device memory; neither clears it."""
        if device_ptr is not None:
            rc = lib().shesolve_set_reference_solution(self._h, self.n, device_ptr, 1)
        elif x_ref is not None:
            self._xref = np.ascontiguousarray(x_ref, np.float64)
            rc = lib().shesolve_set_reference_solution(self._h, len(self._xref), self._xref.ctypes.data, 0)
        else:
            rc = lib().shesolve_set_reference_solution(self._h, 0, None, 0)
        _check(rc, "shesolve_set_reference_solution")

    def ir_history(self):
        """Records of the last solve: initial guess, each refinement step, FP64 hand-over."""
        cnt = C.c_int32()
        _check(lib().shesolve_ir_history(self._h, 0, None, C.byref(cnt)), "shesolve_ir_history")
        out = (IRStep * max(cnt.value, 1))()
        _check(lib().shesolve_ir_history(self._h, cnt.value, out, C.byref(cnt)), "shesolve_ir_history")
        return [out[i].as_dict() for i in range(cnt.value)]

    def ordering(self):
        p = np.zeros(self.n, np.int32)
        _check(lib().shesolve_get_ordering(self._h, p), "shesolve_get_ordering")
        return p

    def levels(self):
        o = np.zeros(2, np.int64)
        _check(lib().shesolve_levels(self._h, o), "shesolve_levels")
        return int(o[0]), int(o[1])

    def set_profiling(self, on=True):
        """on: False/0 off, True/1 every timed launch waited for, 2 SpMV launches timed live."""
        _check(lib().shesolve_set_profiling(self._h, int(on)), "shesolve_set_profiling")

    def kernel_stats(self, kind):
        a = C.c_double()
        n = C.c_int64()
        _check(lib().shesolve_kernel_stats(self._h, kind, C.byref(a), C.byref(n)), "shesolve_kernel_stats")
        return a.value, n.value

    def time_spmv(self, fp32=False, reps=20):
        a = C.c_double()
        _check(lib().shesolve_time_spmv(self._h, 1 if fp32 else 0, reps, C.byref(a)), "shesolve_time_spmv")
        return a.value

    def spmv_bytes(self, fp32=True):
        """Algorithmic bytes of one SpMV in the launched format (16-bit column
        codes in the block multicolour solve space, 4-byte indices otherwise)."""
        b = C.c_int64()
        _check(lib().shesolve_spmv_bytes(self._h, 1 if fp32 else 0, C.byref(b)), "shesolve_spmv_bytes")
        return b.value


def _ptrs(items):
    """ctypes array of void pointers from numpy arrays or integer addresses."""
    arr = (C.c_void_p * len(items))()
    for i, a in enumerate(items):
        arr[i] = a.ctypes.data if isinstance(a, np.ndarray) else a
    return arr


class BatchSolver:
    """Many independent systems solved together on one device
    (shesolve_batch_* in include/shesolve.h)."""

    def __init__(self, device=0, **params):
        self.params = default_params(**params)
        h = C.c_void_p()
        _check(lib().shesolve_batch_create(C.byref(h), device, C.byref(self.params)), "shesolve_batch_create")
        self._h = h
        self.sizes = []

    def close(self):
        if getattr(self, "_h", None):
            lib().shesolve_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        _check(lib().shesolve_batch_set_params(self._h, C.byref(self.params)), "shesolve_batch_set_params")

    def set_fill_levels(self, levels):
        """ILU(k) level per system (0 = ILU(0)) for PRECOND_ILUK; applies at the next set_systems."""
        lv = np.ascontiguousarray(levels, np.int32)
        _check(lib().shesolve_batch_set_fill_levels(self._h, len(lv), lv.ctypes.data if len(lv) else None),
               "shesolve_batch_set_fill_levels")

    def options(self, use_graph=True, poll=8):
        _check(lib().shesolve_batch_options(self._h, 1 if use_graph else 0, poll), "shesolve_batch_options")

    @staticmethod
    def _ranges(block_ranges, count):
        if block_ranges is None:
            return None, None, []
        keep = [np.ascontiguousarray(np.asarray(r, np.int64).reshape(-1)) for r in block_ranges]
        nb = (C.c_int * count)(*[len(r) // 2 for r in keep])
        return nb, _ptrs(keep), keep

    def set_systems(self, systems, block_ranges=None):
        """systems: list of (row_ptr, col, val) host arrays."""
        rps = [np.ascontiguousarray(r, np.int64) for r, _, _ in systems]
        cis = [np.ascontiguousarray(c, np.int32) for _, c, _ in systems]
        vs = [np.ascontiguousarray(v, np.float64) for _, _, v in systems]
        n = np.array([len(r) - 1 for r in rps], np.int64)
        nb, rr, keep = self._ranges(block_ranges, len(systems))
        _check(lib().shesolve_batch_set_systems(self._h, len(systems), n, _ptrs(rps), _ptrs(cis), _ptrs(vs),
                                                C.cast(nb, C.c_void_p) if nb is not None else None, rr),
               "shesolve_batch_set_systems")
        self.sizes = [int(v) for v in n]

    def set_systems_device(self, n, d_row_ptr, d_col, d_val, block_ranges=None):
        """Device CSRs: lists of sizes and device addresses (int)."""
        nn = np.ascontiguousarray(n, np.int64)
        nb, rr, keep = self._ranges(block_ranges, len(nn))
        _check(lib().shesolve_batch_set_systems_device(self._h, len(nn), nn, _ptrs(d_row_ptr), _ptrs(d_col),
                                                       _ptrs(d_val), C.cast(nb, C.c_void_p) if nb is not None else None,
                                                       rr), "shesolve_batch_set_systems_device")
        self.sizes = [int(v) for v in nn]

    def update_values(self, vals):
        vs = [np.ascontiguousarray(v, np.float64) for v in vals]
        _check(lib().shesolve_batch_update_values(self._h, _ptrs(vs)), "shesolve_batch_update_values")

    def update_values_device(self, d_vals):
        _check(lib().shesolve_batch_update_values_device(self._h, _ptrs(d_vals)), "shesolve_batch_update_values_device")

    def solve(self, bs, x0s=None):
        bb = [np.ascontiguousarray(b, np.float64) for b in bs]
        xs = [np.zeros(n) for n in self.sizes] if x0s is None else [np.array(x, np.float64, copy=True) for x in x0s]
        reps = (Report * len(bb))()
        _check(lib().shesolve_batch_solve(self._h, _ptrs(bb), _ptrs(xs), 0 if x0s is None else 1, reps),
               "shesolve_batch_solve")
        return xs, [r.as_dict() for r in reps]

    def solve_device(self, d_bs, d_xs, warm_start=False):
        reps = (Report * len(d_bs))()
        _check(lib().shesolve_batch_solve_device(self._h, _ptrs(d_bs), _ptrs(d_xs), 1 if warm_start else 0, reps),
               "shesolve_batch_solve_device")
        return [r.as_dict() for r in reps]

    def info(self):
        o = np.zeros(6, np.int64)
        _check(lib().shesolve_batch_info(self._h, o), "shesolve_batch_info")
        return dict(zip(("systems", "rows", "nnz", "seg_levels_max", "levels", "factor_nnz"), (int(v) for v in o)))

    def kernel_stats(self, kind):
        """Kernel timing of the underlying concatenated system (see Solver.kernel_stats)."""
        h = lib().shesolve_batch_solver(self._h)
        a = C.c_double()
        n = C.c_int64()
        _check(lib().shesolve_kernel_stats(h, kind, C.byref(a), C.byref(n)), "shesolve_kernel_stats")
        return a.value, n.value

    def set_profiling(self, on=True):
        h = lib().shesolve_batch_solver(self._h)
        _check(lib().shesolve_set_profiling(h, int(on)), "shesolve_set_profiling")

    def precond_apply(self, r):
        """z = M^-1 r with the FP64 factor of the whole block-diagonal system (concatenated vectors)."""
        h = lib().shesolve_batch_solver(self._h)
        z = np.zeros(sum(self.sizes))
        _check(lib().shesolve_precond_apply(h, np.ascontiguousarray(r, np.float64), z), "shesolve_precond_apply")
        return z

    def time_spmv(self, fp32=False, reps=20):
        h = lib().shesolve_batch_solver(self._h)
        a = C.c_double()
        _check(lib().shesolve_time_spmv(h, 1 if fp32 else 0, reps, C.byref(a)), "shesolve_time_spmv")
        return a.value


# ---------------------------------------------------------------- coefficient fill (include/shefill.h)
FILL_EXPORTS = ("shefill_last_error", "shefill_create", "shefill_destroy", "shefill_build", "shefill_sizes",
                "shefill_device_system", "shefill_get_system", "shefill_block_ranges", "shefill_mapping",
                "shefill_last_fill_ms", "shefill_refill", "shefill_recover", "shefill_recover_device",
                "shefill_postproc", "shefill_postproc_device", "shefill_pattern_hash",
                "shefill_density_update", "shefill_density_update_device",
                "shefill_density_update_device_noalloc", "shefill_transfer_device",
                "shefill_terminal_current", "shefill_terminal_current_device")


class FillDevice(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("xv", C.c_void_p), ("yv", C.c_void_p),
                ("doping_n", C.c_void_p), ("doping_p", C.c_void_p), ("potential", C.c_void_p),
                ("contact", C.c_void_p), ("temperature", C.c_double)]


class FillConfig(C.Structure):
    _fields_ = [("carrier", C.c_int), ("L_max", C.c_int), ("energy_spacing", C.c_double),
                ("min_kin_range", C.c_double), ("max_kin_range", C.c_double), ("tau", C.c_double),
                ("acoustic", C.c_int), ("optical", C.c_int), ("scale", C.c_int), ("ordering", C.c_int)]


def _fill_lib():
    L = lib()
    if getattr(L, "_fill_ready", False):
        return L
    L.shefill_last_error.restype = C.c_char_p
    L.shefill_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    L.shefill_destroy.argtypes = [C.c_void_p]
    L.shefill_build.argtypes = [C.c_void_p, C.POINTER(FillDevice), C.POINTER(FillConfig)]
    L.shefill_sizes.argtypes = [C.c_void_p, _i64p]
    L.shefill_refill.argtypes = [C.c_void_p]
    L.shefill_device_system.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
    L.shefill_get_system.argtypes = [C.c_void_p, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p]
    L.shefill_block_ranges.argtypes = [C.c_void_p, _i64p]
    L.shefill_mapping.argtypes = [C.c_void_p, _i64p, _i32p]
    L.shefill_last_fill_ms.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.shefill_recover.argtypes = [C.c_void_p, _f64p, _f64p]
    L.shefill_pattern_hash.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
    L.shefill_terminal_current.argtypes = [C.c_void_p, _f64p, C.c_void_p, C.POINTER(C.c_double)]
    L.shefill_terminal_current_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
    L.shefill_density_update.argtypes = [C.c_void_p, _f64p, _f64p, C.c_double, C.c_int, C.POINTER(C.c_double)]
    L.shefill_density_update_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                                C.POINTER(C.c_double)]
    L.shefill_density_update_device_noalloc.argtypes = L.shefill_density_update_device.argtypes
    L.shefill_transfer_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.shefill_recover_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.shefill_postproc.argtypes = [C.c_void_p, _f64p, C.c_void_p, C.c_void_p]
    L.shefill_postproc_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    for name in FILL_EXPORTS:
        if name != "shefill_last_error":
            getattr(L, name).restype = C.c_int
    L._fill_ready = True
    return L


def _fcheck(rc, what):
    if rc != 0:
        raise SheSolveError(f"{what} failed ({rc}): {_fill_lib().shefill_last_error().decode()}")


class Fill:
    """GPU coefficient fill of one structured device (resident reduced system)."""

    def __init__(self, device=0):
        L = _fill_lib()
        h = C.c_void_p()
        _fcheck(L.shefill_create(C.byref(h), device), "shefill_create")
        self._h = h
        self.sizes = None

    def close(self):
        if getattr(self, "_h", None):
            _fill_lib().shefill_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def build(self, dev, cfg):
        keep = dict(xv=np.ascontiguousarray(dev.xv, np.float64), yv=np.ascontiguousarray(dev.yv, np.float64),
                    dn=np.ascontiguousarray(dev.doping_n, np.float64),
                    dp=np.ascontiguousarray(dev.doping_p, np.float64),
                    psi=np.ascontiguousarray(dev.potential, np.float64),
                    ct=np.ascontiguousarray(dev.contact, np.int32))
        d = FillDevice(dev.dim, dev.nx, dev.ny, keep["xv"].ctypes.data, keep["yv"].ctypes.data,
                       keep["dn"].ctypes.data, keep["dp"].ctypes.data, keep["psi"].ctypes.data,
                       keep["ct"].ctypes.data, dev.temperature)
        c = FillConfig(cfg.carrier, cfg.L_max, cfg.energy_spacing, cfg.min_kin_range, cfg.max_kin_range, cfg.tau,
                       cfg.acoustic, cfg.optical, cfg.scale, int(getattr(cfg, "ordering", 0)))
        _fcheck(_fill_lib().shefill_build(self._h, C.byref(d), C.byref(c)), "shefill_build")
        s = np.zeros(7, np.int64)
        _fcheck(_fill_lib().shefill_sizes(self._h, s), "shefill_sizes")
        self.sizes = dict(NH=int(s[0]), n=int(s[1]), nnz=int(s[2]), n_odd=int(s[3]), nodes=int(s[4]),
                          cells=int(s[5]), facets=int(s[6]))
        self.ncells = dev.ncells
        return self.sizes

    @property
    def n(self):
        return self.sizes["n"]

    @property
    def nnz(self):
        return self.sizes["nnz"]

    def device_system(self):
        ptrs = [C.c_void_p() for _ in range(4)]
        _fcheck(_fill_lib().shefill_device_system(self._h, *[C.byref(p) for p in ptrs]), "shefill_device_system")
        return tuple(p.value for p in ptrs)

    def system(self):
        n, nnz = self.n, self.nnz
        rp = np.zeros(n + 1, np.int64)
        ci = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1))
        rhs, sc, cf = np.zeros(n), np.zeros(n), np.zeros(n)
        _fcheck(_fill_lib().shefill_get_system(self._h, rp, ci, v, rhs, sc, cf), "shefill_get_system")
        return rp, ci[:nnz], v[:nnz], rhs, sc, cf

    def refill(self):
        _fcheck(_fill_lib().shefill_refill(self._h), "shefill_refill")

    def block_ranges(self):
        r = np.zeros(2 * self.sizes["NH"], np.int64)
        _fcheck(_fill_lib().shefill_block_ranges(self._h, r), "shefill_block_ranges")
        return r.reshape(-1, 2)

    def mapping(self):
        m = self.ncells * self.sizes["NH"]
        ei = np.zeros(m, np.int64)
        eo = np.zeros(m, np.int32)
        _fcheck(_fill_lib().shefill_mapping(self._h, ei, eo), "shefill_mapping")
        return ei.reshape(self.ncells, -1), eo.reshape(self.ncells, -1)

    def last_fill_ms(self):
        v = C.c_double()
        _fcheck(_fill_lib().shefill_last_fill_ms(self._h, C.byref(v)), "shefill_last_fill_ms")
        return v.value

    def pattern_hash(self):
        v = C.c_uint64()
        _fcheck(_fill_lib().shefill_pattern_hash(self._h, C.byref(v)), "shefill_pattern_hash")
        return v.value

    def recover(self, x):
        """Full solution (even unknowns scaled back, odd unknowns recovered) on the GPU."""
        xf = np.zeros(self.sizes["n"] + self.sizes["n_odd"])
        _fcheck(_fill_lib().shefill_recover(self._h, np.ascontiguousarray(x, np.float64), xf), "shefill_recover")
        return xf

    def terminal_current(self, facet_current, terminal_mask):
        """Net current of one terminal (cells where terminal_mask is true) from
        the facet currents of postproc, on the GPU."""
        m = np.ascontiguousarray(np.asarray(terminal_mask, bool).astype(np.uint8))
        if m.shape != (self.ncells,):
            raise ValueError(f"terminal mask has {m.size} entries for {self.ncells} cells")
        out = C.c_double()
        _fcheck(_fill_lib().shefill_terminal_current(self._h, np.ascontiguousarray(facet_current, np.float64),
                                                     m.ctypes.data, C.byref(out)), "shefill_terminal_current")
        return out.value

    def density_update_device(self, d_new, d_density, alpha, logarithmic=False):
        """In-place damped update of a device density; returns the relative update norm."""
        nrm = C.c_double()
        _fcheck(_fill_lib().shefill_density_update_device(self._h, d_new, d_density, float(alpha),
                                                          1 if logarithmic else 0, C.byref(nrm)),
                "shefill_density_update_device")
        return nrm.value

    def transfer_device(self, d_x_full_prev, d_density_prev, d_x):
        """Initial guess for the current build from the previous build's full
        solution and density (device addresses), written to d_x (n values):
        the reference's transfer_to_new_h_space on the device."""
        _fcheck(_fill_lib().shefill_transfer_device(self._h, d_x_full_prev, d_density_prev, d_x),
                "shefill_transfer_device")

    def density_update_device_noalloc(self, d_new, d_density, alpha, logarithmic=False):
        """The same update through the allocation-free entry point (its partial
        sums in the library's fixed device scratch; calls on one GPU must not
        overlap)."""
        nrm = C.c_double()
        _fcheck(_fill_lib().shefill_density_update_device_noalloc(self._h, d_new, d_density, float(alpha),
                                                                  1 if logarithmic else 0, C.byref(nrm)),
                "shefill_density_update_device_noalloc")
        return nrm.value

    def density_update(self, new_density, density, alpha=1.0, logarithmic=False):
        """Damped Gummel update of the carrier density on the GPU; returns
        (updated density, relative update norm)."""
        cur = np.array(density, np.float64, copy=True)
        nrm = C.c_double()
        _fcheck(_fill_lib().shefill_density_update(self._h, np.ascontiguousarray(new_density, np.float64), cur,
                                                   float(alpha), 1 if logarithmic else 0, C.byref(nrm)),
                "shefill_density_update")
        return cur, nrm.value

    def recover_device(self, d_x, d_x_full):
        _fcheck(_fill_lib().shefill_recover_device(self._h, d_x, d_x_full), "shefill_recover_device")

    def postproc(self, x_full):
        """Carrier density per cell and current per interior facet on the GPU."""
        dens = np.zeros(self.sizes["cells"])
        cur = np.zeros(max(self.sizes["facets"], 1))
        _fcheck(_fill_lib().shefill_postproc(self._h, np.ascontiguousarray(x_full, np.float64), dens.ctypes.data,
                                             cur.ctypes.data), "shefill_postproc")
        return dens, cur[:self.sizes["facets"]]

    def postproc_device(self, d_x_full, d_density, d_current):
        _fcheck(_fill_lib().shefill_postproc_device(self._h, d_x_full, d_density, d_current),
                "shefill_postproc_device")
