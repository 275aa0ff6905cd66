"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the MI355X path.  Only the test
suite, the smoke check and the CPU-baseline leg of the benchmark import this
module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "liboracle.so")
TABLES = os.path.join(ROOT, "tests", "golden", "coupling_tables.json")

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class _Coupling(C.Structure):
    _fields_ = [("nnz", C.c_int * 6), ("row", C.c_void_p * 6), ("col", C.c_void_p * 6),
                ("val", C.c_void_p * 6)]


class _Device(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("xv", C.c_void_p),
                ("yv", C.c_void_p), ("doping_n", C.c_void_p), ("doping_p", C.c_void_p),
                ("potential", C.c_void_p), ("contact", C.c_void_p), ("temperature", C.c_double)]


class _Config(C.Structure):
    _fields_ = [("carrier", C.c_int), ("L_max", C.c_int), ("energy_spacing", C.c_double),
                ("min_kin_range", C.c_double), ("max_kin_range", C.c_double), ("tau", C.c_double),
                ("acoustic", C.c_int), ("optical", C.c_int), ("scale", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    L.oracle_build.restype = C.c_void_p
    L.oracle_build.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.oracle_free.argtypes = [C.c_void_p]
    L.oracle_last_error.restype = C.c_char_p
    L.oracle_sizes.argtypes = [C.c_void_p, _i64p]
    L.oracle_mapping.argtypes = [C.c_void_p, _i64p, _i32p, _i64p, _i32p]
    L.oracle_facets.argtypes = [C.c_void_p, _i32p, _i32p, _i32p]
    L.oracle_energies.argtypes = [C.c_void_p, _f64p, _f64p]
    L.oracle_reduced.argtypes = [C.c_void_p, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p]
    L.oracle_unscaled_nnz.restype = C.c_int64
    L.oracle_unscaled_nnz.argtypes = [C.c_void_p]
    L.oracle_unscaled.argtypes = [C.c_void_p, _i64p, _i32p, _f64p, _f64p]
    L.oracle_full.argtypes = [C.c_void_p, _i64p, _i32p, _f64p, _f64p]
    L.oracle_block_ranges.argtypes = [C.c_void_p, _i64p]
    L.oracle_finish.restype = C.c_double
    L.oracle_finish.argtypes = [C.c_void_p, _f64p, _f64p]
    L.oracle_postproc.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
    L.oracle_precond_create.restype = C.c_void_p
    L.oracle_precond_create.argtypes = [C.c_int, C.c_int64, _i64p, _i32p, _f64p, C.c_double, C.c_int,
                                        C.c_int, C.c_void_p]
    L.oracle_precond_apply.argtypes = [C.c_void_p, _f64p, _f64p]
    L.oracle_precond_nnz.restype = C.c_int64
    L.oracle_precond_nnz.argtypes = [C.c_void_p]
    L.oracle_precond_factors.argtypes = [C.c_void_p, _i64p, _i32p, _f64p]
    L.oracle_precond_free.argtypes = [C.c_void_p]
    L.oracle_spmv.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p]
    L.oracle_dispersion.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.oracle_bicgstab.restype = C.c_int
    L.oracle_bicgstab.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p, C.c_void_p, C.c_double,
                                  C.c_int, C.c_double, C.POINTER(C.c_double)]
    L.oracle_bicgstab_parallel.restype = C.c_int
    L.oracle_bicgstab_parallel.argtypes = L.oracle_bicgstab.argtypes
    L.oracle_threads.restype = C.c_int
    L.oracle_precond_create_parallel.restype = C.c_void_p
    L.oracle_precond_create_parallel.argtypes = L.oracle_precond_create.argtypes
    L.oracle_precond_blocks.restype = C.c_int
    L.oracle_precond_blocks.argtypes = [C.c_void_p]
    L.oracle_set_plant.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
    _lib = L
    return L


_tables = None


def coupling_tables():
    global _tables
    if _tables is None:
        with open(TABLES) as f:
            _tables = json.load(f)
    return _tables


NAMES = ("a_x", "a_y", "a_z", "b_x", "b_y", "b_z")


class planted:
    """Context manager of the fill's test-only scale factors (oracle_set_plant):
    the diffusion and drift coefficients and the acoustic and optical rates of
    every system built inside the block are multiplied by them; all go back to
    1 on exit.  Used by the physics pins to switch a term off and to plant a
    mistake their checks must catch."""

    def __init__(self, diffusion=1.0, drift=1.0, acoustic=1.0, optical=1.0):
        self.f = (float(diffusion), float(drift), float(acoustic), float(optical))

    def __enter__(self):
        lib().oracle_set_plant(*self.f)
        return self

    def __exit__(self, *exc):
        lib().oracle_set_plant(1.0, 1.0, 1.0, 1.0)
        return False


def dispersion(carrier, eps):
    z, v = C.c_double(), C.c_double()
    lib().oracle_dispersion(carrier, eps, C.byref(z), C.byref(v))
    return z.value, v.value


class CSR:
    def __init__(self, row_ptr, col, val):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)

    @property
    def n(self):
        return len(self.row_ptr) - 1

    @property
    def nnz(self):
        return int(self.row_ptr[-1])

    def matvec(self, x):
        y = np.empty(self.n)
        lib().oracle_spmv(self.n, self.row_ptr, self.col, self.val, np.ascontiguousarray(x, np.float64), y)
        return y

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.row_ptr), shape=(self.n, self.n))


class OracleSystem:
    """The assembled SHE system of one device, restated on the CPU."""

    def __init__(self, dev, cfg):
        L = lib()
        tab = coupling_tables()[str(cfg.L_max)]
        keep = []
        cp = _Coupling()
        for t, name in enumerate(NAMES):
            arr = np.array(tab[name], dtype=np.float64).reshape(-1, 3)
            r = np.ascontiguousarray(arr[:, 0].astype(np.int32))
            c = np.ascontiguousarray(arr[:, 1].astype(np.int32))
            v = np.ascontiguousarray(arr[:, 2])
            keep += [r, c, v]
            cp.nnz[t] = len(v)
            cp.row[t] = r.ctypes.data
            cp.col[t] = c.ctypes.data
            cp.val[t] = v.ctypes.data
        arrs = dict(xv=np.ascontiguousarray(dev.xv, np.float64), yv=np.ascontiguousarray(dev.yv, np.float64),
                    dn=np.ascontiguousarray(dev.doping_n, np.float64),
                    dp=np.ascontiguousarray(dev.doping_p, np.float64),
                    psi=np.ascontiguousarray(dev.potential, np.float64),
                    ct=np.ascontiguousarray(dev.contact, np.int32))
        d = _Device(dev.dim, dev.nx, dev.ny, arrs["xv"].ctypes.data, arrs["yv"].ctypes.data,
                    arrs["dn"].ctypes.data, arrs["dp"].ctypes.data, arrs["psi"].ctypes.data,
                    arrs["ct"].ctypes.data, dev.temperature)
        c = _Config(cfg.carrier, cfg.L_max, cfg.energy_spacing, cfg.min_kin_range, cfg.max_kin_range, cfg.tau,
                    cfg.acoustic, cfg.optical, cfg.scale)
        h = L.oracle_build(C.byref(d), C.byref(c), C.byref(cp))
        if not h:
            raise RuntimeError("oracle_build failed: " + L.oracle_last_error().decode())
        self._h = h
        self.dev, self.cfg = dev, cfg
        s = np.zeros(8, np.int64)
        L.oracle_sizes(h, s)
        (self.NH, self.n_even, self.n_odd, self.nnz_full, self.nnz, self.ncells, self.nfacets,
         self.explicit_zeros) = (int(v) for v in s)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_free(self._h)
            self._h = None

    def mapping(self):
        ei = np.zeros(self.ncells * self.NH, np.int64)
        eo = np.zeros(self.ncells * self.NH, np.int32)
        oi = np.zeros(max(self.nfacets, 1) * self.NH, np.int64)
        oo = np.zeros(max(self.nfacets, 1) * self.NH, np.int32)
        lib().oracle_mapping(self._h, ei, eo, oi, oo)
        n = self.nfacets * self.NH
        return (ei.reshape(self.ncells, self.NH), eo.reshape(self.ncells, self.NH),
                oi[:n].reshape(self.nfacets, self.NH), oo[:n].reshape(self.nfacets, self.NH))

    def facets(self):
        a, b, d = (np.zeros(self.nfacets, np.int32) for _ in range(3))
        lib().oracle_facets(self._h, a, b, d)
        return a, b, d

    def energies(self):
        H = np.zeros(self.NH)
        sh = np.zeros(self.ncells)
        lib().oracle_energies(self._h, H, sh)
        return H, sh

    def reduced(self):
        """(CSR, rhs, unknown scaling, column factors) handed to solvers::solve."""
        n = self.n_even
        rp = np.zeros(n + 1, np.int64)
        ci = np.zeros(self.nnz, np.int32)
        v = np.zeros(self.nnz)
        rhs, sc, cf = np.zeros(n), np.zeros(n), np.zeros(n)
        lib().oracle_reduced(self._h, rp, ci, v, rhs, sc, cf)
        return CSR(rp, ci, v), rhs, sc, cf

    def unscaled(self):
        n = self.n_even
        nnz = int(lib().oracle_unscaled_nnz(self._h))
        rp = np.zeros(n + 1, np.int64)
        ci = np.zeros(nnz, np.int32)
        v = np.zeros(nnz)
        rhs = np.zeros(n)
        lib().oracle_unscaled(self._h, rp, ci, v, rhs)
        return CSR(rp, ci, v), rhs

    def full(self):
        N = self.n_even + self.n_odd
        rp = np.zeros(N + 1, np.int64)
        ci = np.zeros(self.nnz_full, np.int32)
        v = np.zeros(self.nnz_full)
        rhs = np.zeros(N)
        lib().oracle_full(self._h, rp, ci, v, rhs)
        return CSR(rp, ci, v), rhs

    def block_ranges(self):
        r = np.zeros(2 * self.NH, np.int64)
        lib().oracle_block_ranges(self._h, r)
        return r.reshape(self.NH, 2)

    def finish(self, x_reduced):
        xf = np.zeros(self.n_even + self.n_odd)
        res = lib().oracle_finish(self._h, np.ascontiguousarray(x_reduced, np.float64), xf)
        return xf, float(res)

    def postproc(self, x_full):
        dens = np.zeros(self.ncells)
        cur = np.zeros(max(self.nfacets, 1))
        lib().oracle_postproc(self._h, np.ascontiguousarray(x_full, np.float64), dens, cur)
        return dens, cur[: self.nfacets]


PC_NONE, PC_ILU0, PC_ILUT, PC_BLOCK_ILU0, PC_ILUK = 0, 1, 2, 3, 5


class Precond:
    def __init__(self, kind, A: CSR, drop=1e-3, fill=20, ranges=None, parallel=False):
        self.kind = kind
        self.n = A.n
        rr = None
        nb = 0
        if ranges is not None:
            rr = np.ascontiguousarray(np.asarray(ranges, dtype=np.int64).reshape(-1))
            nb = len(rr) // 2
        self._keep = rr
        create = lib().oracle_precond_create_parallel if parallel else lib().oracle_precond_create
        h = create(kind, A.n, A.row_ptr, A.col, A.val, drop, fill, nb, rr.ctypes.data if rr is not None else None)
        if not h:
            raise RuntimeError("precond failed: " + lib().oracle_last_error().decode())
        self._h = h

    def parallel_blocks(self):
        """Blocks the parallel baseline applies concurrently (0 = sequential)."""
        return int(lib().oracle_precond_blocks(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_precond_free(self._h)
            self._h = None

    def apply(self, v):
        out = np.zeros(self.n)
        lib().oracle_precond_apply(self._h, np.ascontiguousarray(v, np.float64), out)
        return out

    def factors(self):
        nnz = int(lib().oracle_precond_nnz(self._h))
        rp = np.zeros(self.n + 1, np.int64)
        ci = np.zeros(nnz, np.int32)
        v = np.zeros(nnz)
        lib().oracle_precond_factors(self._h, rp, ci, v)
        return CSR(rp, ci, v)


def bicgstab(A: CSR, b, x0=None, precond: Precond | None = None, tol=1e-10, max_iters=1000, max_seconds=0.0):
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    res = C.c_double(0.0)
    it = lib().oracle_bicgstab(A.n, A.row_ptr, A.col, A.val, np.ascontiguousarray(b, np.float64), x,
                               precond._h if precond is not None else None, tol, max_iters, max_seconds,
                               C.byref(res))
    return x, int(it), float(res.value)


def bicgstab_parallel(A: CSR, b, x0=None, precond: Precond | None = None, tol=1e-10, max_iters=1000,
                      max_seconds=0.0):
    """bicgstab on all OpenMP threads (oracle_bicgstab_parallel); returns
    (x, iterations, relative residual, threads)."""
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    res = C.c_double(0.0)
    it = lib().oracle_bicgstab_parallel(A.n, A.row_ptr, A.col, A.val, np.ascontiguousarray(b, np.float64), x,
                                        precond._h if precond is not None else None, tol, max_iters, max_seconds,
                                        C.byref(res))
    return x, int(it), float(res.value), int(lib().oracle_threads())


def density_update(new, cur, alpha, logarithmic=False):
    """Damped Gummel update of a carrier density and its relative update norm.

    update_quantity (simulator.hpp:524-600): linear cur (1 - a) + new a, or
    logarithmic cur^(1-a) * max(new, 1e-10 cur)^a.  get_relative_update_norm
    (simulator.hpp:286-320): sqrt(sum zw^2), zw = (new - upd) / upd, or new
    where upd is 0, between the new density and the updated one (the norm the
    reference's loop reports at its next iteration, simulator.hpp:1081-1086).
    Returns (updated density, norm)."""
    new = np.asarray(new, np.float64)
    cur = np.asarray(cur, np.float64)
    if logarithmic:
        nv = np.where(new < cur * 1e-10, cur * 1e-10, new)
        out = np.power(cur, 1.0 - alpha) * np.power(nv, alpha)
    else:
        out = cur * (1.0 - alpha) + new * alpha
    with np.errstate(divide="ignore", invalid="ignore"):
        zw = np.where(np.abs(out) > 0, (new - out) / np.where(out != 0, out, 1.0), new)
    return out, float(np.sqrt(np.sum(zw * zw)))
