"""Batched solves (shesolve_batch_*) against the oracle and against single solves.

Each system of a batch must meet the single-solve bar on its own: FP64
relative residual <= 1e-10, solution within 1e-8 (2-norm, relative) of the
oracle's FP64 solution when solved to 1e-12.  A system's result is the same
alone and in the batch within that tolerance, and the graph-replayed
iteration gives results bit-identical to the same kernels launched one by one.
"""
import numpy as np
import pytest

from conftest import oracle_system

pytestmark = pytest.mark.gpu

MIX = ["diode_L1", "diode_L3", "nin2d_L3", "mos2d_L1", "mos2d_L3"]


def _ref(name):
    from oracle import oracle as O
    cache = _ref.__dict__.setdefault("cache", {})
    if name not in cache:
        s = oracle_system(name)
        A, b, _, _ = s.reduced()
        P = O.Precond(O.PC_ILUK, A, drop=0.0, fill=1) if name == "diode_L5" else O.Precond(O.PC_ILU0, A)
        x, it, res = O.bicgstab(A, b, precond=P, tol=1e-13, max_iters=5000)
        assert res <= 1e-12
        cache[name] = x
    return cache[name]


def _systems(names):
    out = []
    for nm in names:
        A, b, _, _ = oracle_system(nm).reduced()
        out.append((A, b))
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
@pytest.mark.parametrize("kind", ["ilu0", "block_ilu0", "iluk", "block_iluk"])
def test_batch_parity(precision, kind):
    import shesolve
    # ILU(1), whole or per energy slab, also takes the L=5 diode, on which ILU(0) diverges
    names = MIX + ["diode_L5"] if kind in ("iluk", "block_iluk") else MIX
    sy = _systems(names)
    pc = {"ilu0": shesolve.PRECOND_ILU0, "block_ilu0": shesolve.PRECOND_BLOCK_ILU0,
          "iluk": shesolve.PRECOND_ILUK, "block_iluk": shesolve.PRECOND_BLOCK_ILUK}[kind]
    prec = shesolve.FP64 if precision == "fp64" else shesolve.FP32_IR
    bs = shesolve.BatchSolver(0, precond=pc, precision=prec, tol=1e-10, max_iters=5000, fill_level=1)
    ranges = [oracle_system(nm).block_ranges() for nm in names] if kind.startswith("block") else None
    bs.set_systems([(A.row_ptr, A.col, A.val) for A, _ in sy], block_ranges=ranges)
    info = bs.info()
    assert info["systems"] == len(names) and info["rows"] == sum(len(b) for _, b in sy)
    xs, reps = bs.solve([b for _, b in sy])
    for nm, (A, b), x, rep in zip(names, sy, xs, reps):
        print(nm, precision, kind, rep)
        res = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
        assert rep["status"] == 0 and rep["rel_residual"] <= 1e-10 and res <= 1e-10, (nm, rep, res)
    bs.set_params(tol=1e-12)
    xs, reps = bs.solve([b for _, b in sy])
    for nm, x, rep in zip(names, xs, reps):
        assert rep["status"] == 0 and rep["rel_residual"] <= 1e-12, (nm, rep)
        xr = _ref(nm)
        assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr), nm
    bs.close()


def test_batch_equals_single_solves():
    """Each result is the same alone (a single-system handle) and in the batch."""
    import shesolve
    names = MIX
    sy = _systems(names)
    bs = shesolve.BatchSolver(0, precond=shesolve.PRECOND_ILU0, precision=shesolve.FP32_IR, tol=1e-12,
                              max_iters=5000)
    bs.set_systems([(A.row_ptr, A.col, A.val) for A, _ in sy])
    xs, reps = bs.solve([b for _, b in sy])
    for nm, (A, b), xb in zip(names, sy, xs):
        sv = shesolve.Solver(0, precond=shesolve.PRECOND_ILU0, precision=shesolve.FP32_IR, tol=1e-12,
                             max_iters=5000)
        sv.set_matrix(A.row_ptr, A.col, A.val)
        x1, rep = sv.solve(b)
        assert rep["status"] == 0
        assert np.linalg.norm(xb - x1) <= 1e-8 * np.linalg.norm(x1), nm
        sv.close()
    # a one-system batch is the system alone
    for nm, (A, b) in zip(names[:2], sy[:2]):
        one = shesolve.BatchSolver(0, precond=shesolve.PRECOND_ILU0, precision=shesolve.FP32_IR, tol=1e-12,
                                   max_iters=5000)
        one.set_systems([(A.row_ptr, A.col, A.val)])
        (x1,), (rep,) = one.solve([b])
        assert rep["status"] == 0
        xb = xs[names.index(nm)]
        assert np.linalg.norm(xb - x1) <= 1e-8 * np.linalg.norm(x1), nm
        one.close()
    bs.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_graph_replay_matches_direct_launches(precision):
    """The captured iteration, replayed many times with host checks in between,
    gives the same numbers as the kernels launched one by one."""
    import shesolve
    names = ["diode_L1", "nin2d_L3", "mos2d_L1"]
    sy = _systems(names)
    prec = shesolve.FP64 if precision == "fp64" else shesolve.FP32_IR
    out = []
    for graph in (False, True, True):
        # launch shape pinned: the autotuner may pick a different SpMV kernel per
        # handle, and the comparison is between handles
        bs = shesolve.BatchSolver(0, precond=shesolve.PRECOND_ILU0, precision=prec, tol=1e-10, max_iters=5000,
                                  autotune=0)
        bs.options(use_graph=graph, poll=3)
        bs.set_systems([(A.row_ptr, A.col, A.val) for A, _ in sy])
        xs, reps = bs.solve([b for _, b in sy])
        # a second solve on the same handle replays the instantiated graph again
        xs2, reps2 = bs.solve([b for _, b in sy])
        out.append((xs, reps, xs2, reps2))
        bs.close()
    x0, r0, x0b, _ = out[0]
    for xs, reps, xs2, _ in out[1:]:
        for a, b_, c in zip(x0, xs, xs2):
            assert np.array_equal(a, b_) and np.array_equal(a, c)
        assert [r["iterations"] for r in reps] == [r["iterations"] for r in r0]
    for a, b_ in zip(x0, x0b):
        assert np.array_equal(a, b_)


def test_batch_zero_rhs_warm_start_and_update_values():
    import shesolve
    names = ["diode_L1", "mos2d_L1"]
    sy = _systems(names)
    bs = shesolve.BatchSolver(0, precond=shesolve.PRECOND_ILU0, precision=shesolve.FP32_IR, tol=1e-12,
                              max_iters=5000)
    bs.set_systems([(A.row_ptr, A.col, A.val) for A, _ in sy])
    # a zero right-hand side gives x = 0 for that system only (src/solvers/solve.hpp:61-66)
    xs, reps = bs.solve([np.zeros_like(sy[0][1]), sy[1][1]])
    assert not np.any(xs[0]) and reps[0]["status"] == 0 and reps[0]["iterations"] == 0
    assert reps[1]["status"] == 0 and reps[1]["rel_residual"] <= 1e-12
    # warm start from the converged solution needs no iterations
    xs, reps = bs.solve([b for _, b in sy])
    xw, repw = bs.solve([b for _, b in sy], x0s=xs)
    assert all(r["iterations"] == 0 and r["status"] == 0 for r in repw)
    # new values: scale system 1 by 2 -> its solution halves
    bs.update_values([sy[0][0].val, 2.0 * sy[1][0].val])
    xs2, reps2 = bs.solve([b for _, b in sy])
    assert all(r["status"] == 0 for r in reps2)
    assert np.linalg.norm(xs2[1] - 0.5 * xs[1]) <= 1e-8 * np.linalg.norm(xs[1])
    assert np.linalg.norm(xs2[0] - xs[0]) <= 1e-8 * np.linalg.norm(xs[0])
    bs.close()


def test_batch_rejects_bad_input():
    import shesolve
    A, b, _, _ = oracle_system("diode_L1").reduced()
    bs = shesolve.BatchSolver(0)
    with pytest.raises(shesolve.SheSolveError):
        bs.set_systems([])
    bs.set_systems([(A.row_ptr, A.col, A.val)])
    bad = b.copy()
    bad[3] = np.nan
    with pytest.raises(shesolve.SheSolveError):
        bs.solve([bad])
    bs.close()


@pytest.mark.parametrize("kind", ["ilu0", "iluk"])
def test_batch_preconditioner_is_the_per_system_ilu(kind):
    """The one-workgroup-per-system sweeps apply exactly each system's own ILU:
    M^-1 r of the batch equals, system by system, the oracle's ILU applied to
    that system alone (ILU(1) for the L=5 diode under iluk)."""
    import shesolve
    from oracle import oracle as O
    names = ["diode_L1", "diode_L3", "nin2d_L3", "mos2d_L1"] + (["diode_L5"] if kind == "iluk" else [])
    sy = _systems(names)
    pc = shesolve.PRECOND_ILU0 if kind == "ilu0" else shesolve.PRECOND_ILUK
    bs = shesolve.BatchSolver(0, precond=pc, tol=1e-10, max_iters=5000)
    if kind == "iluk":
        bs.set_fill_levels([1 if nm == "diode_L5" else 0 for nm in names])
    bs.set_systems([(A.row_ptr, A.col, A.val) for A, _ in sy])
    rng = np.random.default_rng(9)
    r = rng.standard_normal(sum(len(b) for _, b in sy))
    z = bs.precond_apply(r)
    off = 0
    for nm, (A, b) in zip(names, sy):
        k = 1 if (kind == "iluk" and nm == "diode_L5") else 0
        P = O.Precond(O.PC_ILUK, A, drop=0.0, fill=k)
        zr = P.apply(r[off:off + A.n])
        zs = z[off:off + A.n]
        assert np.linalg.norm(zs - zr) <= 1e-10 * np.linalg.norm(zr), nm
        off += A.n
    bs.close()


def test_batch_block_iluk_is_the_per_slab_ilu():
    """Per-energy block ILU(k) in a batch, one workgroup per slab of every
    system: M^-1 r equals, system by system, the oracle's ILU(k) of that
    system's block-diagonal part (k = 1 for the L=5 diode, 0 otherwise)."""
    import shesolve
    from oracle import oracle as O
    from test_gpu_solver import block_diagonal
    names = ["diode_L1", "diode_L3", "nin2d_L3", "mos2d_L1", "diode_L5"]
    sy = _systems(names)
    ranges = [oracle_system(nm).block_ranges() for nm in names]
    levels = [1 if nm == "diode_L5" else 0 for nm in names]
    bs = shesolve.BatchSolver(0, precond=shesolve.PRECOND_BLOCK_ILUK, tol=1e-10, max_iters=5000)
    bs.set_fill_levels(levels)
    bs.set_systems([(A.row_ptr, A.col, A.val) for A, _ in sy], block_ranges=ranges)
    r = np.random.default_rng(11).standard_normal(sum(len(b) for _, b in sy))
    z = bs.precond_apply(r)
    off = 0
    for nm, (A, b), rg, k in zip(names, sy, ranges, levels):
        P = O.Precond(O.PC_ILUK, block_diagonal(A, rg), drop=0.0, fill=k)
        zr = P.apply(r[off:off + A.n])
        assert np.linalg.norm(z[off:off + A.n] - zr) <= 1e-10 * np.linalg.norm(zr), nm
        off += A.n
    bs.close()


def _wide(n, k, seed):
    """A system whose ILU(0) level schedule is shallow and wide: row i couples
    to i +- k (levels of k rows), diagonally dominant, with random values."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        cols = [c for c in (i - k, i, i + k) if 0 <= c < n]
        rows.append(cols)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate([np.array(r, np.int32) for r in rows])
    val = np.where(ci == np.repeat(np.arange(n), np.diff(rp)), 4.0, -rng.uniform(0.5, 1.5, len(ci)))
    b = rng.standard_normal(n)
    return rp, ci, val, b


def test_batch_level_launch_sweeps():
    """Shallow, wide schedules (levels of 30k-40k rows) are swept by one launch
    per level over all systems instead of one workgroup per system.  That path
    must apply each system's own ILU(0) (against the oracle's, 1e-10), solve
    each system to the bar, and agree with the single solves of the same
    systems."""
    import scipy.sparse as sp
    import shesolve
    from oracle import oracle as O
    systems = [_wide(400_000, 40_000, 1), _wide(300_000, 30_000, 2)]
    bs = shesolve.BatchSolver(0, precond=shesolve.PRECOND_ILU0, precision=shesolve.FP64, tol=1e-10,
                              max_iters=2000)
    bs.set_systems([(rp, ci, v) for rp, ci, v, _ in systems])
    assert bs.info()["seg_levels_max"] == 0  # no segment schedules: the level launches run
    r = np.random.default_rng(3).standard_normal(sum(len(b) for *_, b in systems))
    z = bs.precond_apply(r)
    xs, reps = bs.solve([b for *_, b in systems])
    off = 0
    for (rp, ci, v, b), x, rep in zip(systems, xs, reps):
        n = len(b)
        A = O.CSR(rp, ci, v)
        zr = O.Precond(O.PC_ILU0, A).apply(r[off:off + n])
        assert np.linalg.norm(z[off:off + n] - zr) <= 1e-10 * np.linalg.norm(zr)
        off += n
        M = sp.csr_matrix((v, ci, rp), shape=(n, n))
        assert rep["status"] == 0
        assert np.linalg.norm(b - M @ x) <= 1e-10 * np.linalg.norm(b)
        sv = shesolve.Solver(0, precond=shesolve.PRECOND_ILU0, precision=shesolve.FP64, tol=1e-10, max_iters=2000)
        sv.set_matrix(rp, ci, v)
        x1, rep1 = sv.solve(b)
        sv.close()
        assert rep1["status"] == 0
        assert np.linalg.norm(x1 - x) <= 1e-8 * np.linalg.norm(x1)
    bs.close()


def test_batch_failed_refinement_returns_its_best_iterate():
    """One system of the batch (the L=5 diode under ILU(0), which diverges in
    both precisions, the oracle's too) fails; the others still converge.  The
    failed system reports the failure and gets back its best iterate, whose true
    residual is the reported one and no larger than that of x = 0."""
    import shesolve
    names = ["diode_L1", "diode_L5", "mos2d_L1"]
    sy = _systems(names)
    bs = shesolve.BatchSolver(0, precond=shesolve.PRECOND_ILU0, precision=shesolve.FP32_IR, tol=1e-10,
                              max_iters=2000)
    bs.set_systems([(A.row_ptr, A.col, A.val) for A, _ in sy])
    xs, reps = bs.solve([b for _, b in sy])
    for nm, (A, b), x, rep in zip(names, sy, xs, reps):
        print(nm, rep)
        res = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
        if nm == "diode_L5":
            assert rep["status"] != 0
            assert np.isfinite(res) and res <= 1.0
            assert abs(res - rep["rel_residual"]) <= 1e-6 * res
        else:
            assert rep["status"] == 0 and res <= 1e-10, (nm, rep, res)
    bs.close()
