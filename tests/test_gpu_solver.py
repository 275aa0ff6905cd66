"""GPU parity of the MI355X solver (libshesolve.so) against the CPU oracle.

Inputs: the reduced, normalised SHE systems the oracle assembles for small
devices (the matrix the reference hands to solvers::solve).  Tolerances:
  - SpMV FP64: |y - y_ref| <= 1e-13 * (|A| |x|) per row (summation order differs);
  - SpMV FP32: |y - y_ref| <= 2e-6 * (|A| |x|) per row;
  - ILU0 factors: pattern bit-exact, values within 1e-10 relative;
  - solves stopped at FP64 relative residual 1e-10 (the metric's stop): true
    residual <= 1e-10, solution within 2e-7 of the oracle's FP64 solution
    (condition number ~1e3 times the residual), carrier density within 1e-6
    (2-norm);
  - solves stopped at 1e-12: solution within 1e-8 relative (2-norm), carrier
    density per cell and terminal currents within 1e-6 relative -- the parity
    bars need this stop on these systems.
"""
import numpy as np
import pytest

from conftest import oracle_system

pytestmark = pytest.mark.gpu

SYSTEMS = ["diode_L1", "diode_L3", "diode_L5", "nin2d_L3", "mos2d_L1", "mos2d_L3"]
MC_WEAK = {"diode_L1", "diode_L3", "nin2d_L3"}
# ILU(0) in any ordering diverges on the L=5 diode (the oracle's residual grows
# to 1e5 in 2000 iterations as well); ILU(1) converges in 9 iterations
ILU0_DIVERGES = {"diode_L5"}


def _solver(**kw):
    import shesolve
    return shesolve.Solver(0, **kw)


def _ref_solution(name):
    """Oracle FP64 solution at 1e-13 (ILU0 + BiCGStab; ILU(1) where ILU0 diverges)."""
    from oracle import oracle as O
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    key = ("refsol", name)
    cache = _ref_solution.__dict__.setdefault("cache", {})
    if key not in cache:
        P = O.Precond(O.PC_ILUK, A, drop=0.0, fill=1) if name in ILU0_DIVERGES else O.Precond(O.PC_ILU0, A)
        x, it, res = O.bicgstab(A, b, precond=P, tol=1e-13, max_iters=5000)
        assert res <= 1e-12, res
        cache[key] = x
    return cache[key]


@pytest.mark.parametrize("name", SYSTEMS)
def test_spmv_fp64_fp32(name):
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    rng = np.random.default_rng(7)
    x = rng.standard_normal(A.n)
    sv = _solver()
    sv.set_matrix(A.row_ptr, A.col, A.val)
    y = sv.spmv(x)
    y_ref = A.matvec(x)
    absA = A.to_scipy().copy()
    absA.data = np.abs(absA.data)
    bound = absA @ np.abs(x)
    assert np.all(np.abs(y - y_ref) <= 1e-13 * bound + 1e-300)
    y32 = sv.spmv_f32(x.astype(np.float32))
    assert np.all(np.abs(y32.astype(np.float64) - y_ref) <= 2e-6 * bound + 1e-30)


@pytest.mark.parametrize("name", ["nin2d_L3", "mos2d_L1", "mos2d_L3"])
def test_spmv_with_column_codes(name):
    """In the block multicolour solve space the pipelined SpMV decodes 16-bit
    column codes (a row block's table of 32-row block starts plus the row
    inside the block) instead of reading 4-byte column indices: the codes are
    in use (the launched format's bytes are nnz x 2 B fewer, plus the tables),
    and the product is the oracle's within the SpMV bars, in FP64 and FP32."""
    import shesolve
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    sv = _solver(precond=shesolve.PRECOND_BMC_ILU0, precision=shesolve.FP64)
    sv.set_matrix(A.row_ptr, A.col, A.val)
    csr = A.nnz * 12 + (A.n + 1) * 4 + 2 * A.n * 8
    coded = sv.spmv_bytes(fp32=False)
    assert csr - 2 * A.nnz < coded < csr - A.nnz, (coded, csr)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(A.n)
    y_ref = A.matvec(x)
    absA = A.to_scipy().copy()
    absA.data = np.abs(absA.data)
    bound = absA @ np.abs(x)
    y = sv.spmv(x)
    assert np.all(np.abs(y - y_ref) <= 1e-13 * bound + 1e-300)
    y32 = sv.spmv_f32(x.astype(np.float32))
    assert np.all(np.abs(y32.astype(np.float64) - y_ref) <= 2e-6 * bound + 1e-30)
    sv.close()


@pytest.mark.parametrize("name", ["diode_L1", "mos2d_L3"])
@pytest.mark.parametrize("kind", ["ilu0", "block_ilu0"])
def test_ilu0_factors_match_oracle(name, kind):
    import shesolve
    from oracle import oracle as O
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    ranges = s.block_ranges()
    pc = shesolve.PRECOND_ILU0 if kind == "ilu0" else shesolve.PRECOND_BLOCK_ILU0
    opc = O.PC_ILU0 if kind == "ilu0" else O.PC_BLOCK_ILU0
    sv = _solver(precond=pc)
    sv.set_matrix(A.row_ptr, A.col, A.val, block_ranges=ranges if kind != "ilu0" else None)
    rp, ci, v = sv.factors()
    F = O.Precond(opc, A, ranges=ranges).factors()
    assert np.array_equal(rp, F.row_ptr) and np.array_equal(ci, F.col)
    scale = np.maximum(np.abs(F.val), 1e-300)
    assert np.max(np.abs(v - F.val) / scale) <= 1e-10
    # preconditioner application
    r = np.random.default_rng(3).standard_normal(A.n)
    z = sv.precond_apply(r)
    z_ref = O.Precond(opc, A, ranges=ranges).apply(r)
    assert np.linalg.norm(z - z_ref) <= 1e-10 * np.linalg.norm(z_ref)


def _pc(kind):
    import shesolve
    return {"ilu0": shesolve.PRECOND_ILU0, "block_ilu0": shesolve.PRECOND_BLOCK_ILU0,
            "mc_ilu0": shesolve.PRECOND_MC_ILU0, "iluk": shesolve.PRECOND_ILUK,
            "bmc_ilu0": shesolve.PRECOND_BMC_ILU0, "block_iluk": shesolve.PRECOND_BLOCK_ILUK}[kind]


def block_diagonal(A, ranges):
    """A without the entries that couple two blocks (the matrix a block
    preconditioner factorises, parallel_linear_solver.hpp:79-93)."""
    from oracle import oracle as O
    blk = np.full(A.n, -1)
    for k, (r0, r1) in enumerate(np.asarray(ranges).reshape(-1, 2)):
        blk[r0:r1] = k
    rows = np.repeat(np.arange(A.n), np.diff(A.row_ptr))
    keep = blk[rows] == blk[A.col]
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=A.n))]).astype(np.int64)
    return O.CSR(rp, A.col[keep], A.val[keep])


@pytest.mark.parametrize("name", ["diode_L1", "diode_L5", "mos2d_L3"])
@pytest.mark.parametrize("k", [0, 1, 2])
def test_block_iluk_factors_match_oracle(name, k):
    """ILU(k) of the per-energy slabs: the factor of the block-diagonal part of A
    (pattern bit-exact, values within 1e-10 relative plus 1e-13 of the largest
    entry) and its application, swept one workgroup per slab."""
    import shesolve
    from oracle import oracle as O
    if k == 0 and name in ILU0_DIVERGES:
        pytest.skip("ILU(0) of the L=5 diode's slabs has near-zero pivots (the diverging ILU(0) of this system): "
                    "rounding in the elimination grows past 1e-10; ILU(1) is its preconditioner, checked below")
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    ranges = s.block_ranges()
    sv = _solver(precond=shesolve.PRECOND_BLOCK_ILUK, fill_level=k)
    sv.set_matrix(A.row_ptr, A.col, A.val, block_ranges=ranges)
    rp, ci, v = sv.factors()
    P = O.Precond(O.PC_ILUK, block_diagonal(A, ranges), drop=0.0, fill=k)
    F = P.factors()
    assert np.array_equal(rp, F.row_ptr) and np.array_equal(ci, F.col)
    tol = 1e-10 * np.abs(F.val) + 1e-13 * np.max(np.abs(F.val))
    assert np.all(np.abs(v - F.val) <= tol), np.max(np.abs(v - F.val) - tol)
    r = np.random.default_rng(5).standard_normal(A.n)
    z = sv.precond_apply(r)
    z_ref = P.apply(r)
    assert np.linalg.norm(z - z_ref) <= 1e-10 * np.linalg.norm(z_ref)


@pytest.mark.parametrize("name", ["diode_L1", "diode_L5", "mos2d_L3"])
@pytest.mark.parametrize("k,drop", [(1, 0.0), (2, 0.0), (2, 1e-3)])
def test_iluk_factors_match_oracle(name, k, drop):
    """ILU(k): level-of-fill pattern bit-exact against the oracle, values within
    1e-10 relative (plus 1e-13 of the largest entry for cancelled fill), and the
    same preconditioner application."""
    import shesolve
    from oracle import oracle as O
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    sv = _solver(precond=shesolve.PRECOND_ILUK, fill_level=k, drop_tol=drop)
    sv.set_matrix(A.row_ptr, A.col, A.val)
    rp, ci, v = sv.factors()
    P = O.Precond(O.PC_ILUK, A, drop=drop, fill=k)
    F = P.factors()
    assert np.array_equal(rp, F.row_ptr) and np.array_equal(ci, F.col)
    tol = 1e-10 * np.abs(F.val) + 1e-13 * np.max(np.abs(F.val))
    assert np.all(np.abs(v - F.val) <= tol), np.max(np.abs(v - F.val) - tol)
    r = np.random.default_rng(4).standard_normal(A.n)
    z = sv.precond_apply(r)
    z_ref = P.apply(r)
    assert np.linalg.norm(z - z_ref) <= 1e-10 * np.linalg.norm(z_ref)


def _colour_order(A, block=1, with_colours=False):
    """Greedy colouring of the symmetrised pattern (explicit zeros included)
    in natural order, smallest free colour; rows sorted by colour (stable).
    With block > 1 the units coloured are blocks of `block` consecutive rows,
    two blocks conflicting when an entry links them.
    Returns (new -> old, the matrix permuted into that order as an oracle CSR
    [, the colour of each natural row])."""
    import scipy.sparse as sp
    from oracle import oracle as O
    n = A.n
    pattern = sp.csr_matrix((np.ones(A.nnz), A.col, A.row_ptr), shape=(n, n))
    S = (pattern + pattern.T).tocsr()
    nb = (n + block - 1) // block
    R = sp.csr_matrix((np.ones(n), (np.arange(n) // block, np.arange(n))), shape=(nb, n))
    SB = (R @ S @ R.T).tocsr()
    bcol = np.full(nb, -1)
    for i in range(nb):
        nbrs = SB.indices[SB.indptr[i]:SB.indptr[i + 1]]
        used = set(bcol[nbrs[nbrs != i]].tolist())
        q = 0
        while q in used:
            q += 1
        bcol[i] = q
    colour = bcol[np.arange(n) // block]
    perm = np.argsort(colour, kind="stable")
    iperm = np.empty(n, np.int64)
    iperm[perm] = np.arange(n)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(np.diff(A.row_ptr)[perm])
    ci = np.empty(A.nnz, np.int32)
    va = np.empty(A.nnz)
    for p, i in enumerate(perm):
        cols = iperm[A.col[A.row_ptr[i]:A.row_ptr[i + 1]]]
        order = np.argsort(cols)
        ci[rp[p]:rp[p + 1]] = cols[order]
        va[rp[p]:rp[p + 1]] = A.val[A.row_ptr[i]:A.row_ptr[i + 1]][order]
    if with_colours:
        return perm, O.CSR(rp, ci, va), colour
    return perm, O.CSR(rp, ci, va)


def _natural_entries(rp, ci, v, perm):
    """Factor entries as (natural row, natural column, value), sorted."""
    rows = perm[np.repeat(np.arange(len(rp) - 1), np.diff(rp))]
    cols = perm[ci]
    order = np.lexsort((cols, rows))
    return rows[order], cols[order], v[order]


@pytest.mark.parametrize("name", ["diode_L1", "mos2d_L1", "mos2d_L3", "nin2d_L3"])
def test_mc_ilu0_factors_match_oracle(name):
    """Multicolour ILU0 is ILU(0) eliminated colour by colour (greedy colouring
    of the symmetrised pattern in natural order, smallest free colour, rows
    sorted by colour).  The device may store the rows in another order (by
    chunks of natural rows, colours sorted inside each; mos2d_L3 spans two
    chunks): its factor, mapped back to natural row and column indices, must
    equal the oracle's ILU(0) of the colour-ordered matrix (same entries, values
    within 1e-10), and M^-1 r in natural order must agree to 1e-10."""
    from oracle import oracle as O
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    sv = _solver(precond=_pc("mc_ilu0"))
    sv.set_matrix(A.row_ptr, A.col, A.val)
    perm_s = sv.ordering().astype(np.int64)
    assert np.array_equal(np.sort(perm_s), np.arange(A.n))
    perm_c, Ap = _colour_order(A)
    P = O.Precond(O.PC_ILU0, Ap)
    F = P.factors()
    rp, ci, v = sv.factors()
    r1, c1, v1 = _natural_entries(rp, ci, v, perm_s)
    r2, c2, v2 = _natural_entries(F.row_ptr, F.col, F.val, perm_c)
    assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
    assert np.max(np.abs(v1 - v2) / np.maximum(np.abs(v2), 1e-300)) <= 1e-10
    lv = sv.levels()
    assert lv[0] <= 16 and lv[1] <= 16, lv
    # application in the natural ordering
    r = np.random.default_rng(5).standard_normal(A.n)
    z = sv.precond_apply(r)
    z_ref = np.empty(A.n)
    z_ref[perm_c] = P.apply(r[perm_c])
    assert np.linalg.norm(z - z_ref) <= 1e-10 * np.linalg.norm(z_ref)


@pytest.mark.parametrize("name", ["diode_L1", "mos2d_L1", "mos2d_L3", "nin2d_L3"])
def test_bmc_ilu0_factors_match_oracle(name):
    """Block multicolour ILU0 is ILU(0) in the order of a greedy colouring of
    blocks of 32 consecutive rows (natural block order, smallest free colour,
    rows sorted by colour, stable).  The device stores exactly that order; its
    factor must equal the oracle's ILU(0) of the reordered matrix (same entries,
    values within 1e-10), blocks of one colour must be independent (no entry
    links two of them), and M^-1 r must agree to 1e-10."""
    from oracle import oracle as O
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    sv = _solver(precond=_pc("bmc_ilu0"))
    sv.set_matrix(A.row_ptr, A.col, A.val)
    perm_s = sv.ordering().astype(np.int64)
    perm_c, Ap, colour = _colour_order(A, block=32, with_colours=True)
    assert np.array_equal(perm_s, perm_c)
    # independence: an entry between two rows of one colour stays inside one
    # block, and a block is one consecutive run of the order
    rows = perm_c[np.repeat(np.arange(A.n), np.diff(Ap.row_ptr))]
    cols = perm_c[Ap.col]
    same_colour = colour[rows] == colour[cols]
    assert np.all((rows // 32 == cols // 32)[same_colour])
    blk_in_order = perm_c // 32
    assert len(np.unique(blk_in_order)) == 1 + np.count_nonzero(np.diff(blk_in_order))
    P = O.Precond(O.PC_ILU0, Ap)
    F = P.factors()
    rp, ci, v = sv.factors()
    assert np.array_equal(rp, F.row_ptr) and np.array_equal(ci, F.col)
    assert np.max(np.abs(v - F.val) / np.maximum(np.abs(F.val), 1e-300)) <= 1e-10
    r = np.random.default_rng(6).standard_normal(A.n)
    z = sv.precond_apply(r)
    z_ref = np.empty(A.n)
    z_ref[perm_c] = P.apply(r[perm_c])
    assert np.linalg.norm(z - z_ref) <= 1e-10 * np.linalg.norm(z_ref)


@pytest.mark.parametrize("name", SYSTEMS)
@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
@pytest.mark.parametrize("kind", ["ilu0", "block_ilu0", "mc_ilu0", "iluk", "bmc_ilu0", "block_iluk"])
def test_solve_parity(name, precision, kind):
    import shesolve
    if kind not in ("iluk", "block_iluk") and name in ILU0_DIVERGES:
        pytest.skip("ILU(0) in natural, block or multicolour form diverges on the L=5 diode (so does the "
                    "oracle's); ILU(k) is the preconditioner for it")
    if kind == "mc_ilu0" and name in MC_WEAK:
        pytest.skip("multicolour ILU0 on a 1D chain (or a thin 2D strip) decouples neighbouring cells across colours; "
                    "BiCGStab then needs thousands of iterations (diode_L1: 2541 vs 26 with natural ILU0), so natural "
                    "ILU0 is the preconditioner for these devices")
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    pc = _pc(kind)
    prec = shesolve.FP64 if precision == "fp64" else shesolve.FP32_IR
    sv = _solver(precond=pc, precision=prec, tol=1e-10, max_iters=5000, fill_level=1)
    sv.set_matrix(A.row_ptr, A.col, A.val, block_ranges=s.block_ranges() if kind.startswith("block") else None)
    # FP64 relative residual <= 1e-10
    x, rep = sv.solve(b)
    assert rep["status"] == 0
    res = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    assert res <= 1e-10, res
    assert rep["rel_residual"] <= 1e-10
    # At the metric's 1e-10 stop the solution error is bounded by the condition
    # number (about 1e3 on these systems) times the residual: measured up to
    # 1.2e-7 (the 1D diodes), so 2e-7 here; the density holds the 1e-6 bar in the
    # 2-norm, per cell it and the terminal currents miss it by up to 3x.
    x_ref = _ref_solution(name)
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print(name, precision, kind, rep, "error at the 1e-10 stop", err)
    assert err <= 2e-7, err
    assert density_2norm(s, x, x_ref) <= 1e-6
    # The parity bars -- solution within 1e-8, density per cell and terminal
    # currents within 1e-6 -- hold with the stop at 1e-12.
    sv.set_params(tol=1e-12)
    x, rep = sv.solve(b)
    assert rep["status"] == 0 and rep["rel_residual"] <= 1e-12
    err12 = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print(name, precision, kind, "error at the 1e-12 stop", err12)
    assert err12 <= 1e-8, err12
    check_postproc(s, x, x_ref)


def terminal_groups(dev):
    """Contact cells split by connectivity order (left/right in 1D;
    source/drain/body in 2D)."""
    ids = np.nonzero(np.asarray(dev.contact, bool))[0]
    nx = dev.nx
    if dev.dim == 1:
        return [np.array([ids[0]]), np.array([ids[-1]])]
    if dev.name.startswith("nin2d"):
        return [ids[ids % nx == 0], ids[ids % nx == nx - 1]]
    top = ids[ids // nx == dev.ny - 1]
    bottom = ids[ids // nx == 0]
    return [top[top % nx < nx // 2], top[top % nx >= nx // 2], bottom]


def terminal_currents(s, j, absolute=False):
    """Net current out of each contact group: sum of facet currents (times the
    facet area) over facets with exactly one cell in the group, oriented from
    the contact into the device.  Groups: contact cells split by connectivity
    order (left/right in 1D; source/drain/body in 2D)."""
    dev = s.dev
    c1, c2, d = s.facets()
    out = []
    for g in terminal_groups(dev):
        inside = np.zeros(dev.ncells, bool)
        inside[g] = True
        m1, m2 = inside[c1], inside[c2]
        sel = m1 ^ m2
        area = np.where(d == 0, np.diff(dev.yv)[c1 // dev.nx] if dev.dim > 1 else 1.0,
                        np.diff(dev.xv)[c1 % dev.nx])
        sign = np.where(m1, 1.0, -1.0)
        terms = (j * area * sign)[sel]
        out.append(float(np.sum(np.abs(terms) if absolute else terms)))
    return np.array(out)


def density_2norm(s, x, x_ref):
    """Relative 2-norm difference of the carrier densities of two reduced solutions."""
    n, _ = s.postproc(s.finish(x)[0])
    n_ref, _ = s.postproc(s.finish(x_ref)[0])
    return np.linalg.norm(n - n_ref) / np.linalg.norm(n_ref)


def check_postproc(s, x, x_ref):
    """Density within 1e-6 (2-norm relative; per cell where the density is within
    four decades of its maximum) and terminal currents within 1e-6 relative.
    Minority-carrier densities 10+ decades below the majority are below the
    resolution of any residual-based stop on this formulation: an FP64 direct
    solve and BiCGStab to 1e-14 already disagree there."""
    xf, _ = s.finish(x)
    xf_ref, _ = s.finish(x_ref)
    n, j = s.postproc(xf)
    n_ref, j_ref = s.postproc(xf_ref)
    assert np.linalg.norm(n - n_ref) <= 1e-6 * np.linalg.norm(n_ref)
    big = np.abs(n_ref) >= 1e-4 * np.max(np.abs(n_ref))
    assert np.max(np.abs(n - n_ref)[big] / np.abs(n_ref[big])) <= 1e-6
    # facet currents within 1e-6 of their largest
    assert np.max(np.abs(j - j_ref)) <= 1e-6 * np.max(np.abs(j_ref))
    I, I_ref = terminal_currents(s, j), terminal_currents(s, j_ref)
    if s.dev.name.startswith("mosfet2d"):
        # no channel: each terminal's net current is the small difference of
        # opposite facet fluxes (the off-state current of the p-body is carried
        # by minority electrons), so it is compared on the scale of those
        # fluxes, the sum of their magnitudes over the terminal's facets
        flux = terminal_currents(s, j_ref, absolute=True)
        assert np.max(np.abs(I - I_ref)) <= 1e-6 * np.max(flux), (I, I_ref, flux)
        return
    scale = np.max(np.abs(I_ref))
    assert np.max(np.abs(I - I_ref)) <= 1e-6 * scale, (I, I_ref)
    # current conservation (tests/src/ushape_2d.cpp:196-204): |I_L + I_R| / max < 1e-5
    assert abs(I_ref[0] + I_ref[1]) <= 1e-5 * scale


def test_zero_rhs_and_warm_start():
    s = oracle_system("diode_L1")
    A, b, _, _ = s.reduced()
    sv = _solver(tol=1e-10)
    sv.set_matrix(A.row_ptr, A.col, A.val)
    x, rep = sv.solve(np.zeros(A.n))
    assert np.all(x == 0) and rep["iterations"] == 0
    x, rep = sv.solve(b)
    assert rep["status"] == 0
    x2, rep2 = sv.solve(b, x0=x)
    assert rep2["iterations"] == 0 and rep2["rel_residual"] <= 1e-10


def test_invalid_rhs_is_rejected():
    import shesolve
    s = oracle_system("diode_L1")
    A, b, _, _ = s.reduced()
    sv = _solver()
    sv.set_matrix(A.row_ptr, A.col, A.val)
    bb = b.copy()
    bb[5] = np.nan
    with pytest.raises(shesolve.SheSolveError, match="non-finite"):
        sv.solve(bb)


def test_failed_setup_leaves_no_matrix():
    """A matrix the analysis rejects (a row without a diagonal entry, a row
    pointer that decreases) leaves the handle without a matrix: a later solve
    reports "no matrix set" instead of running on a half-built factor."""
    import shesolve
    s = oracle_system("diode_L1")
    A, b, _, _ = s.reduced()
    sv = _solver(precond=shesolve.PRECOND_ILU0)
    sv.set_matrix(A.row_ptr, A.col, A.val)
    # drop the diagonal entry of row 3
    rp, ci, v = A.row_ptr.copy(), A.col.copy(), A.val.copy()
    k = int(np.nonzero(ci[rp[3]:rp[4]] == 3)[0][0]) + int(rp[3])
    ci, v = np.delete(ci, k), np.delete(v, k)
    rp[4:] -= 1
    with pytest.raises(shesolve.SheSolveError, match="no diagonal"):
        sv.set_matrix(rp, ci, v)
    with pytest.raises(shesolve.SheSolveError, match="no matrix set"):
        sv.solve(b)
    bad = A.row_ptr.copy()
    bad[5], bad[6] = bad[6], bad[5]
    with pytest.raises(shesolve.SheSolveError, match="decreases"):
        sv.set_matrix(bad, A.col, A.val)
    with pytest.raises(shesolve.SheSolveError, match="no matrix set"):
        sv.solve(b)
    # the handle takes a good matrix again
    sv.set_matrix(A.row_ptr, A.col, A.val)
    x, rep = sv.solve(b)
    assert rep["status"] == 0


def test_unconverged_fp64_reports_the_true_residual():
    """An FP64 solve stopped by its iteration cap reports ||b - A x|| / ||b||
    of the returned iterate, not the BiCGStab recurrence residual."""
    import shesolve
    s = oracle_system("mos2d_L3")
    A, b, _, _ = s.reduced()
    sv = _solver(precond=shesolve.PRECOND_ILU0, precision=shesolve.FP64, max_iters=3, tol=1e-14)
    sv.set_matrix(A.row_ptr, A.col, A.val)
    x, rep = sv.solve(b)
    assert rep["status"] != 0
    true = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    assert rep["rel_residual"] == pytest.approx(true, rel=1e-8)


def test_update_values_refactorises():
    s = oracle_system("diode_L3")
    A, b, _, _ = s.reduced()
    sv = _solver(tol=1e-10)
    sv.set_matrix(A.row_ptr, A.col, A.val)
    x1, _ = sv.solve(b)
    sv.update_values(2.0 * A.val)
    x2, rep = sv.solve(b)
    assert rep["status"] == 0
    assert np.linalg.norm(2.0 * x2 - x1) <= 1e-8 * np.linalg.norm(x1)


@pytest.mark.parametrize("name,kind", [("diode_L1", "ilu0"), ("diode_L5", "iluk"), ("mos2d_L3", "mc_ilu0")])
def test_refinement_history_against_the_oracle(name, kind):
    """Per-refinement-step report of the FP32 + FP64-refinement solve: each
    record holds the FP64 residual of the iterate and, with a reference
    solution set, its error against the oracle's FP64 solution (1e-13)."""
    import shesolve
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    x_ref = _ref_solution(name)
    sv = _solver(precond=_pc(kind), precision=shesolve.FP32_IR, tol=1e-10, max_iters=5000, fill_level=1)
    sv.set_matrix(A.row_ptr, A.col, A.val)
    x0, rep0 = sv.solve(b)
    h0 = sv.ir_history()
    assert all(np.isnan(r["rel_error"]) for r in h0)
    sv.set_reference_solution(x_ref)
    x, rep = sv.solve(b)
    h = sv.ir_history()
    print(name, kind, rep, h)
    # the reference only adds the error evaluation: same iterates, same residuals
    assert np.array_equal(x, x0)
    assert [r["rel_residual"] for r in h] == [r["rel_residual"] for r in h0]
    # one record for the initial guess (x = 0: residual and error exactly 1),
    # then one per refinement step
    assert h[0]["step"] == 0 and h[0]["inner_iterations"] == 0 and h[0]["kind"] == 0
    assert abs(h[0]["rel_residual"] - 1.0) <= 1e-12 and abs(h[0]["rel_error"] - 1.0) <= 1e-12
    assert [r["step"] for r in h if r["kind"] != 2] == list(range(len([r for r in h if r["kind"] != 2])))
    assert sum(r["inner_iterations"] for r in h) == rep["iterations"]
    assert h[-1]["rel_residual"] == rep["rel_residual"] <= 1e-10
    # the last error is the host's ||x - x_ref|| / ||x_ref||
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    assert abs(h[-1]["rel_error"] - err) <= 1e-9 * err + 1e-15, (h[-1]["rel_error"], err)
    assert err <= 1e-6
    # accepted corrections reduce the FP64 residual
    acc = [r["rel_residual"] for r in h if r["kind"] == 0]
    assert all(r2 < r1 for r1, r2 in zip(acc, acc[1:]))
    # cleared: errors are NaN again
    sv.set_reference_solution(None)
    sv.solve(b)
    assert all(np.isnan(r["rel_error"]) for r in sv.ir_history())
    # a plain FP64 solve leaves one record, kind 2
    sv.set_reference_solution(x_ref)
    sv.set_params(precision=shesolve.FP64)
    x64, rep64 = sv.solve(b)
    h64 = sv.ir_history()
    assert len(h64) == 1 and h64[0]["kind"] == 2 and h64[0]["inner_iterations"] == rep64["iterations"]
    assert h64[0]["rel_residual"] == rep64["rel_residual"]
    err64 = np.linalg.norm(x64 - x_ref) / np.linalg.norm(x_ref)
    assert abs(h64[0]["rel_error"] - err64) <= 1e-9 * err64 + 1e-15


def test_failed_refinement_returns_its_best_iterate():
    """ILU(0) diverges on the L=5 diode (so does the oracle's): the FP32 inner
    solves blow up, the refinement hands over to FP64 BiCGStab, which diverges
    as well.  The solve reports the failure and returns the best iterate it
    reached: the reported residual is the true residual of the returned x, no
    larger than that of x = 0 or of the best accepted refinement step."""
    import shesolve
    s = oracle_system("diode_L5")
    A, b, _, _ = s.reduced()
    sv = _solver(precond=_pc("ilu0"), precision=shesolve.FP32_IR, tol=1e-10, max_iters=2000)
    sv.set_matrix(A.row_ptr, A.col, A.val)
    x, rep = sv.solve(b)
    h = sv.ir_history()
    print(rep, h)
    assert rep["status"] != 0
    res = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    assert np.isfinite(res) and res <= 1.0
    assert abs(res - rep["rel_residual"]) <= 1e-6 * res
    assert h[-1]["kind"] == 2 and h[-1]["rel_residual"] == rep["rel_residual"]
    assert rep["rel_residual"] <= min(r["rel_residual"] for r in h if r["kind"] == 0)
