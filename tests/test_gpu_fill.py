"""GPU coefficient fill (shefill) against the CPU oracle's assembly.

Bit-exact: energy count, unknown mapping, expansion orders, CSR pattern,
block ranges.  FP64 values (matrix, right-hand side, unknown scaling, column
factors) within 1e-11 relative per entry: the fill sums the same terms in a
different order (and with fused multiply-adds), so it agrees to rounding.
"""
import numpy as np
import pytest

from conftest import oracle_system

pytestmark = pytest.mark.gpu

CASES = ["diode_L1", "diode_L3", "diode_L5", "nin2d_L3", "mos2d_L1", "mos2d_L3"]


@pytest.mark.parametrize("name", CASES)
def test_fill_matches_oracle(name):
    import shesolve
    s = oracle_system(name)
    F = shesolve.Fill(0)
    sz = F.build(s.dev, s.cfg)
    print(name, sz, "fill ms", F.last_fill_ms())
    assert sz["NH"] == s.NH and sz["n"] == s.n_even and sz["n_odd"] == s.n_odd and sz["nnz"] == s.nnz
    ei, eo = F.mapping()
    oei, oeo, _, _ = s.mapping()
    assert np.array_equal(ei, oei) and np.array_equal(eo, oeo)
    rp, ci, v, rhs, sc, cf = F.system()
    A, b, osc, ocf = s.reduced()
    assert np.array_equal(rp, A.row_ptr) and np.array_equal(ci, A.col)
    assert np.array_equal(F.block_ranges(), s.block_ranges())
    def rel(x, y):
        return np.max(np.abs(x - y) / np.maximum(np.abs(y), 1e-300))
    assert rel(sc, osc) <= 1e-12
    # geometric means of |entries| (exp of a mean of logs): small entries left by
    # cancellation in the eliminated sums carry ~1e-9 relative rounding
    assert rel(cf, ocf) <= 1e-10
    # rows have unit 2-norm after the normalisation; the eliminated sums cancel
    # terms up to ~1e5 times the row norm in the 2D cases, so the absolute
    # agreement is 1e-10; entries above 1e-3 agree to 1e-10 relative
    assert np.max(np.abs(v - A.val)) <= 1e-10
    big = np.abs(A.val) > 1e-3
    assert rel(v[big], A.val[big]) <= 1e-10, rel(v[big], A.val[big])
    assert np.linalg.norm(rhs - b) <= 1e-11 * np.linalg.norm(b)


@pytest.mark.parametrize("name", ["diode_L1", "diode_L3", "diode_L5", "nin2d_L3", "mos2d_L1", "mos2d_L3"])
def test_recovery_and_postproc_match_oracle(name):
    """Scale back + odd recovery and the post-processing (carrier density per
    cell, current per interior facet) on the GPU against the oracle, applied to
    the same vectors: a random reduced vector (the recovery is linear, so every
    odd row is exercised) and the oracle's converged solution."""
    import shesolve
    from oracle import oracle as O
    s = oracle_system(name)
    F = shesolve.Fill(0)
    F.build(s.dev, s.cfg)
    A, b, _, _ = s.reduced()
    rng = np.random.default_rng(11)
    xr = rng.standard_normal(A.n)
    xf = F.recover(xr)
    xf_ref, _ = s.finish(xr)
    assert xf.shape == xf_ref.shape
    assert np.linalg.norm(xf - xf_ref) <= 1e-12 * np.linalg.norm(xf_ref)
    assert np.max(np.abs(xf - xf_ref)) <= 1e-10 * np.max(np.abs(xf_ref))
    P = O.Precond(O.PC_ILUK, A, drop=0.0, fill=1)
    x, it, res = O.bicgstab(A, b, precond=P, tol=1e-12, max_iters=3000)
    assert res <= 1e-12
    xf = F.recover(x)
    xf_ref, full_res = s.finish(x)
    assert np.linalg.norm(xf - xf_ref) <= 1e-12 * np.linalg.norm(xf_ref)
    dens, cur = F.postproc(xf_ref)
    dens_ref, cur_ref = s.postproc(xf_ref)
    assert np.max(np.abs(dens - dens_ref)) <= 1e-12 * np.max(np.abs(dens_ref))
    if len(cur_ref):
        assert np.max(np.abs(cur - cur_ref)) <= 1e-11 * np.max(np.abs(cur_ref))
        # terminal currents (postproc/terminal_current.hpp) on the GPU: against
        # the same oriented sum over the same facet currents on the CPU (1e-12 of
        # the sum of |terms|), and against the oracle's facet currents (1e-8:
        # the facet currents agree to 1e-11 of their maximum, and a terminal
        # current cancels terms of either sign)
        from test_gpu_solver import terminal_currents, terminal_groups
        I_same = terminal_currents(s, cur)
        I_abs = terminal_currents(s, cur, absolute=True)
        I_ref = terminal_currents(s, cur_ref)
        for g, Is, Ia, Ir in zip(terminal_groups(s.dev), I_same, I_abs, I_ref):
            mask = np.zeros(s.dev.ncells, bool)
            mask[g] = True
            I = F.terminal_current(cur, mask)
            assert abs(I - Is) <= 1e-12 * abs(Ia), (I, Is)
            assert abs(I - Ir) <= 1e-8 * np.max(np.abs(I_ref)), (I, Ir)
    F.close()


@pytest.mark.parametrize("logarithmic", [False, True])
@pytest.mark.parametrize("alpha", [1.0, 0.3])
def test_density_update_matches_oracle(logarithmic, alpha):
    """Damped Gummel update of the density and the relative update norm on the
    GPU against the oracle: the oracle's converged density of a 2D device as
    the new value, against a perturbed old density that has zeros (the norm
    takes the new value there) and cells whose new value collapses below
    1e-10 of the old one (the logarithmic form limits them).  Within 1e-14
    relative per cell (pow and fused multiply-adds differ in the last bit),
    1e-12 for the norm (summation order)."""
    import shesolve
    from oracle import oracle as O
    s = oracle_system("nin2d_L3")
    F = shesolve.Fill(0)
    F.build(s.dev, s.cfg)
    A, b, _, _ = s.reduced()
    P = O.Precond(O.PC_ILUK, A, drop=0.0, fill=1)
    x, _, res = O.bicgstab(A, b, precond=P, tol=1e-12, max_iters=3000)
    xf, _ = s.finish(x)
    new, _ = s.postproc(xf)
    rng = np.random.default_rng(5)
    old = new * np.exp(rng.uniform(-1.0, 1.0, new.shape))
    old[::17] = 0.0
    new = new.copy()
    new[5::23] = old[5::23] * 1e-13
    ref, ref_norm = O.density_update(new, old, alpha, logarithmic)
    out, nrm = F.density_update(new, old, alpha, logarithmic)
    assert out.shape == ref.shape
    assert np.max(np.abs(out - ref) / np.maximum(np.abs(ref), 1e-300)) <= 1e-14
    assert nrm == pytest.approx(ref_norm, rel=1e-12)
    with pytest.raises(shesolve.SheSolveError):
        F.density_update(new, old, 0.0)
    F.close()


@pytest.mark.parametrize("name", ["diode_L1", "diode_L3", "diode_L5"])
def test_two_sided_ordering_is_a_permutation(name):
    """ordering = 1 numbers the cells of each energy level from both ends of a
    1D device: the system must be the oracle's with rows and columns permuted
    (pattern bit-exact, values within the fill's 1e-10), its ILU level
    schedule must be shallower, and a solve in that numbering must
    give the same carrier density (1e-6) as the oracle's solution."""
    import dataclasses
    import shesolve
    from oracle import oracle as O
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    oei, oeo, _, _ = s.mapping()
    F = shesolve.Fill(0)
    F.build(s.dev, dataclasses.replace(s.cfg, ordering=1))
    ei, eo = F.mapping()
    assert np.array_equal(eo, oeo)
    assert np.array_equal(ei >= 0, oei >= 0) and not np.array_equal(ei, oei)
    # new row -> old row: the rows of each (cell, energy) node keep their harmonic order
    starts = np.sort(oei[oei >= 0])
    cnt = np.diff(np.append(starts, A.n))
    count_of = dict(zip(starts.tolist(), cnt.tolist()))
    perm = np.empty(A.n, np.int64)
    for e_new, e_old in zip(ei[oei >= 0], oei[oei >= 0]):
        k = count_of[int(e_old)]
        perm[e_new:e_new + k] = np.arange(e_old, e_old + k)
    assert np.array_equal(np.sort(perm), np.arange(A.n))
    iperm = np.empty_like(perm)
    iperm[perm] = np.arange(A.n)
    rp, ci, v, rhs, _, _ = F.system()
    lens = np.diff(A.row_ptr)[perm]
    assert np.array_equal(np.diff(rp), lens)
    for r in range(0, A.n, max(1, A.n // 2000)):
        o = perm[r]
        cols = iperm[A.col[A.row_ptr[o]:A.row_ptr[o + 1]]]
        order = np.argsort(cols)
        assert np.array_equal(ci[rp[r]:rp[r + 1]], cols[order])
        assert np.max(np.abs(v[rp[r]:rp[r + 1]] - A.val[A.row_ptr[o]:A.row_ptr[o + 1]][order])) <= 1e-10
    assert np.linalg.norm(rhs - b[perm]) <= 1e-11 * np.linalg.norm(b)
    kind = dict(precond=shesolve.PRECOND_ILUK, fill_level=1) if name == "diode_L5" else \
        dict(precond=shesolve.PRECOND_ILU0)
    deep = []
    for rows_cols_vals in ((A.row_ptr, A.col, A.val), (rp, ci, v)):
        sv = shesolve.Solver(0, precision=shesolve.FP64, tol=1e-12, max_iters=3000, **kind)
        sv.set_matrix(*rows_cols_vals)
        deep.append(max(sv.levels()))
        sv.close()
    # the cell chain sets most of the depth; the energy couplings add a part
    # that does not shrink (60-100 cell test diodes: 0.53-0.73 of the depth)
    assert deep[1] <= 0.75 * deep[0], deep
    sv = shesolve.Solver(0, precision=shesolve.FP64, tol=1e-12, max_iters=3000, **kind)
    sv.set_matrix(rp, ci, v)
    x, rep = sv.solve(rhs)
    sv.close()
    assert rep["status"] == 0
    dens, _ = F.postproc(F.recover(x))
    P = O.Precond(O.PC_ILUK, A, drop=0.0, fill=1 if name == "diode_L5" else 0)
    x_ref, _, res = O.bicgstab(A, b, precond=P, tol=1e-12, max_iters=3000)
    dens_ref, _ = s.postproc(s.finish(x_ref)[0])
    assert np.linalg.norm(dens - dens_ref) <= 1e-6 * np.linalg.norm(dens_ref)
    F.close()
