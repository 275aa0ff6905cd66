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

CASES = ["cfg1", "diode_L3", "diode_L5", "nin2d_L3", "mos2d_L1", "mos2d_L3"]


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
