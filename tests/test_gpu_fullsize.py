"""Parity at full size: the 2D L=3 MOSFET-like device of 161 x 81 vertices
(200 x 100 nm, 10 meV to 1.5 eV; n = 16.1 M unknowns, 165 M entries), the
system of the plain FP64 workload and of its FP32 + FP64-refinement form.

- Fill: the GPU coefficient fill against the oracle's assembly of the same
  device (she::assemble + she::solve up to the solver call): mapping, block
  ranges and pattern bit-exact, unknown scaling within 1e-12.  The column
  factors of row_normalize_system (math/linalg_util.hpp:200-282) are geometric
  means over every stored entry of a column, and at this size some columns
  hold an entry that the odd elimination cancels to ~1e-11 of its neighbours:
  that entry carries a relative rounding of ~1e-5 (one ulp of its O(1) terms,
  summed in another order), which moves its column's factor by ~1e-6 and with
  it every value of the column (145 752 of 165 M entries differ by more than
  1e-10).  So the values are compared with both scalings undone -- each row's
  entries relative to its diagonal, (v_ij / c_j) / (v_ii / c_i), which is the
  unscaled eliminated system up to a row factor -- within 1e-10 of the row's
  largest, the right-hand side likewise, the column factors within 1e-4.  The
  oracle needs about 75 s and 26 GB for this build.
  Solutions are compared in the GPU system's scaling (the oracle's solution
  times its column factors over the GPU's), each residual on its own system.
- Solves at the metric's stop, 1e-10: the plain FP64 BiCGStab with
  per-energy block ILU0 (the workload's method) and the FP32 + FP64
  refinement with block multicolour ILU0 (the benchmark's), each with the
  FP64 relative residual recomputed here from the fill's system <= 1e-10,
  the solution within 2e-7 of the oracle's FP64 solution and the carrier
  density within 1e-6 (2-norm).
- The parity bars at the 1e-12 stop: solution within 1e-8, density within 1e-6
  per cell (where within four decades of the maximum).
The oracle's solution: its block-ILU0 BiCGStab on all host threads to 1e-12.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def shesolve_csr(rp, ci, v):
    import scipy.sparse as sp
    return sp.csr_matrix((v, ci, rp), shape=(len(rp) - 1, len(rp) - 1))


def _say(t0, msg):
    print(f"[full size {time.perf_counter() - t0:6.1f} s] {msg}", flush=True)


def test_full_size_mosfet_fill_and_solves():
    import shesolve
    from oracle import oracle as O
    from shesolve import devices as D
    t0 = time.perf_counter()
    dev, cfg = D.mosfet_200x100nm(vx=161, vy=81, L=3)
    _say(t0, "device built (nonlinear Poisson potential)")
    F = shesolve.Fill(0)
    sz = F.build(dev, cfg)
    _say(t0, f"GPU fill: {sz}, device fill {F.last_fill_ms():.1f} ms")
    s = O.OracleSystem(dev, cfg)
    _say(t0, f"oracle assembly: n {s.n_even}, odd {s.n_odd}")
    assert sz["NH"] == s.NH and sz["n"] == s.n_even and sz["n_odd"] == s.n_odd and sz["nnz"] == s.nnz
    ei, eo = F.mapping()
    oei, oeo, _, _ = s.mapping()
    assert np.array_equal(ei, oei) and np.array_equal(eo, oeo)
    del ei, eo, oei, oeo
    rp, ci, v, rhs, sc, cf = F.system()
    A, b, osc, ocf = s.reduced()
    assert np.array_equal(rp, A.row_ptr) and np.array_equal(ci, A.col)
    ranges = F.block_ranges()
    assert np.array_equal(ranges, s.block_ranges())

    def rel(x, y):
        return np.max(np.abs(x - y) / np.maximum(np.abs(y), 1e-300))
    assert rel(sc, osc) <= 1e-12
    assert rel(cf, ocf) <= 1e-4
    rows = np.repeat(np.arange(A.n), np.diff(A.row_ptr))
    is_diag = ci == rows
    assert np.array_equal(np.bincount(rows[is_diag], minlength=A.n), np.ones(A.n, np.int64))
    dg = np.zeros(A.n)
    odg = np.zeros(A.n)
    # a normalised value is v_ij = a_ij c_j / r_i (c the column factors, r the
    # row norms): v_ij / c_j over v_ii / c_i is a_ij / a_ii
    dg[rows[is_diag]] = v[is_diag] / cf[rows[is_diag]]
    odg[rows[is_diag]] = A.val[is_diag] / ocf[rows[is_diag]]
    r_g = v / cf[ci] / dg[rows]
    r_o = A.val / ocf[ci] / odg[rows]
    scale = np.maximum.reduceat(np.abs(r_o), A.row_ptr[:-1])
    assert np.max(np.abs(r_g - r_o) / scale[rows]) <= 1e-10
    assert np.max(np.abs(rhs / dg - b / odg)) <= 1e-10 * np.max(np.abs(b / odg))
    del rows, is_diag, dg, odg, r_g, r_o, scale
    gsys = shesolve_csr(rp, ci, v)
    del rp, ci, v, osc
    _say(t0, "fill matches the oracle (both scalings undone)")

    P = O.Precond(O.PC_BLOCK_ILU0, A, ranges=ranges, parallel=True)
    x_ref, it_ref, res_ref, threads = O.bicgstab_parallel(A, b, precond=P, tol=1e-12, max_iters=5000)
    del P
    assert res_ref <= 1e-12, res_ref
    _say(t0, f"oracle block-ILU0 BiCGStab to 1e-12 on {threads} threads: {it_ref} iterations")
    bn = np.linalg.norm(rhs)
    # the oracle's solution in the GPU system's column scaling (x c is the same
    # unknown in both); the 2-norm bar is taken there, in the balanced space the
    # solver works in -- without the column factors (x c alone) the hot-carrier
    # unknowns, up to many decades above exp(-E/kT), would dominate the norm
    xs_ref = x_ref * ocf / cf
    xf_ref, _ = s.finish(x_ref)
    n_ref, _ = s.postproc(xf_ref)
    del xf_ref

    d_rp, d_ci, d_val, d_rhs = F.device_system()
    from shesolve import serving
    dx = serving.DeviceBuffer(A.n)
    try:
        cases = [("fp64 block_ilu0", dict(precond=shesolve.PRECOND_BLOCK_ILU0, precision=shesolve.FP64), ranges),
                 ("fp32+ir bmc_ilu0", dict(precond=shesolve.PRECOND_BMC_ILU0, precision=shesolve.FP32_IR), None)]
        for name, kw, rg in cases:
            sv = shesolve.Solver(0, tol=1e-10, max_iters=5000, **kw)
            sv.set_matrix_device(A.n, A.nnz, d_rp, d_ci, d_val, block_ranges=rg)
            rep = sv.solve_device(d_rhs, dx.ptr)
            x = dx.download()
            res = np.linalg.norm(rhs - gsys @ x) / bn
            err = np.linalg.norm(x - xs_ref) / np.linalg.norm(xs_ref)
            n, _ = F.postproc(F.recover(x))
            dn = np.linalg.norm(n - n_ref) / np.linalg.norm(n_ref)
            _say(t0, f"{name} to 1e-10: {rep['iterations']} iterations, residual {res:.2e}, "
                     f"error {err:.2e}, density {dn:.2e}")
            assert rep["status"] == 0 and res <= 1e-10 and rep["rel_residual"] <= 1e-10, (rep, res)
            assert err <= 2e-7, err
            assert dn <= 1e-6, dn
            sv.set_params(tol=1e-12)
            rep = sv.solve_device(d_rhs, dx.ptr)
            x = dx.download()
            res = np.linalg.norm(rhs - gsys @ x) / bn
            err = np.linalg.norm(x - xs_ref) / np.linalg.norm(xs_ref)
            n, _ = F.postproc(F.recover(x))
            bigc = np.abs(n_ref) >= 1e-4 * np.max(np.abs(n_ref))
            dcell = np.max(np.abs(n - n_ref)[bigc] / np.abs(n_ref[bigc]))
            _say(t0, f"{name} to 1e-12: {rep['iterations']} iterations, residual {res:.2e}, "
                     f"error {err:.2e}, density per cell {dcell:.2e}")
            assert rep["status"] == 0 and res <= 1e-12, (rep, res)
            assert err <= 1e-8, err
            assert dcell <= 1e-6, dcell
            sv.close()
    finally:
        dx.close()
        F.close()
