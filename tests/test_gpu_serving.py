"""Serving loop (shesolve.serving): continuous admission of Gummel-like solve
sequences on one GPU.

Each request's result must be the same alone and in the mix:
- With the launch shape pinned (autotune off), the final iterate and carrier
  density are bit-identical: the requests share the device but no state.
- With the autotuner on (the serving default), the launch shapes and with
  them the rounding can differ, so the two runs are two solves to 1e-10 along
  different paths: the solutions agree within 1e-7 (2-norm relative; a
  condition number of ~1e3 times the 1e-10 residual bar) and the carrier
  densities within the 1e-6 bar.
- Every solve of every request reaches FP64 relative residual 1e-10.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests():
    from shesolve import serving
    # small members of the serving mix: three diodes and two MOSFET-like meshes
    reqs = serving.make_requests(5, rate=200.0, seed=11, mos_frac=0.4, mesh=(21, 29), steps=3,
                                 diode_length=(60, 120))
    kinds = sorted({r.kind for r in reqs})
    assert kinds == ["diode", "mosfet"], kinds
    return reqs


@pytest.mark.parametrize("autotune", [False, True])
def test_results_same_alone_and_in_the_mix(autotune):
    from shesolve import serving
    reqs = _requests()
    alone = serving.Server(slots=1, autotune=autotune).run(reqs, realtime=False)
    srv = serving.Server(slots=4, autotune=autotune)
    mix = srv.run(reqs, realtime=True)
    assert srv.stats["max_inflight"] >= 2
    for a, m in zip(alone, mix):
        assert a.rid == m.rid
        for rep in a.reports + m.reports:
            assert rep["status"] == 0 and rep["rel_residual"] <= 1e-10, (a.desc, rep)
        assert a.n == m.n
        if not autotune:
            assert np.array_equal(a.x, m.x) and np.array_equal(a.density, m.density)
        else:
            assert np.linalg.norm(a.x - m.x) <= 1e-7 * np.linalg.norm(a.x)
            assert np.linalg.norm(a.density - m.density) <= 1e-6 * np.linalg.norm(a.density)


def test_admission_respects_the_memory_budget():
    """With a budget of one request's footprint, requests run one at a time;
    a request larger than the whole budget still runs (alone)."""
    from shesolve import serving
    reqs = _requests()
    fps = [serving.footprint_estimate(r.dev, r.cfg) for r in reqs]
    srv = serving.Server(slots=4, hbm_budget=max(fps) * 1.01)
    res = srv.run(reqs, realtime=False)
    assert srv.stats["peak_footprint"] <= max(fps) * 1.01 + 1.0 or srv.stats["max_inflight"] == 1
    assert all(rep["status"] == 0 for r in res for rep in r.reports)
    srv = serving.Server(slots=4, hbm_budget=min(fps) * 0.5)
    res = srv.run(reqs[:2], realtime=False)
    assert srv.stats["max_inflight"] == 1


def _transfer_reference(s_old, s_new, xf_old, dens_old):
    """she::transfer_to_new_h_space restated for the even unknowns
    (she/transfer_to_new_h_space.hpp:58-196): per cell and new level, the old
    level of nearest kinetic energy (energy_index_H walked on the old grid,
    she/assemble_scattering.hpp:64-105), moved up past old levels of order 0,
    the shared harmonics copied; then every cell scaled to the old density and
    the result divided by the new system's unknown scaling and column factors."""
    ei0, eo0, _, _ = s_old.mapping()
    ei1, eo1, _, _ = s_new.mapping()
    H0, sh0 = s_old.energies()
    H1, sh1 = s_new.energies()
    nh0, nh1 = H0.size, H1.size
    f = np.zeros(s_new.n_even + s_new.n_odd)
    even = lambda L: L * (L + 1) // 2
    for c in range(s_new.ncells):
        k0 = lambda r: H0[r] - sh0[c]
        for h in range(1, nh1 - 1):
            u = ei1[c, h]
            if u < 0:
                continue
            E = H1[h] - sh1[c]
            r = min(h, nh0 - 1)
            if E != k0(r):
                inc = 1
                diff = diff0 = E - k0(r)
                if diff < 0:
                    inc = -1
                while nh0 - 1 > r > 0 and diff * diff0 > 0:
                    r += inc
                    diff = E - k0(r)
                if diff * diff0 > 0:
                    continue
                e1 = abs(diff)
                e2 = abs(E - k0(r - inc)) if 0 <= r - inc < nh0 else e1
                if e2 < e1:
                    r -= inc
            while eo0[c, r] == 0 and r < nh0 - 1:
                r += 1
            if eo0[c, r] <= 0 or ei0[c, r] < 0:
                continue
            m = min(even(eo1[c, h]), even(eo0[c, r]))
            f[u:u + m] = xf_old[ei0[c, r]:ei0[c, r] + m]
    dens_new, _ = s_new.postproc(f)
    for c in range(s_new.ncells):
        if dens_new[c] > 0 and dens_old[c] > 0:
            for h in range(nh1):
                u = ei1[c, h]
                if u >= 0:
                    f[u:u + even(eo1[c, h])] *= dens_old[c] / dens_new[c]
    _, _, scal, colf = s_new.reduced()
    return f[:s_new.n_even] / (scal * colf)


@pytest.mark.parametrize("kind", ["diode", "mosfet"])
def test_transfer_is_the_reference_rule(kind):
    """shefill_transfer_device against the restated transfer_to_new_h_space on a
    potential change large enough to move band edges across energy levels."""
    import dataclasses
    import scipy.sparse.linalg as spl
    import shesolve
    from oracle import oracle as O
    from shesolve import devices as D, serving
    if kind == "diode":
        dev0 = D.diode_1d(length_nm=100.0, vertices=101, bias=0.3)
        cfg = D.SheConfig(L_max=3, energy_spacing=0.01 * D.EV, min_kin_range=0.6 * D.EV)
    else:
        dev0 = D.mosfet_2d(vx=21, vy=11)
        cfg = D.SheConfig(L_max=1, energy_spacing=0.02 * D.EV, min_kin_range=0.6 * D.EV)
    rng = np.random.default_rng(5)
    dev1 = dataclasses.replace(dev0, potential=dev0.potential + serving.vertex_perturbation(dev0, rng, 0.03))
    s0, s1 = O.OracleSystem(dev0, cfg), O.OracleSystem(dev1, cfg)
    A0, b0, _, _ = s0.reduced()
    xf0, _ = s0.finish(spl.spsolve(A0.to_scipy().tocsc(), b0))
    dens0, _ = s0.postproc(xf0)
    expect = _transfer_reference(s0, s1, xf0, dens0)
    F = shesolve.Fill(0)
    F.build(dev0, cfg)
    F.build(dev1, cfg)
    bx, bd, bo = serving.DeviceBuffer(xf0.size), serving.DeviceBuffer(dens0.size), serving.DeviceBuffer(expect.size)
    try:
        bx.upload(xf0)
        bd.upload(dens0)
        F.transfer_device(bx.ptr, bd.ptr, bo.ptr)
        got = bo.download()
    finally:
        for b in (bx, bd, bo):
            b.close()
        F.close()
    assert np.array_equal(got == 0, expect == 0)
    assert np.linalg.norm(got - expect) <= 1e-13 * np.linalg.norm(expect)


def test_request_matches_the_oracle_gummel_sequence():
    """A serving request (5 warm-started solves, the potential perturbed per
    vertex between them, damped density updates) ends at the carrier density
    of the oracle's sequence: the same potentials, each system solved directly,
    the same damping (simulator.hpp:524-600); within 1e-6."""
    import dataclasses
    import scipy.sparse.linalg as spl
    from oracle import oracle as O
    from shesolve import devices as D, serving
    cases = [serving.Request(0, 0.0, "diode", D.diode_1d(length_nm=120.0, vertices=121, bias=0.4),
                             D.SheConfig(L_max=3, energy_spacing=0.008 * D.EV, min_kin_range=0.8 * D.EV), seed=17),
             serving.Request(1, 0.0, "mosfet", D.mosfet_2d(vx=25, vy=13),
                             D.SheConfig(L_max=1, energy_spacing=0.01 * D.EV, min_kin_range=1.0 * D.EV), seed=23)]
    for req in cases:
        res = serving.run_request(req, tol=1e-10, max_iters=5000, autotune=False)
        assert all(rep["status"] == 0 and rep["rel_residual"] <= 1e-10 for rep in res.reports), res.reports
        rng = np.random.default_rng(req.seed)
        base = np.array(req.dev.potential, np.float64)
        cur = None
        for step in range(req.steps):
            pot = base if step == 0 else base + serving.vertex_perturbation(req.dev, rng, req.sigma)
            s = O.OracleSystem(dataclasses.replace(req.dev, potential=pot), req.cfg)
            A, b, _, _ = s.reduced()
            xf, _ = s.finish(spl.spsolve(A.to_scipy().tocsc(), b))
            dens, _ = s.postproc(xf)
            cur = dens if step == 0 else O.density_update(dens, cur, req.damping)[0]
        err = np.linalg.norm(res.density - cur) / np.linalg.norm(cur)
        print(req.kind, "density error vs the oracle sequence", err, flush=True)
        assert err <= 1e-6, (req.kind, err)
