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
