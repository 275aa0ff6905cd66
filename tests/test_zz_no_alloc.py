"""The allocation-free Gummel density update (shefill_density_update_device_noalloc)
against the allocating entry point it mirrors (shefill_density_update_device):
the updated density and the update norm must be the same bit for bit, on two
calls in a row through the C ABI, the second on a smaller device (fewer cells,
so fewer partial sums than the first call left in the shared scratch)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def _upload(hip, a):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 8))) == 0
    assert hip.hipMemcpy(p, a.ctypes.data, ctypes.c_size_t(a.nbytes), 1) == 0  # host -> device
    return p


def _download(hip, p, n):
    out = np.zeros(n)
    assert hip.hipMemcpy(out.ctypes.data, p, ctypes.c_size_t(8 * n), 2) == 0  # device -> host
    return out


@pytest.mark.parametrize("logarithmic,alpha", [(False, 0.3), (True, 0.5)])
def test_noalloc_density_update_matches_the_allocating_one(logarithmic, alpha):
    import shesolve
    from shesolve import devices as D
    hip = _hip()
    F = shesolve.Fill(0)
    rng = np.random.default_rng(7)
    # first a 2D device (12 800 cells, 50 partial sums), then a smaller 1D one (99 cells, 1)
    cases = [(D.mosfet_2d(vx=161, vy=81, poisson=False), D.SheConfig(L_max=1, energy_spacing=0.05 * D.EV,
                                                                      min_kin_range=0.3 * D.EV)),
             (D.diode_1d(length_nm=100.0, vertices=100, bias=0.2), D.SheConfig(L_max=1, energy_spacing=0.02 * D.EV,
                                                                             min_kin_range=0.3 * D.EV))]
    for dev, cfg in cases:
        F.build(dev, cfg)
        nc = dev.ncells
        print(f"[no-alloc] {dev.name}: {nc} cells, logarithmic={logarithmic}, alpha={alpha}", flush=True)
        cur = 10.0 ** rng.uniform(10, 24, nc)
        new = cur * rng.uniform(0.5, 2.0, nc)
        new[::17] = 0.0
        d_new = _upload(hip, new)
        d_a, d_b = _upload(hip, cur), _upload(hip, cur)
        try:
            n_a = F.density_update_device(d_new.value, d_a.value, alpha, logarithmic)
            n_b = F.density_update_device_noalloc(d_new.value, d_b.value, alpha, logarithmic)
            a, b = _download(hip, d_a, nc), _download(hip, d_b, nc)
        finally:
            for p in (d_new, d_a, d_b):
                hip.hipFree(p)
        assert np.array_equal(a, b)
        assert n_a == n_b, (n_a, n_b)
        assert n_a > 0
    F.close()
