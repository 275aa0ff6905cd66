"""Request loop for SHE solve sequences on one GPU, with continuous admission.

A request is one Gummel-like sequence of SHE solves on one device. It mirrors what
simulator::run does per nonlinear iteration (simulator.hpp:895-1096):
1. set up energies, expansion orders and mapping again (the pattern follows the potential);
2. assemble;
3. solve, warm-started from the previous iterate transferred to the new energy grid
   on the device (shefill_transfer_device, the reference's
   she::transfer_to_new_h_space);
4. perturb the potential before the next step (here: the request's potential plus
   N(0, sigma) per mesh vertex drawn afresh each step, each cell taking the mean of
   its corner vertices).

Serving is continuous:
- A request is admitted as soon as it has arrived, a slot is free and its estimated
  device footprint fits the HBM budget. Nothing waits for a batch to fill or drain.
- Each slot is a host thread that owns its own fill and solver handles, so its own HIP
  stream. The kernels of admitted requests therefore overlap on the device.
- The library calls release the GIL, so the slots' host work (pattern analysis, level
  schedules) also runs in parallel.
- The matrix, preconditioner, Krylov vectors and previous iterate of a request stay
  resident for its whole sequence.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import threading
import time

import numpy as np

from . import devices as D

_hip = None


def _hiplib():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip


class DeviceBuffer:
    """A plain device allocation of n doubles (freed on close)."""

    def __init__(self, n):
        self.n = n
        p = C.c_void_p()
        rc = _hiplib().hipMalloc(C.byref(p), C.c_size_t(8 * max(n, 1)))
        if rc != 0:
            raise MemoryError(f"hipMalloc of {8 * n} bytes failed ({rc})")
        self.ptr = p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        if a.size != self.n:
            raise ValueError(f"upload of {a.size} values into a buffer of {self.n}")
        rc = _hiplib().hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy host to device failed ({rc})")

    def download(self):
        a = np.empty(self.n)
        rc = _hiplib().hipMemcpy(a.ctypes.data, self.ptr, 8 * self.n, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy device to host failed ({rc})")
        return a

    def close(self):
        if self.ptr:
            _hiplib().hipFree(self.ptr)
            self.ptr = None


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the run
    kind: str                 # "diode" or "mosfet"
    dev: D.Device
    cfg: D.SheConfig
    seed: int
    steps: int = 5
    sigma: float = 5e-3       # potential perturbation per mesh vertex between steps [V]
    damping: float = 0.3      # Gummel damping of the density (solvers/config.hpp:257 default)


@dataclasses.dataclass
class Result:
    rid: int
    kind: str
    arrival: float
    start: float
    finish: float
    reports: list
    n: list
    x: np.ndarray             # last reduced solution
    density: np.ndarray       # damped carrier density after the last step
    footprint: float
    desc: str = ""
    update_norms: list = dataclasses.field(default_factory=list)  # relative density update per step

    @property
    def latency(self):
        return self.finish - self.arrival


def _even_count(L):
    """Even harmonics up to order L (odd L): 1, 6, 15, 28 for L = 1, 3, 5, 7."""
    return L * (L + 1) // 2


def footprint_estimate(dev: D.Device, cfg: D.SheConfig) -> float:
    """Upper estimate of the device bytes one request holds.

    The count covers:
    - the fill's reduced system (values, columns, row pointer);
    - the solver's FP64 and FP32 copies;
    - the level-of-fill factor;
    - 18 Krylov vectors;
    - per-row fill data.

    The energy count follows setup_energies (she/setup_energies.hpp:108-120)."""
    span = float(np.max(dev.potential) - np.min(dev.potential))
    nh = int((D.Q * span + 2 * cfg.energy_spacing + cfg.min_kin_range) / cfg.energy_spacing) + 1
    rows = dev.ncells * max(nh - 2, 1) * _even_count(cfg.L_max)
    per_row = {1: {1: 5.0, 3: 7.0, 5: 11.0, 7: 16.0}, 2: {1: 6.0, 3: 11.0, 5: 20.0, 7: 30.0}}[dev.dim][cfg.L_max]
    return rows * (per_row * 52.0 + 220.0)


def vertex_perturbation(dev: D.Device, rng, sigma):
    """Potential perturbation per cell from N(0, sigma) drawn at every mesh
    vertex: each cell takes the mean of its corner vertices (2 in 1D, 4 in 2D),
    the cell value of the piecewise (bi)linear vertex field."""
    if dev.dim == 1:
        v = rng.normal(0.0, sigma, dev.nx + 1)
        return 0.5 * (v[:-1] + v[1:])
    v = rng.normal(0.0, sigma, (dev.ny + 1, dev.nx + 1))
    return (0.25 * (v[:-1, :-1] + v[:-1, 1:] + v[1:, :-1] + v[1:, 1:])).ravel()


def _solver_params(dev, cfg):
    """2D meshes: block multicolour ILU0, FP32 + FP64 refinement (one sweep
    launch per colour, blocks of consecutive rows solved inside).  1D chains:
    ILU(k) of each per-energy slab (the block ranges of she::solve,
    she/linear_solver.hpp:42-83), k = 0, and k = 1 for L = 5 where ILU(0)
    diverges, in FP64.  The slabs are swept side by side, one wavefront each,
    so a sweep is as deep as one slab, not as the whole chain: on 1D diodes of
    100-1000 nm this solves 4.6x (L <= 3) and over 20x (L = 5) faster than a
    natural-order ILU of the whole system, whose level schedule is thousands
    of levels deep.  FP32 corrections of these systems need about twice the
    iterations, so FP64 is faster."""
    import shesolve
    if dev.dim > 1:
        return dict(precond=shesolve.PRECOND_BMC_ILU0, precision=shesolve.FP32_IR)
    return dict(precond=shesolve.PRECOND_BLOCK_ILUK, fill_level=1 if cfg.L_max >= 5 else 0,
                precision=shesolve.FP64)


def run_request(req: Request, tol=1e-10, max_iters=1000, device=0, clock=None, autotune=True):
    """One request, start to finish, on the calling thread.  autotune=False
    pins the launch shape, so that the result cannot depend on timings.
    max_iters: the reference's default iteration cap (solvers/config.hpp:62-63)."""
    import shesolve
    start = clock() if clock else time.perf_counter()
    F = shesolve.Fill(device)
    params = _solver_params(req.dev, req.cfg)
    slabs = params["precond"] == shesolve.PRECOND_BLOCK_ILUK
    sv = shesolve.Solver(device, tol=tol, max_iters=max_iters, autotune=1 if autotune else 0, **params)
    rng = np.random.default_rng(req.seed)
    base = np.array(req.dev.potential, np.float64)
    dev = dataclasses.replace(req.dev, potential=base.copy())
    reps, ns, norms = [], [], []
    pattern = None
    dx = dxf = dnew = dcur = dj = None
    try:
        for step in range(req.steps):
            if step > 0:
                # each step perturbs the request's potential afresh: successive
                # potentials differ by sqrt(2) sigma per vertex.  A random walk
                # (each step perturbing the last) grows as sqrt(step) sigma; on
                # the 154x77 L=3 device its 4th system is not solved to 1e-10 in
                # 3000 FP64 iterations by natural ILU(0), ILU(1) or block
                # multicolour ILU(0), while the fresh draw converges in 117-495
                dev = dataclasses.replace(dev, potential=base + vertex_perturbation(dev, rng, req.sigma))
            sz = F.build(dev, req.cfg)
            n, nnz = sz["n"], sz["nnz"]
            d_rp, d_ci, d_val, d_rhs = F.device_system()
            h = F.pattern_hash()
            if h == pattern and dx is not None and dx.n == n:
                sv.update_values_device(d_val)
            else:
                sv.set_matrix_device(n, nnz, d_rp, d_ci, d_val, block_ranges=F.block_ranges() if slabs else None)
                pattern = h
            if dx is None or dx.n != n:
                if dx is not None:
                    dx.close()
                dx = DeviceBuffer(n)
            warm = step > 0
            if warm:
                # the previous full solution and its density, still resident, onto
                # the new unknowns (transfer_to_new_h_space, on the device)
                F.transfer_device(dxf.ptr, dnew.ptr, dx.ptr)
            rep = sv.solve_device(d_rhs, dx.ptr, warm_start=warm)
            reps.append(rep)
            ns.append(n)
            # the tail of the Gummel step on the device: full solution, carrier
            # density and facet currents, then the damped density update and its
            # relative update norm (simulator.hpp:1081-1086)
            nfull = n + F.sizes["n_odd"]
            if dxf is None or dxf.n != nfull:
                if dxf is not None:
                    dxf.close()
                dxf = DeviceBuffer(nfull)
            if dnew is None:
                dnew, dcur = DeviceBuffer(F.sizes["cells"]), DeviceBuffer(F.sizes["cells"])
                dj = DeviceBuffer(max(F.sizes["facets"], 1))
            F.recover_device(dx.ptr, dxf.ptr)
            F.postproc_device(dxf.ptr, dnew.ptr, dj.ptr)
            if step == 0:
                dcur.upload(dnew.download())
            else:
                norms.append(F.density_update_device(dnew.ptr, dcur.ptr, req.damping))
        dens = dcur.download()
        x_last = dx.download()
    finally:
        for b in (dx, dxf, dnew, dcur, dj):
            if b is not None:
                b.close()
        sv.close()
        F.close()
    finish = clock() if clock else time.perf_counter()
    desc = (f"{req.kind} {req.dev.nx}x{req.dev.ny} L={req.cfg.L_max} "
            f"dE={req.cfg.energy_spacing / D.EV * 1e3:.2f} meV")
    return Result(req.rid, req.kind, req.arrival, start, finish, reps, ns, x_last, dens,
                  footprint_estimate(req.dev, req.cfg), desc, norms)


_mos_cache = {}


def _mosfet(vx, vy):
    key = (vx, vy)
    if key not in _mos_cache:
        _mos_cache[key] = D.mosfet_2d(vx=vx, vy=vy)
    return _mos_cache[key]


def make_requests(count, rate=200.0, seed=2024, mos_frac=0.2, mesh=(41, 161), steps=5, sigma=5e-3,
                  diode_length=(100, 1000)):
    """Poisson arrivals at `rate` per second.

    - A fraction `mos_frac` are 2D MOSFET-like devices: the 200x100 nm geometry, vx uniform in
      `mesh`, vy = (vx - 1) / 2 + 1, L in {1, 3}, 10 meV to 1.5 eV.
    - The rest are 1D diodes drawn like the batched workload: length uniform in
      `diode_length` nm with 1 nm cells, n+ 20 % of the length each side, L in
      {1, 3, 5}, energy step U[2, 10] meV to 1.0 eV, bias U[0, 1.5] V."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        if rng.random() < mos_frac:
            vx = int(rng.integers(mesh[0], mesh[1] + 1))
            vx += 1 - vx % 2
            vy = (vx - 1) // 2 + 1
            L = int(rng.choice([1, 3]))
            dev = _mosfet(vx, vy)
            cfg = D.SheConfig(L_max=L, energy_spacing=0.010 * D.EV, min_kin_range=1.5 * D.EV)
            kind = "mosfet"
        else:
            length = float(rng.integers(diode_length[0], diode_length[1] + 1))
            L = int(rng.choice([1, 3, 5]))
            de = float(rng.uniform(2.0, 10.0)) * 1e-3
            bias = float(rng.uniform(0.0, 1.5))
            dev = D.diode_1d(length_nm=length, vertices=int(length) + 1, bias=bias)
            cfg = D.SheConfig(L_max=L, energy_spacing=de * D.EV, min_kin_range=1.0 * D.EV)
            kind = "diode"
        out.append(Request(i, t, kind, dev, cfg, seed=int(rng.integers(1 << 30)), steps=steps, sigma=sigma))
    return out


class Server:
    """Slots (host threads) that admit arrived requests in arrival order while
    the summed footprint estimate of the requests in flight stays within the
    HBM budget."""

    def __init__(self, slots=8, hbm_budget=200e9, tol=1e-10, device=0, autotune=True, max_iters=5000):
        self.slots = slots
        self.budget = hbm_budget
        self.tol = tol
        self.device = device
        self.autotune = autotune
        self.max_iters = max_iters

    def run(self, requests, realtime=True, on_done=None):
        """Serves `requests`.  realtime: requests become eligible at their
        arrival time (else all at once).  on_done(result, finished, total) is
        called as each request completes.  Returns the results in request order."""
        pending = sorted(requests, key=lambda r: r.arrival)
        lock = threading.Condition()
        state = {"in_use": 0.0, "next": 0, "peak": 0.0, "inflight": 0, "max_inflight": 0, "done": 0}
        results = [None] * len(pending)
        errors = []
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0

        def worker():
            while True:
                with lock:
                    while True:
                        if errors:
                            return
                        i = state["next"]
                        if i >= len(pending):
                            return
                        req = pending[i]
                        fp = footprint_estimate(req.dev, req.cfg)
                        wait = req.arrival - clock() if realtime else 0.0
                        fits = state["in_use"] + fp <= self.budget or state["inflight"] == 0
                        if wait <= 0 and fits:
                            state["next"] += 1
                            state["in_use"] += fp
                            state["inflight"] += 1
                            state["peak"] = max(state["peak"], state["in_use"])
                            state["max_inflight"] = max(state["max_inflight"], state["inflight"])
                            break
                        lock.wait(timeout=max(min(wait, 0.05), 0.001) if wait > 0 else 0.05)
                try:
                    res = run_request(req, tol=self.tol, max_iters=self.max_iters, device=self.device, clock=clock,
                                      autotune=self.autotune)
                except Exception as e:  # noqa: BLE001 - surfaced to the caller below
                    with lock:
                        errors.append((req.rid, repr(e)))
                        lock.notify_all()
                    return
                with lock:
                    results[i] = res
                    state["in_use"] -= fp
                    state["inflight"] -= 1
                    state["done"] += 1
                    if on_done is not None:
                        on_done(res, state["done"], len(pending))
                    lock.notify_all()

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.slots)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise RuntimeError(f"requests failed: {errors}")
        self.stats = {"peak_footprint": state["peak"], "max_inflight": state["max_inflight"],
                      "makespan": clock()}
        by_id = {r.rid: r for r in results}
        return [by_id[r.rid] for r in requests]


def summarize(results, makespan):
    lat = np.array([r.latency for r in results])
    solves = sum(len(r.reports) for r in results)
    per_solve = []
    for r in results:
        per_solve.extend(rep["solve_ms"] for rep in r.reports)
    return {
        "requests": len(results),
        "solves": solves,
        "solves_per_s": solves / makespan if makespan > 0 else math.nan,
        "p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "solve_p50_ms": float(np.percentile(per_solve, 50)),
        "solve_p99_ms": float(np.percentile(per_solve, 99)),
        "max_rel_residual": max(rep["rel_residual"] for r in results for rep in r.reports),
        "all_converged": all(rep["status"] == 0 for r in results for rep in r.reports),
        "not_converged": sum(1 for r in results for rep in r.reports if rep["status"] != 0),
        "makespan_s": makespan,
    }
