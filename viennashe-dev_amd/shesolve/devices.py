"""Synthetic structured devices for the SHE linear solve.

A device is the input of the coefficient fill: a structured 1D line or 2D
quadrilateral mesh (cell ids y-outer, x-inner as in the reference's
generator, util/generate_device.hpp:249-268 and :426-450), per-cell donor and
acceptor doping, the electrostatic potential and the cells that carry a
contact potential.

The potential is the built-in potential of the local doping
(physics/physics.hpp:141-163, the reference's initial guess
``built_in_potential_accessor``, accessors.hpp:160-200) plus a linear ramp of
the applied bias between the contacts -- a drift-diffusion-like initial guess
that needs no Poisson solve.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

Q = 1.602176565e-19
KB = 1.3806488e-23
EV = Q


@dataclasses.dataclass
class Device:
    dim: int
    nx: int
    ny: int
    xv: np.ndarray          # vertex x coordinates [m]
    yv: np.ndarray          # vertex y coordinates [m] (length 2 in 1D)
    doping_n: np.ndarray    # per cell [m^-3]
    doping_p: np.ndarray
    potential: np.ndarray   # per cell [V]
    contact: np.ndarray     # int32 per cell
    temperature: float = 300.0
    name: str = ""

    @property
    def ncells(self) -> int:
        return self.nx * self.ny


@dataclasses.dataclass
class SheConfig:
    L_max: int = 1
    energy_spacing: float = 0.031 * EV      # [J]
    min_kin_range: float = 1.0 * EV         # [J]
    max_kin_range: float = 5.0 * EV         # [J]
    tau: float = 1e-16                      # G/R boundary parameter [s]
    carrier: int = 0                        # 0 electrons, 1 holes
    acoustic: int = 1
    optical: int = 1
    scale: int = 1
    ordering: int = 0                       # 1: cells of a 1D device from both ends (the fill's numbering)


def _band_weight(T, ct):
    Ncv0 = 2.50933670314359713002e19 * 100 * 100 * 100
    me = 0.33 if ct == 0 else 0.6
    return Ncv0 * (me * (T / 300.0)) ** 1.5 * 6.0


def _aux(T, ct):
    edge = (1.0 if ct == 0 else -1.0) * 1.11 * Q * 0.5
    pol = -1.0 if ct == 0 else 1.0
    return _band_weight(T, ct) * math.exp(pol * edge / (KB * T))


def built_in_potential(T, dn, dp):
    """physics/physics.hpp:141-163, vectorised."""
    dn = np.asarray(dn, dtype=np.float64)
    dp = np.asarray(dp, dtype=np.float64)
    net = dn - dp
    naux, paux = _aux(T, 0), _aux(T, 1)
    x = np.sqrt(net * net + 4.0 * naux * paux)
    with np.errstate(divide="ignore", invalid="ignore"):
        pos = T * np.log((x + net) * 0.5 / naux) * KB / Q
        neg = -T * np.log((x - net) * 0.5 / paux) * KB / Q
    out = np.where(net >= 0.0, pos, neg)
    return np.where((dn <= 0) & (dp <= 0), 0.0, out)


EPS_SI = 11.9 * 1.0e7 / (4.0 * 3.1415926535897932384626433832795 * 299792458.0 * 299792458.0)


def poisson_potential(dev, quasi_fermi, max_newton=100, tol=1e-10):
    """Electrostatic potential of a structured device from the nonlinear
    Poisson equation with Boltzmann carriers at given quasi-Fermi levels:
        div(eps grad psi) = -q (p - n + N_D - N_A),
        n = n_aux exp((psi - phi)/V_T),  p = p_aux exp((phi - psi)/V_T),
    cell-centred finite volumes on the device's own mesh, Dirichlet
    psi = built-in potential + applied voltage on contact cells, zero flux
    elsewhere; damped Newton from the built-in guess.  This is the smooth
    drift-diffusion-like initial guess the SHE solve starts from."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    T = dev.temperature
    Vt = KB * T / Q
    naux, paux = _aux(T, 0), _aux(T, 1)
    nx, ny = dev.nx, dev.ny
    nc = nx * ny
    dx = np.diff(dev.xv)
    dy = np.diff(dev.yv) if dev.dim > 1 else np.array([1.0])
    xc = 0.5 * (dev.xv[:-1] + dev.xv[1:])
    yc = 0.5 * (dev.yv[:-1] + dev.yv[1:]) if dev.dim > 1 else np.array([0.0])
    vol = (dx[None, :] * dy[:, None]).ravel()
    rows, cols, vals = [], [], []
    # face couplings eps * area / distance
    i = np.arange(nc).reshape(ny, nx)
    if nx > 1:
        a = i[:, :-1].ravel(); b = i[:, 1:].ravel()
        area = np.repeat(dy, nx - 1) if dev.dim > 1 else np.ones(nx - 1)
        g = EPS_SI * area / np.tile(np.diff(xc), ny)
        rows += [a, b]; cols += [b, a]; vals += [g, g]
    if dev.dim > 1 and ny > 1:
        a = i[:-1, :].ravel(); b = i[1:, :].ravel()
        g = EPS_SI * np.tile(dx, ny - 1) / np.repeat(np.diff(yc), nx)
        rows += [a, b]; cols += [b, a]; vals += [g, g]
    rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    G = sp.csr_matrix((vals, (rows, cols)), shape=(nc, nc))
    Lap = G - sp.diags(np.asarray(G.sum(axis=1)).ravel())
    C = dev.doping_n - dev.doping_p
    phi = np.asarray(quasi_fermi, dtype=np.float64)
    contact = np.asarray(dev.contact, bool)
    psi = built_in_potential(T, dev.doping_n, dev.doping_p) + phi
    fixed = psi.copy()
    free = ~contact
    for _ in range(max_newton):
        n = naux * np.exp((psi - phi) / Vt)
        p = paux * np.exp((phi - psi) / Vt)
        F = Lap @ psi + Q * vol * (p - n + C)
        J = Lap + sp.diags(-Q * vol * (p + n) / Vt)
        Jf = J[free][:, free].tocsc()
        d = np.zeros(nc)
        d[free] = spl.spsolve(Jf, -F[free])
        step = np.max(np.abs(d))
        if step > 10 * Vt:
            d *= 10 * Vt / step
        psi = psi + d
        psi[contact] = fixed[contact]
        if step < tol:
            break
    return psi


def _vertices(start, length, points):
    j = np.arange(points, dtype=np.float64)
    return start + length * j / (float(points) - 1.0)


def diode_1d(length_nm=250.0, vertices=251, nplus_frac=0.2, n_plus=1e24, n_low=1e21,
             bias=0.5, T=300.0, poisson=False):
    """n+ - n - n+ diode on [0, length]; n+ regions are nplus_frac of the length
    at each end, the first and last cells are the contacts (0 V and ``bias``).
    poisson=True replaces the built-in guess by the nonlinear Poisson potential."""
    L = length_nm * 1e-9
    xv = _vertices(0.0, L, vertices)
    nx = vertices - 1
    xc = 0.5 * (xv[:-1] + xv[1:])
    dn = np.where((xc < nplus_frac * L) | (xc > (1.0 - nplus_frac) * L), n_plus, n_low)
    dp = np.zeros(nx)
    psi = built_in_potential(T, dn, dp) + bias * (xc - xc[0]) / (xc[-1] - xc[0])
    contact = np.zeros(nx, dtype=np.int32)
    contact[0] = contact[-1] = 1
    dev = Device(1, nx, 1, xv, np.array([0.0, 1.0]), dn, dp, psi, contact, T, name=f"diode1d_{vertices}v")
    if poisson:
        dev.potential = poisson_potential(dev, bias * (xc - xc[0]) / (xc[-1] - xc[0]))
    return dev


def mosfet_2d(width_nm=200.0, height_nm=100.0, vx=161, vy=81, pocket_w_nm=50.0,
              pocket_h_nm=20.0, n_plus=1e26, p_body=1e23, v_drain=0.5, T=300.0, poisson=True):
    """MOSFET-like rectangle: n+ source/drain pockets at the top corners of a
    p-body.  Contacts: top-row cells inside the pockets (source 0 V, drain
    ``v_drain``) and the bottom row (body, 0 V)."""
    W, Hh = width_nm * 1e-9, height_nm * 1e-9
    xv = _vertices(0.0, W, vx)
    yv = _vertices(0.0, Hh, vy)
    nx, ny = vx - 1, vy - 1
    xc = 0.5 * (xv[:-1] + xv[1:])
    yc = 0.5 * (yv[:-1] + yv[1:])
    X, Y = np.meshgrid(xc, yc)                     # [ny, nx]: cell id = j*nx + i
    pw, ph = pocket_w_nm * 1e-9, pocket_h_nm * 1e-9
    pocket = ((X < pw) | (X > W - pw)) & (Y > Hh - ph)
    dn = np.where(pocket, n_plus, 0.0).ravel()
    dp = np.where(pocket, 0.0, p_body).ravel()
    psi = built_in_potential(T, dn, dp) + v_drain * ((X - xc[0]) / (xc[-1] - xc[0])).ravel()
    contact = np.zeros((ny, nx), dtype=np.int32)
    contact[ny - 1, :] = np.where(pocket[ny - 1, :], 1, 0)
    contact[0, :] = 1
    dev = Device(2, nx, ny, xv, yv, dn, dp, psi, contact.ravel().astype(np.int32), T, name=f"mosfet2d_{vx}x{vy}")
    if poisson:
        dev.potential = poisson_potential(dev, v_drain * ((X - xc[0]) / (xc[-1] - xc[0])).ravel())
    return dev


def nin_2d(width_nm=100.0, height_nm=40.0, vx=41, vy=17, nplus_frac=0.2, n_plus=1e24, n_low=1e21,
           bias=0.3, T=300.0):
    """Two-dimensional n+ n n+ structure (a diode extruded in y); contacts are the
    left and right cell columns (0 V and ``bias``)."""
    W, Hh = width_nm * 1e-9, height_nm * 1e-9
    xv = _vertices(0.0, W, vx)
    yv = _vertices(0.0, Hh, vy)
    nx, ny = vx - 1, vy - 1
    xc = 0.5 * (xv[:-1] + xv[1:])
    yc = 0.5 * (yv[:-1] + yv[1:])
    X, Y = np.meshgrid(xc, yc)
    dn = np.where((X < nplus_frac * W) | (X > (1.0 - nplus_frac) * W), n_plus, n_low).ravel()
    dp = np.zeros(nx * ny)
    psi = built_in_potential(T, dn, dp) + bias * ((X - xc[0]) / (xc[-1] - xc[0])).ravel()
    contact = np.zeros((ny, nx), dtype=np.int32)
    contact[:, 0] = 1
    contact[:, nx - 1] = 1
    return Device(2, nx, ny, xv, yv, dn, dp, psi, contact.ravel().astype(np.int32), T, name=f"nin2d_{vx}x{vy}")


def nin_diode_250nm():
    """1D n+ n n+ diode, 251 vertices on [0,250] nm, L=1, 3.1 meV to 1.0 eV."""
    return diode_1d(), SheConfig(L_max=1, energy_spacing=0.0031 * EV, min_kin_range=1.0 * EV)


def mosfet_200x100nm(vx=161, vy=81, L=3):
    """2D MOSFET-like 200x100 nm, L=3, 10 meV to 1.5 eV."""
    return mosfet_2d(vx=vx, vy=vy), SheConfig(L_max=L, energy_spacing=0.010 * EV, min_kin_range=1.5 * EV)


def batch_diodes(count=96, seed=1234):
    """Seeded family of 1D diodes: length U[100,1000] nm with 1 nm cells, n+ 20%
    each side, L in {1,3,5}, energy step U[2,10] meV to 1.0 eV, bias U[0,1.5] V."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        length = float(rng.integers(100, 1001))
        L = int(rng.choice([1, 3, 5]))
        de = float(rng.uniform(2.0, 10.0)) * 1e-3
        bias = float(rng.uniform(0.0, 1.5))
        dev = diode_1d(length_nm=length, vertices=int(length) + 1, bias=bias)
        out.append((dev, SheConfig(L_max=L, energy_spacing=de * EV, min_kin_range=1.0 * EV)))
    return out
