"""Pins of the CPU oracle (no GPU).

The oracle restates the reference's SHE assembly, elimination, normalisation
and ILU/BiCGStab solve.  Each check below is independent of the oracle's own
code: golden data held by the reference (coupling tables, recorded dispersion
samples), closed forms, textbook identities (ILU0 reproduces A on its
pattern, exact LU on a tridiagonal matrix), direct solves, and the physical
invariants the reference's own tests check (Maxwellian equilibrium,
tests/src/equilibrium_resistor.cpp:121-197; current conservation,
tests/src/ushape_2d.cpp:196-204).
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl
from scipy.special import lpmv

from conftest import oracle_system

Q = 1.602176565e-19
KB = 1.3806488e-23


# ------------------------------------------------------------------ coupling tables
def _Y(n, m, th, ph):
    """Real spherical harmonic as math/spherical_harmonics.hpp:98-120 defines it."""
    am = abs(m)
    norm = math.sqrt((2 * n + 1) * math.factorial(n - am) / (2 * math.pi * math.factorial(n + am)))
    P = lpmv(am, n, np.cos(th))
    if m > 0:
        return norm * P * np.cos(m * ph)
    if m == 0:
        return norm * P / math.sqrt(2.0)
    return norm * P * np.sin(m * ph)


def _quadrature_tables(L):
    """a and b coupling integrals (harmonics_coupling.hpp:35-205) by Gauss-Legendre
    in theta and the trapezoid rule in phi (exact for the trigonometric part)."""
    xg, wg = np.polynomial.legendre.leggauss(64)
    th = 0.5 * np.pi * (xg + 1.0)
    wt = 0.5 * np.pi * wg
    nph = 64
    ph = 2 * np.pi * np.arange(nph) / nph
    wp = 2 * np.pi / nph
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    W = wt[:, None] * wp
    K = (L + 1) ** 2
    idx = [(l, m) for l in range(L + 1) for m in range(-l, l + 1)]
    Y = {(l, m): _Y(l, m, TH, PH) for (l, m) in idx}
    h = 1e-6
    dth = {(l, m): (_Y(l, m, TH + h, PH) - _Y(l, m, TH - h, PH)) / (2 * h) for (l, m) in idx}
    dph = {(l, m): (m * _Y(l, -m, TH, PH) if m != 0 else 0.0 * TH) for (l, m) in idx}
    S, C_, St, Ct = np.sin(TH), np.cos(TH), np.sin(PH), np.cos(PH)
    out = {k: np.zeros((K, K)) for k in ("a_x", "a_y", "a_z", "b_x", "b_y", "b_z")}
    for (l, m) in idx:
        i = l * l + l + m
        for (l1, m1) in idx:
            if abs(l - l1) != 1:
                continue
            j = l1 * l1 + l1 + m1
            Y1 = Y[(l1, m1)]
            out["a_x"][i, j] = np.sum(W * Y[(l, m)] * S * Ct * Y1 * S)
            out["a_y"][i, j] = np.sum(W * Y[(l, m)] * S * St * Y1 * S)
            out["a_z"][i, j] = np.sum(W * Y[(l, m)] * C_ * Y1 * S)
            out["b_x"][i, j] = np.sum(W * (dth[(l, m)] * C_ * Ct * S - dph[(l, m)] * St) * Y1)
            out["b_y"][i, j] = np.sum(W * (dth[(l, m)] * C_ * St * S + dph[(l, m)] * Ct) * Y1)
            out["b_z"][i, j] = np.sum(W * (dth[(l, m)] * (-1.0) * S) * Y1 * S)
    return out


@pytest.mark.parametrize("L", [1, 3])
def test_coupling_fixture_is_the_coupling_integral(L):
    from oracle import oracle as O
    tab = O.coupling_tables()[str(L)]
    ref = _quadrature_tables(L)
    K = (L + 1) ** 2
    for name, entries in tab.items():
        M = np.zeros((K, K))
        for i, j, v in entries:
            assert M[i, j] == 0.0, f"{name}({i},{j}) assigned twice"
            M[i, j] = v
        Q_ = ref[name]
        # every stored constant is the integral (the tables carry ~1e-9 quadrature error)
        assert np.max(np.abs(M - np.where(M != 0, Q_, 0.0))) <= 2e-7, name
        # and every integral above 1e-8 is stored, except the l=0 row of b whose
        # derivative vanishes (b(0, l=1) = 0) -- nothing is missing or transposed
        missing = (np.abs(Q_) > 1e-6) & (M == 0.0)
        assert not missing.any(), (name, np.argwhere(missing)[:5])


def test_dispersion_known_values():
    """Modena band at 0.1 eV: density of states and velocity against the closed
    forms of physics/dispersion.hpp:233-270 (eps (1 + alpha eps) = hbar^2 k^2 / 2 m_d);
    their magnitudes in the fill are fixed by tests/test_physics_pins.py."""
    from oracle import oracle as O
    Z, v = O.dispersion(0, 0.1 * Q)
    hbar = 6.62606896e-34 / (2 * math.pi)
    m0 = 9.10938291e-31
    md, ml, mt = 0.32 * m0, 0.98 * m0, 0.19 * m0
    mc = 3.0 / (1 / ml + 2 / mt)
    a, e = 0.5 / Q, 0.1 * Q
    Zc = md * math.sqrt(2 * md) / (2 * math.pi ** 2 * hbar ** 3) * (1 + 2 * a * e) * math.sqrt(e * (1 + a * e))
    vc = math.sqrt(mc / md) * math.sqrt(2 * e * (1 + a * e) / md) / (1 + 2 * a * e)
    assert abs(Z / Zc - 1) < 1e-13 and abs(v / vc - 1) < 1e-13
    assert O.dispersion(0, -1e-21) == (0.0, 0.0)


# ------------------------------------------------------------------ solvers
def _tridiag(n, seed=0):
    rng = np.random.default_rng(seed)
    d = 4.0 + rng.random(n)
    lo, up = -1.0 - rng.random(n - 1), -1.0 - rng.random(n - 1)
    return sp.diags([lo, d, up], [-1, 0, 1]).tocsr()


def _csr(M):
    from oracle import oracle as O
    M = sp.csr_matrix(M)
    M.sort_indices()
    return O.CSR(M.indptr, M.indices, M.data)


def test_ilu0_is_exact_lu_on_tridiagonal():
    from oracle import oracle as O
    M = _tridiag(200)
    A = _csr(M)
    b = np.arange(200.0) + 1
    P = O.Precond(O.PC_ILU0, A)
    x, it, res = O.bicgstab(A, b, precond=P, tol=1e-13)
    assert it == 1 and res <= 1e-13
    assert np.allclose(x, spl.spsolve(M.tocsc(), b), rtol=1e-12)


def test_block_ilu0_exact_on_block_diagonal_and_ilut_full_fill_is_lu():
    from oracle import oracle as O
    M = sp.block_diag([_tridiag(50, 1), _tridiag(70, 2)]).tocsr()
    A = _csr(M)
    b = np.ones(120)
    P = O.Precond(O.PC_BLOCK_ILU0, A, ranges=[[0, 50], [50, 120]])
    x, it, res = O.bicgstab(A, b, precond=P, tol=1e-13)
    assert it == 1
    # a dense-ish random matrix: ILUT without dropping equals LU
    rng = np.random.default_rng(3)
    D = sp.random(60, 60, density=0.2, random_state=3) + sp.eye(60) * 8
    A2 = _csr(D)
    P2 = O.Precond(O.PC_ILUT, A2, drop=0.0, fill=60)
    x, it, res = O.bicgstab(A2, rng.random(60), precond=P2, tol=1e-12)
    assert it == 1 and res <= 1e-12


def _laplace5(nx, ny):
    T = sp.diags([-1.0, 4.0, -1.0], [-1, 0, 1], shape=(nx, nx))
    S = sp.diags([-1.0, -1.0], [-1, 1], shape=(ny, ny))
    return (sp.kron(sp.eye(ny), T) + sp.kron(S, sp.eye(nx))).tocsr()


def test_iluk_level_of_fill_on_the_five_point_laplacian():
    """Level-of-fill pattern (Saad, Iterative Methods, 2nd ed., sec. 10.3.3):
    ILU(0) keeps the five diagonals; ILU(1) adds exactly the two fill diagonals
    at offsets -(nx-1) and +(nx-1) (eliminating the south neighbour brings in its
    east neighbour, eliminating the west neighbour brings in its north
    neighbour); ILU(k) with k >= nx reproduces the exact LU of this banded
    matrix, so M = A."""
    from oracle import oracle as O
    nx, ny = 7, 6
    M = _laplace5(nx, ny)
    A = _csr(M)
    n = nx * ny
    F0 = O.Precond(O.PC_ILUK, A, drop=0.0, fill=0).factors()
    assert np.array_equal(F0.row_ptr, A.row_ptr) and np.array_equal(F0.col, A.col)
    F1 = O.Precond(O.PC_ILUK, A, drop=0.0, fill=1).factors()
    for i in range(n):
        gx, gy = i % nx, i // nx
        want = {i}
        if gx > 0: want.add(i - 1)
        if gx < nx - 1: want.add(i + 1)
        if gy > 0: want.add(i - nx)
        if gy < ny - 1: want.add(i + nx)
        if gy > 0 and gx < nx - 1: want.add(i - nx + 1)
        if gy < ny - 1 and gx > 0: want.add(i + nx - 1)
        got = set(F1.col[F1.row_ptr[i]:F1.row_ptr[i + 1]].tolist())
        assert got == want, (i, sorted(got ^ want))
    P = O.Precond(O.PC_ILUK, A, drop=0.0, fill=2 * nx)
    x = np.random.default_rng(0).random(n)
    assert np.allclose(P.apply(M @ x), x, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", ["diode_L1", "diode_L5"])
def test_iluk_zero_level_is_ilu0_and_dropping_rules(name):
    """ILU(k) at k = 0 is ILU(0) entry for entry; ILU(k) keeps (L U)_ij = a_ij on
    A's pattern (its defining property on the extended pattern); with a drop
    tolerance every stored multiplier and upper entry is either zero or at
    least the row's threshold, and a drop tolerance of 1e-300 drops nothing."""
    from oracle import oracle as O
    A, b, _, _ = oracle_system(name).reduced()
    F0 = O.Precond(O.PC_ILU0, A).factors()
    Fk0 = O.Precond(O.PC_ILUK, A, drop=0.0, fill=0).factors()
    assert np.array_equal(F0.row_ptr, Fk0.row_ptr) and np.array_equal(F0.col, Fk0.col)
    assert np.array_equal(F0.val, Fk0.val)
    P1 = O.Precond(O.PC_ILUK, A, drop=0.0, fill=1)
    F1 = P1.factors()
    Fs = F1.to_scipy().tocsr()
    L = sp.tril(Fs, -1) + sp.eye(A.n)
    U = sp.triu(Fs)
    LU = (L @ U).tocsr()
    Mx = A.to_scipy().tocsr()
    vals = np.asarray(LU[Mx.nonzero()]).ravel()
    assert np.max(np.abs(vals - Mx.data)) <= 1e-11 * max(np.abs(L).max() * np.abs(U).max(), 1.0)
    Ft = O.Precond(O.PC_ILUK, A, drop=1e-300, fill=1).factors()
    assert np.array_equal(Ft.val, F1.val)
    drop = 1e-3
    Fd = O.Precond(O.PC_ILUK, A, drop=drop, fill=2).factors()
    rn = np.sqrt(np.add.reduceat(A.val ** 2, A.row_ptr[:-1]))
    rows = np.repeat(np.arange(A.n), np.diff(Fd.row_ptr))
    off = Fd.col != rows
    v = np.abs(Fd.val[off])
    assert np.all((v == 0) | (v >= drop * rn[rows[off]]))
    assert np.count_nonzero(Fd.val[off] == 0) > 0


@pytest.mark.parametrize("name", ["diode_L1", "diode_L3", "mos2d_L3"])
def test_ilu0_reproduces_A_on_its_pattern(name):
    """ILU(0) defining property: (L U)_ij = a_ij for every (i, j) in the pattern of A."""
    from oracle import oracle as O
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    F = O.Precond(O.PC_ILU0, A).factors().to_scipy().tocsr()
    assert np.array_equal(F.indptr, A.row_ptr) and np.array_equal(F.indices, A.col)
    L = sp.tril(F, -1) + sp.eye(A.n)
    U = sp.triu(F)
    LU = (L @ U).tocsr()
    M = A.to_scipy().tocsr()
    vals = np.asarray(LU[M.nonzero()]).ravel()
    ref = M.data
    scale = np.abs(L).max() * np.abs(U).max()
    assert np.max(np.abs(vals - ref)) <= 1e-12 * max(scale, 1.0)


@pytest.mark.parametrize("name", ["diode_L1", "diode_L3", "mos2d_L3"])
def test_bicgstab_matches_direct_solve(name):
    from oracle import oracle as O
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    x, it, res = O.bicgstab(A, b, precond=O.Precond(O.PC_ILU0, A), tol=1e-13, max_iters=5000)
    xd = spl.spsolve(A.to_scipy().tocsc(), b)
    assert res <= 1e-13
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)


# ------------------------------------------------------------------ front end
@pytest.mark.parametrize("name", ["diode_L1", "diode_L3", "mos2d_L3"])
def test_elimination_and_recovery_solve_the_full_system(name):
    """recover(eliminate(A)) gives the solution of the full even+odd system, and
    the odd-odd block is diagonal (she/eliminate.hpp:80-354)."""
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    xd = spl.spsolve(A.to_scipy().tocsc(), b)
    xf, res = s.finish(xd)
    assert res <= 1e-10
    Fm, fb = s.full()
    M = Fm.to_scipy().tocsr()
    ne = s.n_even
    oo = M[ne:, ne:].tocoo()
    assert np.all(oo.row == oo.col) and np.all(oo.data > 0)
    assert np.linalg.norm(M @ xf - fb) <= 1e-10 * np.linalg.norm(fb)


@pytest.mark.parametrize("name", ["diode_L1", "mos2d_L3"])
def test_normalisation_identities(name):
    """row_normalize_system (math/linalg_util.hpp:200-282): A = R * A0 * S * C with
    S the unknown scaling exp(-E_kin/kT), C the column factors (geometric mean of
    each column set to one), R the inverse row 2-norms; rows end with unit norm
    and positive diagonal."""
    s = oracle_system(name)
    A, b, sc, cf = s.reduced()
    A0, b0 = s.unscaled()
    M, M0 = A.to_scipy().tocsr(), A0.to_scipy().tocsr()
    rows = np.sqrt(np.asarray(M.multiply(M).sum(axis=1)).ravel())
    assert np.allclose(rows, 1.0, rtol=1e-12)
    assert np.all(M.diagonal() > 0)
    Cs = M0 @ sp.diags(sc * cf)
    rn = np.sqrt(np.asarray(Cs.multiply(Cs).sum(axis=1)).ravel())
    R = sp.diags(1.0 / rn)
    assert abs(R @ Cs - M).max() <= 1e-12
    assert np.allclose(b, b0 / rn, rtol=1e-12, atol=0)
    # geometric mean of |column| after unknown scaling and column factor = 1
    Sc = (M0 @ sp.diags(sc * cf)).tocsc()
    for j in np.random.default_rng(0).choice(A.n, 200, replace=False):
        col = Sc.data[Sc.indptr[j]:Sc.indptr[j + 1]]
        assert abs(np.mean(np.log(np.abs(col)))) <= 1e-9


@pytest.mark.parametrize("name", ["diode_L1", "diode_L3", "nin2d_L3", "mos2d_L3"])
def test_mapping_rules(name):
    """Energy-major numbering (mapping.hpp:204-271): H outer, cells inner, no
    unknowns at H = 0 and N_H - 1, none where the averaged density of states is
    zero; even_unknowns_on_node(L) per node; orders L_max or 1 (band edge)."""
    s = oracle_system(name)
    ei, eo, oi, oo = s.mapping()
    H, sh = s.energies()
    kin = H[None, :] - sh[:, None]
    assert np.all(ei[:, 0] < 0) and np.all(ei[:, -1] < 0)
    nodes = [(h, c) for h in range(s.NH) for c in range(s.ncells) if ei[c, h] >= 0]
    expect = 0
    for h, c in nodes:
        assert ei[c, h] == expect
        L = eo[c, h]
        assert L in (1, s.cfg.L_max) or (s.cfg.L_max >= 5 and L in (3,))
        expect += sum(2 * l + 1 for l in range(0, L + 1, 2))
        # some point of the energy box is above the band edge
        assert max(kin[c, h], (kin[c, h] + kin[c, h + 1]) / 2) > 0
    assert expect == s.n_even
    # cells strictly inside the band (all three box points below zero) carry nothing
    inside_gap = (kin[:, 1:-1] <= 0) & (kin[:, 2:] + kin[:, 1:-1] <= 0) & (kin[:, :-2] + kin[:, 1:-1] <= 0)
    assert np.all(ei[:, 1:-1][inside_gap] < 0)


@pytest.mark.parametrize("name", ["diode_L1", "mos2d_L3"])
def test_pattern_couples_only_facet_neighbours_and_phonon_partners(name):
    """Reduced rows couple a node to itself, its facet neighbours at the same H,
    and (harmonic 0 only) to the same cell at H +- hbar*omega/dE."""
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    ei, eo, _, _ = s.mapping()
    n = A.n
    node_c = np.zeros(n, np.int64)
    node_h = np.zeros(n, np.int64)
    harm = np.zeros(n, np.int64)
    for c in range(s.ncells):
        for h in range(s.NH):
            u = ei[c, h]
            if u >= 0:
                e = sum(2 * l + 1 for l in range(0, eo[c, h] + 1, 2))
                node_c[u:u + e] = c
                node_h[u:u + e] = h
                harm[u:u + e] = np.arange(e)
    k = round(0.062 / (s.cfg.energy_spacing / Q))
    M = A.to_scipy().tocoo()
    r, c = M.row, M.col
    same_h = node_h[r] == node_h[c]
    nx = s.dev.nx
    dc = np.abs(node_c[r] - node_c[c])
    neighbour = (dc == 0) | (dc == 1) | (dc == nx)
    assert np.all(neighbour[same_h])
    other = ~same_h
    assert np.all(node_c[r][other] == node_c[c][other])
    assert np.all(np.abs(node_h[r] - node_h[c])[other] == k)
    assert np.all((harm[r] == 0)[other] & (harm[c] == 0)[other])


# ------------------------------------------------------------------ physics invariants
def _resistor(dim, T, dE):
    from shesolve import devices as D
    if dim == 1:
        dev = D.diode_1d(length_nm=60.0, vertices=31, n_plus=1e23, n_low=1e23, bias=0.0, T=T)
    else:
        dev = D.nin_2d(width_nm=100.0, height_nm=20.0, vx=11, vy=3, n_plus=1e23, n_low=1e23, bias=0.0, T=T)
    cfg = D.SheConfig(L_max=1, energy_spacing=dE * Q, min_kin_range=0.5 * Q)
    return dev, cfg


@pytest.mark.parametrize("dim,T", [(1, 300.0), (2, 300.0), (2, 400.0)])
def test_equilibrium_is_maxwellian(dim, T):
    """Zero bias, uniform doping: f_00(H) proportional to exp(-E_kin/kT) to 1e-3 and
    the odd (current-carrying) coefficients vanish (equilibrium_resistor.cpp:121-197)."""
    import scipy.sparse.linalg as spl2
    from oracle import oracle as O
    dev, cfg = _resistor(dim, T, 0.0062)
    s = O.OracleSystem(dev, cfg)
    A, b, _, _ = s.reduced()
    x = spl2.spsolve(A.to_scipy().tocsc(), b)
    xf, res = s.finish(x)
    ei, eo, oi, oo = s.mapping()
    H, sh = s.energies()
    ratios = []
    for c in range(s.ncells):
        for h in range(s.NH):
            if ei[c, h] >= 0:
                ratios.append(xf[ei[c, h]] / math.exp(-(H[h] - sh[c]) / (KB * T)))
    ratios = np.array(ratios)
    assert np.max(np.abs(ratios / np.median(ratios) - 1)) <= 1e-3
    odd = xf[s.n_even:]
    assert np.max(np.abs(odd)) <= 1e-12 * np.max(np.abs(xf[:s.n_even]))


@pytest.mark.parametrize("name", ["diode_L1", "diode_L3", "nin2d_L3"])
def test_current_conservation(name):
    """Terminal currents of the two contacts cancel to 1e-5 (ushape_2d.cpp:196-204)
    and the contact density follows the Maxwellian boundary normalisation."""
    from test_gpu_solver import terminal_currents
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    x = spl.spsolve(A.to_scipy().tocsc(), b)
    xf, _ = s.finish(x)
    n, j = s.postproc(xf)
    I = terminal_currents(s, j)
    assert abs(I[0] + I[1]) <= 1e-5 * np.max(np.abs(I))
    # majority density in the contact cells is the doping (G/R contacts, tau = 1e-16 s)
    ct = np.nonzero(s.dev.contact)[0]
    assert np.allclose(n[ct], s.dev.doping_n[ct], rtol=2e-2)


def test_density_update_closed_forms():
    """The damped Gummel update: a = 1 takes the new density (norm 0), a = 1/2
    gives the arithmetic (linear) or geometric (logarithmic) mean, the
    logarithmic form limits a collapsing density to 1e-10 of the old one, and
    the norm is the 2-norm of (new - upd) / upd with the new value where the
    update is 0: (new - old) / (new + old) per cell for the linear a = 1/2."""
    from oracle import oracle as O
    cur = np.array([1e10, 4.0, 2.5e16, 0.0, 3.0])
    new = np.array([4e10, 9.0, 2.5e16, 7.0, -1.0])
    out, nrm = O.density_update(new, cur, 1.0)
    assert np.array_equal(out, new) and nrm == 0.0
    out, nrm = O.density_update(new, cur, 0.5)
    assert np.allclose(out, (new + cur) / 2, rtol=1e-15)
    assert nrm == pytest.approx(np.sqrt(np.sum(((new - cur) / (new + cur)) ** 2)), rel=1e-14)
    assert nrm == pytest.approx(np.sqrt(0.6 ** 2 + (5 / 13) ** 2 + 0 + 1 + 2 ** 2), rel=1e-14)
    pos = slice(0, 3)
    out, _ = O.density_update(new[pos], cur[pos], 0.5, logarithmic=True)
    assert np.allclose(out, np.sqrt(new[pos] * cur[pos]), rtol=1e-15)
    out, _ = O.density_update(np.array([-5.0]), np.array([2.0]), 0.25, logarithmic=True)
    assert np.isclose(out[0], 2.0 ** 0.75 * 2e-10 ** 0.25, rtol=1e-15)
    out, _ = O.density_update(np.array([1e-20]), np.array([2.0]), 1.0, logarithmic=True)
    assert out[0] == 2e-10
    out, nrm = O.density_update(new, cur, 0.3)
    brute = 0.0
    for n_, u_ in zip(new, out):
        brute += ((n_ - u_) / u_) ** 2 if u_ != 0 else n_ ** 2
    assert nrm == pytest.approx(brute ** 0.5, rel=1e-15)
    out, nrm = O.density_update(np.array([0.0, 3.0]), np.array([0.0, 0.0]), 0.3)
    assert np.allclose(out, [0.0, 0.9], rtol=1e-15) and nrm == pytest.approx(2.1 / 0.9, rel=1e-14)


def _greedy_block_colours(A, block):
    """Greedy colouring of blocks of `block` consecutive rows in natural order
    (smallest free colour; two blocks conflict when an entry of the symmetrised
    pattern links them); the colour of each row."""
    n = A.n
    P = sp.csr_matrix((np.ones(A.nnz), A.col, A.row_ptr), shape=(n, n))
    S = (P + P.T).tocsr()
    nb = (n + block - 1) // block
    R = sp.csr_matrix((np.ones(n), (np.arange(n) // block, np.arange(n))), shape=(nb, n))
    SB = (R @ S @ R.T).tocsr()
    col = np.full(nb, -1)
    for i in range(nb):
        nbrs = SB.indices[SB.indptr[i]:SB.indptr[i + 1]]
        used = set(col[nbrs[nbrs != i]].tolist())
        q = 0
        while q in used:
            q += 1
        col[i] = q
    return col[np.arange(n) // block]


def _bicgstab_iterations_in_order(A, b, perm):
    from oracle import oracle as O
    M = sp.csr_matrix((A.val, A.col, A.row_ptr), shape=(A.n, A.n))[perm][:, perm].tocsr()
    M.sort_indices()
    Ap = O.CSR(M.indptr.astype(np.int64), M.indices.astype(np.int32), M.data.copy())
    x, it, res = O.bicgstab(Ap, b[perm], precond=O.Precond(O.PC_ILU0, Ap), tol=1e-10, max_iters=3000)
    assert res <= 1e-10, res
    return it


@pytest.mark.parametrize("name", ["mos2d_L3", "nin2d_L3"])
def test_block_multicolour_ordering(name):
    """The block multicolour ordering the GPU solver uses for 2D devices (blocks
    of 32 consecutive rows coloured greedily, rows sorted by colour): blocks of
    one colour are independent, and ILU(0)-preconditioned BiCGStab in that order
    converges in no more iterations than in the point multicolour order (block
    size 1).  A colouring that let two blocks of one colour touch fails the
    first check."""
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    colour = _greedy_block_colours(A, 32)
    rows = np.repeat(np.arange(A.n), np.diff(A.row_ptr))
    same = colour[rows] == colour[A.col]
    assert np.all((rows // 32 == A.col // 32)[same])
    bad = colour.copy()
    bad[:32] = colour[32]
    same_bad = bad[rows] == bad[A.col]
    assert not np.all((rows // 32 == A.col // 32)[same_bad])
    it_block = _bicgstab_iterations_in_order(A, b, np.argsort(colour, kind="stable"))
    it_point = _bicgstab_iterations_in_order(A, b, np.argsort(_greedy_block_colours(A, 1), kind="stable"))
    print(name, "colours", colour.max() + 1, "iterations block", it_block, "point", it_point)
    assert it_block <= it_point


@pytest.mark.parametrize("name", ["mos2d_L3", "nin2d_L3", "diode_L3"])
def test_parallel_baseline_matches_the_sequential_solver(name):
    """The parallel CPU baseline (oracle_bicgstab_parallel) runs the same
    BiCGStab with per-energy block ILU0 on all OpenMP threads.  Its block
    application is the sequential one restricted to each block (no factor entry
    leaves a block), so one iteration gives the sequential iterate up to the
    summation order of the dot products, and the solve reaches the same
    solution."""
    from oracle import oracle as O
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    ranges = s.block_ranges()
    P = O.Precond(O.PC_BLOCK_ILU0, A, ranges=ranges)
    # every nonempty energy slab is one parallel block
    assert P.parallel_blocks() == int(np.count_nonzero(ranges[:, 1] > ranges[:, 0])) > 1
    # the blocks factorised on all threads give the same factor bit for bit
    Pp = O.Precond(O.PC_BLOCK_ILU0, A, ranges=ranges, parallel=True)
    F, Fp = P.factors(), Pp.factors()
    assert np.array_equal(F.row_ptr, Fp.row_ptr) and np.array_equal(F.col, Fp.col)
    assert np.array_equal(F.val, Fp.val)
    x1, _, _ = O.bicgstab(A, b, precond=P, tol=1e-300, max_iters=1)
    x2, _, _, threads = O.bicgstab_parallel(A, b, precond=P, tol=1e-300, max_iters=1)
    assert threads >= 1
    assert np.linalg.norm(x2 - x1) <= 1e-12 * np.linalg.norm(x1)
    x1, it1, r1 = O.bicgstab(A, b, precond=P, tol=1e-12, max_iters=5000)
    x2, it2, r2, _ = O.bicgstab_parallel(A, b, precond=P, tol=1e-12, max_iters=5000)
    print(name, "iterations", it1, it2, "threads", threads)
    assert r1 <= 1e-12 and r2 <= 1e-12
    assert np.linalg.norm(x2 - x1) <= 1e-9 * np.linalg.norm(x1)
