"""Pins the oracle's step composition and stencil operators against what
the mathematics fixes, independently of the oracle's own code:

- the low-storage RK3 / Crank-Nicolson recurrence (rk3.c:3-6,
  chanfast.c:444-483) on a spanwise Fourier mode times a discrete z
  eigenvector, and on the plane-mean mode under a constant pressure
  gradient with the forcing folded into the next zeta term
  (chanfast.c:485-530);
- a flow moving with the walls (uframe, vframe != 0) staying steady, which
  fixes the sign of the boundary lift (chanfast.c:472-477);
- polynomial exactness of the x-y Laplacian, the z Helmholtz operators
  (2nd order exact for cubics, the 4th-order compact-staggered forms for
  quintics, on uniform spacing) and their convergence on the stretched grid;
- QUICK face values (Leonard 1979: exact for quadratics, error -/+ h^3 q'''/16
  on the upwind / downwind-biased stencils) and the scalar flux divergence;
- buoyancy interpolation exactness (linear at 2nd, cubic at 4th order);
- the 4th-order wall closure rows (diffuseu.c:703-802, diffusew.c:691-766)
  as relations on the filled ghosts, and the closure-reduced banded solve
  against a dense solve of the same rows.

CPU only.
"""
import numpy as np
import pytest

import oracle_lib as ol

L = ol.lib()
dp = ol.dptr

A1, A2 = 9.0 / 8.0, -1.0 / 24.0           # fourth.c:3-6
RK_A = (29 / 96, -3 / 40, 1 / 6)            # rk3.c:3-6
RK_B = (37 / 160, 5 / 24, 1 / 6)
RK_G = (8 / 15, 5 / 12, 3 / 4)
RK_Z = (0.0, -17 / 60, -5 / 12)


def tanh_dz(nz, beta=2.0, lz=2.0):
    j = np.arange(nz + 1)
    zf = 0.5 * lz * (1.0 + np.tanh(beta * (2.0 * j / nz - 1.0)) / np.tanh(beta))
    return np.diff(zf)


def grid(nx=8, ny=8, nz=12, dz=None, lx=2 * np.pi, ly=np.pi):
    if dz is None:
        dz = tanh_dz(nz)
    return ol.PaddedGrid(nx, ny, nz, lx, ly, dz, ng=3)


def coords(g):
    """Index coordinates of the padded block: x_i = (i - ng) dx, y_j = (j - ng) dy,
    cell centres z_k and the faces above them (from the mirrored dz)."""
    ng = g.ng
    x = (np.arange(g.g.n0) - ng) * g.g.dx
    y = (np.arange(g.g.n1) - ng) * g.g.dy
    d = g.dzbuf
    top = np.concatenate([[0.0], np.cumsum(d)]) - d[:ng].sum()   # face below level k at top[k]
    zc = top[:-1] + 0.5 * d
    zf = top[1:]
    return x, y, zc, zf


def poly(coef, t):
    return sum(c * t ** p for p, c in enumerate(coef))


def dpoly(coef, t, m=1):
    c = np.polynomial.polynomial.polyder(np.asarray(coef, float), m)
    return np.polynomial.polynomial.polyval(t, c)


# --------------------------------------------------- x-y Laplacian

@pytest.mark.parametrize("order,deg", [(2, 3), (4, 5)])
@pytest.mark.parametrize("kind", ["cell", "face"])
def test_laplacexy_polynomial_exact(order, deg, kind):
    """add_laplacexy_operator (diffuseu.c:286-359, diffusew.c:290-365): the
    2nd-order three-point form is exact for cubics, the 4th-order
    compact-staggered form d(a1 D1 + a2 D3) for quintics (any other a1, a2
    leaves an h^2 or h^4 term)."""
    g = grid(nx=12, ny=12, nz=8)
    x, y, zc, _ = coords(g)
    rng = np.random.default_rng(order + deg)
    px, py = rng.standard_normal(deg + 1), rng.standard_normal(deg + 1)
    q = poly(px, x)[:, None, None] + poly(py, y)[None, :, None] + np.zeros(g.shape)
    h = np.zeros(g.shape)
    nu = 0.37
    fn = L.orc_laplacexy_cell if kind == "cell" else L.orc_laplacexy_face
    fn(dp(q), dp(h), g.ref, order, nu)
    want = nu * (dpoly(px, x, 2)[:, None, None] + dpoly(py, y, 2)[None, :, None]) + np.zeros(g.shape)
    ng = g.ng
    ktop = g.g.n2 - ng - (0 if kind == "cell" else 1)
    sl = (slice(ng, -ng), slice(ng, -ng), slice(ng, ktop))
    assert np.abs(h[sl] - want[sl]).max() <= 1e-9 * np.abs(want[sl]).max()
    assert np.all(h[:, :, ktop:] == 0.0)          # rows above the last one are not touched


# --------------------------------------------- z Helmholtz operators

@pytest.mark.parametrize("order,deg", [(2, 3), (4, 5)])
@pytest.mark.parametrize("kind", ["cell", "face"])
def test_helmholtz1d_polynomial_exact_uniform(order, deg, kind):
    """helmholtz1d_operator (diffuseu.c:107-170, diffusew.c:107-172) on
    uniform spacing: h = q + nudt q'' exactly for cubics (2nd order) and
    quintics (4th order), ghosts filled with the same polynomial."""
    nz = 14
    g = grid(nx=6, ny=6, nz=nz, dz=np.full(nz, 2.0 / nz))
    _, _, zc, zf = coords(g)
    z = zc if kind == "cell" else zf
    rng = np.random.default_rng(10 + order + deg)
    pz = rng.standard_normal(deg + 1)
    q = np.broadcast_to(poly(pz, z), g.shape).copy()
    h = np.zeros(g.shape)
    nudt = 0.013
    (L.orc_helmholtz1d_cell if kind == "cell" else L.orc_helmholtz1d_face)(dp(q), dp(h), g.ref, order, nudt)
    ng = g.ng
    ktop = g.g.n2 - ng - (0 if kind == "cell" else 1)
    want = q + nudt * np.broadcast_to(dpoly(pz, z, 2), g.shape)
    sl = (slice(None), slice(None), slice(ng, ktop))
    assert np.abs(h[sl] - want[sl]).max() <= 1e-11 * np.abs(want[sl]).max()


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("kind", ["cell", "face"])
def test_helmholtz1d_converges_on_tanh_grid(order, kind):
    """On the tanh-stretched grid the operator approximates q + nudt q''
    with an error falling at least as h^2 away from the walls (the metric
    terms: centre distances, cell widths and the stencil's half-level
    spacings).  In the rows next to the walls the mirrored ghost widths
    (ghost.c:31-54) break the smooth mapping: the 2nd-order rows still
    converge as h there; the 4th-order rows, whose 3h differences span the
    mirror, do not converge for a function sampled at the mirrored points
    (they are pinned on uniform spacing above and through the closures
    below)."""
    errs, errw = [], []
    for nz in (32, 64, 128):
        g = grid(nx=6, ny=6, nz=nz)
        _, _, zc, zf = coords(g)
        z = zc if kind == "cell" else zf
        q = np.broadcast_to(np.cos(1.7 * z + 0.3), g.shape).copy()
        h = np.zeros(g.shape)
        (L.orc_helmholtz1d_cell if kind == "cell" else L.orc_helmholtz1d_face)(dp(q), dp(h), g.ref, order, 1.0)
        ng = g.ng
        ktop = g.g.n2 - ng - (0 if kind == "cell" else 1)
        want = -1.7 ** 2 * np.cos(1.7 * z + 0.3)
        e = np.abs(h[2, 2, ng:ktop] - q[2, 2, ng:ktop] - want[ng:ktop])
        errs.append(e[3:-3].max())
        errw.append(max(e[:3].max(), e[-3:].max()))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    ratew = np.log2(np.array(errw[:-1]) / np.array(errw[1:]))
    assert rates.min() >= 1.6 and rates[-1] >= 1.85, (errs, rates)
    if order == 2:
        assert ratew.min() >= 0.9, (errw, ratew)


# --------------------------------------------- 4th-order wall closures

def dstag(q, dzb, k):
    """A1/A2 staggered derivative of q at the level between k and k+1,
    over the half-level spacing there (the reference's D_{k+1/2})."""
    return (A1 * (q[k + 1] - q[k]) + A2 * (q[k + 2] - q[k - 1])) / (0.5 * (dzb[k] + dzb[k + 1]))


def test_cell_closure_rows_order4():
    """ghost_stencil_bot4/top4 (diffuseu.c:703-802), Dirichlet: the first
    ghost repeats the first interior value, the 4-point wall interpolant
    equals the wall value, and the second difference of the staggered
    derivative across the wall vanishes."""
    nz = 12
    g = grid(nz=nz)
    rng = np.random.default_rng(21)
    u = rng.standard_normal(g.shape)
    bcb, bct = 0.3, -0.7
    L.orc_apply_bc_cell(dp(u), g.ref, 4, bcb, bct, 0, 0)
    ng, n2, d = g.ng, g.g.n2, g.dzbuf
    col = u[4, 5]
    kb, kt = ng - 1, n2 - ng                      # first ghost at each wall
    assert col[kb] == col[kb + 1] and col[kt] == col[kt - 1]
    wall_b = 0.5 * A2 * (col[kb - 1] + col[kb + 2]) + 0.5 * (A1 + 2 * A2) * (col[kb] + col[kb + 1])
    wall_t = 0.5 * A2 * (col[kt + 1] + col[kt - 2]) + 0.5 * (A1 + 2 * A2) * (col[kt] + col[kt - 1])
    assert abs(wall_b - bcb) <= 1e-14 and abs(wall_t - bct) <= 1e-14
    # bottom: D_{k-1/2} - 2 D_{k+1/2} + D_{k+3/2} = 0 at the first ghost k
    k = kb
    sb = dstag(col, d, k - 1) - 2 * dstag(col, d, k) + dstag(col, d, k + 1)
    k = kt
    st = dstag(col, d, k - 2) - 2 * dstag(col, d, k - 1) + dstag(col, d, k)
    scale = np.abs(col).max() / d.min()
    assert abs(sb) <= 1e-12 * scale and abs(st) <= 1e-12 * scale


def test_face_closure_rows_order4():
    """ghost_stencil for w (diffusew.c:691-766): the wall face holds the
    wall value, the first ghost is the odd reflection 2 bc - w_1, and the
    third ghost equalises the staggered derivative in the two cells next
    to the wall face."""
    nz = 12
    g = grid(nz=nz)
    rng = np.random.default_rng(22)
    w = rng.standard_normal(g.shape)
    bcb, bct = 0.2, -0.1
    L.orc_apply_bc_face(dp(w), g.ref, 4, bcb, bct)
    ng, n2, d = g.ng, g.g.n2, g.dzbuf
    col = w[3, 2]
    kb, kt = ng - 1, n2 - ng - 1                  # wall faces
    assert col[kb] == bcb and col[kt] == bct
    assert abs(col[kb - 1] - (2 * bcb - col[kb + 1])) <= 1e-15
    assert abs(col[kt + 1] - (2 * bct - col[kt - 1])) <= 1e-15

    def dface(k):   # derivative of w in cell k (between faces k-1 and k)
        return (A1 * (col[k] - col[k - 1]) + A2 * (col[k + 1] - col[k - 2])) / d[k]
    scale = np.abs(col).max() / d.min()
    assert abs(dface(kb) - dface(kb + 1)) <= 1e-12 * scale
    assert abs(dface(kt) - dface(kt + 1)) <= 1e-12 * scale


def probe_matrix(g, kind, order, nudt):
    """Dense (n2 x n2) matrix of q -> helmholtz1d(q) on one padded column
    (rows of the interior levels; other rows zero)."""
    n2 = g.g.n2
    M = np.zeros((n2, n2))
    fn = L.orc_helmholtz1d_cell if kind == "cell" else L.orc_helmholtz1d_face
    for c in range(n2):
        q = g.zeros()
        q[:, :, c] = 1.0
        h = g.zeros()
        fn(dp(q), dp(h), g.ref, order, nudt)
        M[:, c] = h[2, 2]
    return M


@pytest.mark.parametrize("kind", ["cell", "face"])
def test_solve_order4_vs_dense(kind):
    """solve_helmholtz1d_{u,w}4 (diffuseu.c:173-283, diffusew.c:175-287):
    the ghost-row reduction, heptadiagonal solve and ghost recovery equal a
    dense solve of the interior operator rows plus the three closure rows
    per wall (as in the two tests above)."""
    nz = 13
    g = grid(nx=6, ny=6, nz=nz)
    ng, n2, d = g.ng, g.g.n2, g.dzbuf
    nudt = -0.021
    bcb, bct = (0.25, -0.5) if kind == "cell" else (0.0, 0.0)
    M = probe_matrix(g, kind, 4, nudt)
    rng = np.random.default_rng(23)
    h = rng.standard_normal(g.shape)
    A = np.zeros((n2, n2))
    rhs = np.zeros(n2)
    ktop = n2 - ng - (0 if kind == "cell" else 1)
    A[ng:ktop] = M[ng:ktop]
    rhs[ng:ktop] = h[2, 2, ng:ktop]
    rows = []
    if kind == "cell":
        kb, kt = ng - 1, n2 - ng
        # first ghost = first interior
        rows.append(({kb: 1.0, kb + 1: -1.0}, 0.0))
        rows.append(({kt: 1.0, kt - 1: -1.0}, 0.0))
        # wall interpolant = bc
        rows.append(({kb - 1: 0.5 * A2, kb + 2: 0.5 * A2, kb: 0.5 * (A1 + 2 * A2), kb + 1: 0.5 * (A1 + 2 * A2)}, bcb))
        rows.append(({kt + 1: 0.5 * A2, kt - 2: 0.5 * A2, kt: 0.5 * (A1 + 2 * A2), kt - 1: 0.5 * (A1 + 2 * A2)}, bct))

        def drow(k, s):   # s * D_{k+1/2} as a row
            hh = 0.5 * (d[k] + d[k + 1])
            return {k + 1: s * A1 / hh, k: -s * A1 / hh, k + 2: s * A2 / hh, k - 1: -s * A2 / hh}
        for base in (kb - 1, kt - 2):
            r = {}
            for kk, s in ((base, 1.0), (base + 1, -2.0), (base + 2, 1.0)):
                for c, v in drow(kk, s).items():
                    r[c] = r.get(c, 0.0) + v
            rows.append((r, 0.0))
        out = g.zeros()
        L.orc_solve1d_cell(dp(out), dp(h), g.ref, 4, nudt, bcb, bct, 0, 0)
        lo, hi = ng - 3, n2 - ng + 3
    else:
        kb, kt = ng - 1, n2 - ng - 1
        rows.append(({kb: 1.0}, bcb))
        rows.append(({kt: 1.0}, bct))
        rows.append(({kb - 1: 1.0, kb + 1: 1.0}, 2 * bcb))
        rows.append(({kt + 1: 1.0, kt - 1: 1.0}, 2 * bct))

        def drow(k, s):   # s * derivative of w in cell k
            return {k: s * A1 / d[k], k - 1: -s * A1 / d[k], k + 1: s * A2 / d[k], k - 2: -s * A2 / d[k]}
        for k in (kb, kt):
            r = {}
            for kk, s in ((k, 1.0), (k + 1, -1.0)):
                for c, v in drow(kk, s).items():
                    r[c] = r.get(c, 0.0) + v
            rows.append((r, 0.0))
        out = g.zeros()
        L.orc_solve1d_face(dp(out), dp(h), g.ref, 4, nudt, bcb, bct)
        lo, hi = ng - 3, n2 - ng + 2
    # the unknowns are levels lo..hi-1; rows: interior operator + 6 closure rows
    idx = list(range(lo, hi))
    B = np.zeros((len(idx), len(idx)))
    b = np.zeros(len(idx))
    nint = ktop - ng
    B[:nint] = A[ng:ktop][:, lo:hi]
    b[:nint] = rhs[ng:ktop]
    for q, (r, val) in enumerate(rows):
        for c, v in r.items():
            B[nint + q, c - lo] += v
        b[nint + q] = val
    ref = np.linalg.solve(B, b)
    got = out[2, 2, lo:hi]
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_operator_matrix_order2_textbook():
    """The probed 2nd-order cell operator with its Dirichlet ghost
    (0.5 (q_g + q_1) = 0) equals the textbook finite-volume matrix on the
    stretched grid: centre-to-centre spacings inside, cell width outside."""
    nz = 11
    g = grid(nx=6, ny=6, nz=nz)
    ng, n2, d = g.ng, g.g.n2, g.dzbuf
    Lz = operator_z(g, "cell", 2)
    T = np.zeros((nz, nz))
    for k in range(nz):
        K = k + ng
        hp, hm = 0.5 * (d[K] + d[K + 1]), 0.5 * (d[K] + d[K - 1])
        T[k, k] = -1 / (hp * d[K]) - 1 / (hm * d[K])
        if k + 1 < nz:
            T[k, k + 1] = 1 / (hp * d[K])
        else:
            T[k, k] -= 1 / (hp * d[K])
        if k > 0:
            T[k, k - 1] = 1 / (hm * d[K])
        else:
            T[k, k] -= 1 / (hm * d[K])
    assert np.abs(Lz - T).max() <= 1e-12 * np.abs(T).max()


# ------------------------------------------------- QUICK, scalar, buoyancy

@pytest.mark.parametrize("direction", [0, 1, 2])
@pytest.mark.parametrize("sign", [+1.0, -1.0])
def test_quick_face_values_cubic(direction, sign):
    """flux_quick (scalar.c:113-162): the face value is the quadratic
    through the two neighbours and the next upwind point (q = -1/8).  On a
    cubic the face value exceeds the exact one by h^3 q'''/16 with the
    upwind stencil (positive flux) and falls short by as much with the
    other: the Lagrange remainder (t - t_a)(t - t_b)(t - t_c) q'''/6 at the
    face is -3h^3/8 q'''/6 for the upwind stencil, +3h^3/8 q'''/6 for the
    other."""
    nz = 12
    h_z = 2.0 / nz
    g = grid(nx=10, ny=10, nz=nz, dz=np.full(nz, h_z))
    x, y, zc, zf = coords(g)
    pc = np.array([0.3, -0.7, 0.45, 0.9])                 # c = cubic along the direction
    c = np.zeros(g.shape)
    if direction == 0:
        c[:] = poly(pc, x)[:, None, None]; hh = g.g.dx; tf = x + 0.5 * hh
    elif direction == 1:
        c[:] = poly(pc, y)[None, :, None]; hh = g.g.dy; tf = y + 0.5 * hh
    else:
        c[:] = poly(pc, zc)[None, None, :]; hh = h_z; tf = zf
    vel = [np.zeros(g.shape) for _ in range(3)]
    vel[direction][:] = sign * 0.8
    m = [np.zeros(g.shape) for _ in range(3)]
    L.orc_flux_quick(dp(vel[0]), dp(vel[1]), dp(vel[2]), dp(c), dp(m[0]), dp(m[1]), dp(m[2]), g.ref)
    area = [g.g.dy * g.dzbuf[None, None, :], g.g.dx * g.dzbuf[None, None, :], g.g.dx * g.g.dy]
    face = m[direction] / (area[direction] * vel[direction])
    exact = poly(pc, tf) + sign * hh ** 3 * 6 * pc[3] / 16
    sl = [slice(1, -2)] * 3
    shape = [1, 1, 1]
    shape[direction] = -1
    want = np.broadcast_to(exact.reshape(shape), g.shape)[tuple(sl)]
    assert np.abs(face[tuple(sl)] - want).max() <= 1e-12 * np.abs(want).max()


def test_quick_advect_scalar_quadratic_exact():
    """flux_quick + advect_scalar (scalar.c:79-162): with a uniform velocity
    the face values of a quadratic are exact, so the flux divergence is
    -U dc/dx (and likewise in y and z) at every interior cell."""
    nz = 10
    h_z = 2.0 / nz
    g = grid(nx=10, ny=10, nz=nz, dz=np.full(nz, h_z))
    x, y, zc, _ = coords(g)
    px, py, pz = [0.2, -0.4, 0.35], [-0.1, 0.6, -0.25], [0.5, 0.3, 0.15]
    c = poly(px, x)[:, None, None] + poly(py, y)[None, :, None] + poly(pz, zc)[None, None, :]
    U, V, W = 0.7, -0.45, 0.3
    u, v, w = np.full(g.shape, U), np.full(g.shape, V), np.full(g.shape, W)
    m = [np.zeros(g.shape) for _ in range(3)]
    L.orc_flux_quick(dp(u), dp(v), dp(w), dp(c), dp(m[0]), dp(m[1]), dp(m[2]), g.ref)
    f = np.zeros(g.shape)
    L.orc_advect_scalar(dp(m[0]), dp(m[1]), dp(m[2]), dp(f), g.ref)
    want = -(U * dpoly(px, x)[:, None, None] + V * dpoly(py, y)[None, :, None] + W * dpoly(pz, zc)[None, None, :])
    ng = g.ng
    sl = (slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
    assert np.abs(f[sl] - want[sl]).max() <= 1e-11 * np.abs(want[sl]).max()


@pytest.mark.parametrize("order,deg", [(2, 1), (4, 3)])
def test_buoyancy_interpolation_exact(order, deg):
    """add_buoyancy2/4 (scalar.c:6-76): c is interpolated to the u, v, w
    points (the faces between cells) by the 2-point mean (exact for linear
    c) or the 4-point 9/16, -1/16 form (exact for cubics), times betg."""
    nz = 12
    h_z = 2.0 / nz
    g = grid(nx=10, ny=10, nz=nz, dz=np.full(nz, h_z))
    x, y, zc, zf = coords(g)
    rng = np.random.default_rng(31 + order)
    px, py, pz = (rng.standard_normal(deg + 1) for _ in range(3))
    bx, by, bz = 0.3, -1.1, 2.5
    ng = g.ng
    sl = (slice(ng, -ng), slice(ng, -ng))
    for which, cfield, want in (
            (0, poly(px, x)[:, None, None] + np.zeros(g.shape), bx * poly(px, x + 0.5 * g.g.dx)[:, None, None]),
            (1, poly(py, y)[None, :, None] + np.zeros(g.shape), by * poly(py, y + 0.5 * g.g.dy)[None, :, None]),
            (2, poly(pz, zc)[None, None, :] + np.zeros(g.shape), bz * poly(pz, zf)[None, None, :])):
        fu, fv, fw = g.zeros(), g.zeros(), g.zeros()
        L.orc_add_buoyancy(dp(np.ascontiguousarray(cfield)), dp(fu), dp(fv), dp(fw), bx, by, bz, g.ref, order)
        got = (fu, fv, fw)[which]
        ktop = g.g.n2 - ng - (1 if which == 2 else 0)
        want = np.broadcast_to(want, g.shape)
        err = np.abs(got[sl + (slice(ng, ktop),)] - want[sl + (slice(ng, ktop),)]).max()
        assert err <= 1e-12 * np.abs(want).max(), which


# ------------------------------------------------------------ RK3 step

def operator_z(g, kind, order):
    """nz x nz matrix of the homogeneous-Dirichlet z operator d2/dz2 the
    solver sees: interior rows of helmholtz1d with ghosts filled by
    apply_bc (bc = 0), minus the identity."""
    ng, n2 = g.ng, g.g.n2
    nzk = g.g.nz - (0 if kind == "cell" else 1)
    Lz = np.zeros((nzk, nzk))
    for c in range(nzk):
        q = g.zeros()
        q[:, :, ng + c] = 1.0
        if kind == "cell":
            L.orc_apply_bc_cell(dp(q), g.ref, order, 0.0, 0.0, 0, 0)
        else:
            L.orc_apply_bc_face(dp(q), g.ref, order, 0.0, 0.0)
        h = g.zeros()
        (L.orc_helmholtz1d_cell if kind == "cell" else L.orc_helmholtz1d_face)(dp(q), dp(h), g.ref, order, 1.0)
        Lz[:, c] = h[2, 2, ng:ng + nzk] - q[2, 2, ng:ng + nzk]
    return Lz


def rk_cfg(nx, ny, nz, order, dz, **kw):
    cfg = dict(nx=nx, ny=ny, nz=nz, order=order, scalar=0, lx=2 * np.pi, ly=np.pi, dt=1e-2, want_cfl=0.5,
               cflmode=0, nu=0.1, prandtl=1.0, dpdx=0.0, want_uvolavg=0.0, uvolavgmode=0, want_cvolavg=0.0,
               cvolavgmode=0, uframe=0.0, vframe=0.0, betg_x=0.0, betg_y=0.0, betg_z=0.0, dz=dz)
    cfg.update(kw)
    return cfg


def lam_xy(order, theta):
    """Modified wavenumber of the x-y Laplacian: 4 sin^2(t/2) (2nd order),
    (2 a1 sin(t/2) + 2 a2 sin(3t/2))^2 for the compact-staggered 4th order."""
    if order == 2:
        return 4 * np.sin(theta / 2) ** 2
    return (2 * A1 * np.sin(theta / 2) + 2 * A2 * np.sin(1.5 * theta)) ** 2


@pytest.mark.parametrize("order", [2, 4])
def test_rk3_spanwise_mode_times_z_eigenvector(order):
    """u = sin(2 pi j / ny) phi(z), v = w = 0, phi an eigenvector of the
    discrete z operator (Lz phi = mu phi): advection and pressure vanish,
    the x-y diffusion is explicit (gamma, zeta terms) and the z diffusion
    Crank-Nicolson (alpha, beta), so each substep multiplies the amplitude by
    a_{s+1} = [(1 + nu alpha dt mu) a_s + dt F (gamma a_s + zeta a_{s-1})]
              / (1 - nu beta dt mu),  F = -nu lam(theta) / dy^2."""
    nx, ny, nz = 6, 12, 15
    dz = tanh_dz(nz)
    cfg = rk_cfg(nx, ny, nz, order, dz)
    g = grid(nx=nx, ny=ny, nz=nz, dz=dz)
    Lz = operator_z(g, "cell", order)
    ev, V = np.linalg.eig(Lz)
    real = np.where((np.abs(ev.imag) < 1e-9 * np.abs(ev).max()) & (ev.real < 0))[0]
    pick = real[np.argsort(-ev.real[real])][2]          # the third slowest real mode
    mu = ev.real[pick]
    phi = V[:, pick].real
    phi /= np.abs(phi).max()
    theta = 2 * np.pi / ny
    jj = np.arange(ny)
    u0 = np.sin(theta * jj)[None, :, None] * phi[None, None, :] + np.zeros((nx, ny, nz))
    o = ol.OracleRun(cfg)
    o.put("u", u0)
    o.put("v", np.zeros((nx, ny, nz)))
    o.put("w", np.zeros((nx, ny, nz)))
    nsteps = 4
    o.steps(nsteps)
    nu, dt = cfg["nu"], cfg["dt"]
    F = -nu * lam_xy(order, theta) / g.g.dy ** 2
    a = 1.0
    for _ in range(nsteps):
        a_prev = None
        for s in range(3):
            num = (1 + nu * RK_A[s] * dt * mu) * a + dt * F * RK_G[s] * a
            if s > 0:
                num += dt * F * RK_Z[s] * a_prev
            a_prev, a = a, num / (1 - nu * RK_B[s] * dt * mu)
    got = o.get("u")
    assert abs(a - 1.0) > 1e-2                           # the mode really decayed
    assert np.abs(got - a * u0).max() <= 1e-13
    assert np.abs(o.get("v")).max() == 0.0 and np.abs(o.get("w")).max() <= 1e-15


@pytest.mark.parametrize("order", [2, 4])
def test_rk3_mean_mode_under_constant_pressure_gradient(order):
    """A plane-uniform u = U(z) under a fixed dpdx (uvolavgmode 0): the
    forcing m = -dpdx enters each substep through the unit-forcing profile
    h1 = S^-1 (gamma dt) (chanfast.c:498-518) and, folded into f
    (f_u += m), through the next substep's zeta term, so
    U_{s+1} = S^-1 [(I + nu alpha dt Lz) U_s + zeta dt m] + S^-1 (gamma dt) m,
    S = I - nu beta dt Lz (a numpy matrix recurrence)."""
    nx, ny, nz = 6, 6, 15
    dz = tanh_dz(nz)
    dpdx = 0.3
    cfg = rk_cfg(nx, ny, nz, order, dz, dpdx=dpdx, nu=0.05)
    g = grid(nx=nx, ny=ny, nz=nz, dz=dz)
    Lz = operator_z(g, "cell", order)
    zc = np.cumsum(dz) - 0.5 * dz
    U = 1.5 * (1 - (zc - 1.0) ** 2) + 0.2 * np.sin(3 * zc)
    o = ol.OracleRun(cfg)
    o.put("u", np.broadcast_to(U, (nx, ny, nz)).copy())
    o.put("v", np.zeros((nx, ny, nz)))
    o.put("w", np.zeros((nx, ny, nz)))
    nsteps = 5
    log = o.steps(nsteps)
    nu, dt, m = cfg["nu"], cfg["dt"], -dpdx
    I = np.eye(nz)
    one = np.ones(nz)
    for _ in range(nsteps):
        for s in range(3):
            S = I - nu * RK_B[s] * dt * Lz
            rhs = (I + nu * RK_A[s] * dt * Lz) @ U + RK_Z[s] * dt * m * one
            U = np.linalg.solve(S, rhs) + np.linalg.solve(S, RK_G[s] * dt * one) * m
    got = o.get("u")
    assert np.abs(got - U[None, None, :]).max() <= 1e-13 * np.abs(U).max()
    assert np.allclose(log[:, 3], m, rtol=0, atol=1e-15)


@pytest.mark.parametrize("order", [2, 4])
def test_flow_moving_with_the_walls_is_steady(order):
    """uframe, vframe != 0 put the walls at u = -uframe, v = -vframe
    (chanfast.c:59-64); a uniform field equal to the wall velocity with
    the bulk target at that value (want_uvolavg = 0) is a steady solution,
    which holds only with the boundary lift h -= H(bc-only field) of the
    right sign (chanfast.c:472-477)."""
    nx, ny, nz = 8, 8, 13
    dz = tanh_dz(nz)
    uf, vf = 0.4, -0.25
    cfg = rk_cfg(nx, ny, nz, order, dz, uframe=uf, vframe=vf, uvolavgmode=1, want_uvolavg=0.0, nu=0.05)
    o = ol.OracleRun(cfg)
    o.put("u", np.full((nx, ny, nz), -uf))
    o.put("v", np.full((nx, ny, nz), -vf))
    o.put("w", np.zeros((nx, ny, nz)))
    log = o.steps(3)
    assert np.abs(o.get("u") + uf).max() <= 1e-14
    assert np.abs(o.get("v") + vf).max() <= 1e-14
    assert np.abs(o.get("w")).max() <= 1e-14
    assert np.abs(log[:, 3]).max() <= 1e-12          # (target - bulk) / vol1 at round-off


# ------------------------------------------- pressure scale, reductions, logs

B1, B2 = 9.0 / 16.0, -1.0 / 16.0           # fourth.c:3-6


@pytest.mark.parametrize("order", [2, 4])
def test_hydrostatic_pressure_balances_buoyancy(order):
    """A fluid at rest with the conduction profile c(z) and buoyancy
    betg_z c: the projection removes the whole forcing, so the pressure of
    every substep satisfies the discrete hydrostatic balance
    grad_z p = betg_z c (interpolated to the w faces) exactly.  This fixes
    the pressure scale: the Poisson and projection use (alpha + beta) dt
    (chanfast.c:639-655), which equals gamma + zeta per substep, so the
    forcing's (gamma + zeta) dt over (alpha + beta) dt is one."""
    nx, ny, nz = 8, 8, 13
    dz = tanh_dz(nz)
    bz = 2.0
    cfg = rk_cfg(nx, ny, nz, order, dz, scalar=1, betg_z=bz, nu=1e-13, prandtl=1.0)
    zc = np.cumsum(dz) - 0.5 * dz
    c = zc / 2.0 - 0.5                               # walls at -0.5 / +0.5 (chanfast.c:67-68)
    o = ol.OracleRun(cfg)
    o.put("u", np.zeros((nx, ny, nz)))
    o.put("v", np.zeros((nx, ny, nz)))
    o.put("w", np.zeros((nx, ny, nz)))
    o.put("c", np.broadcast_to(c, (nx, ny, nz)).copy())
    o.steps(1)
    p = o.get("p")[3, 4]
    cc = o.get("c")[3, 4]
    dzh = 0.5 * (dz[:-1] + dz[1:])
    if order == 2:
        grad = (p[1:] - p[:-1]) / dzh
        want = bz * 0.5 * (cc[:-1] + cc[1:])
    else:
        k = np.arange(1, nz - 2)
        grad = (A1 * (p[k + 1] - p[k]) + A2 * (p[k + 2] - p[k - 1])) / dzh[k]
        want = bz * (B1 * (cc[k] + cc[k + 1]) + B2 * (cc[k - 1] + cc[k + 2]))
    assert np.abs(grad - want).max() <= 1e-11 * np.abs(want).max()
    assert np.abs(o.get("w")).max() <= 1e-12 and np.abs(o.get("u")).max() <= 1e-12


def test_cfl_and_plane_averages():
    """get_cfl (get.c:390-417): the largest of |u| dt/dx, |v| dt/dy and
    |w| dt / dz_face over interior points (not their sum); get_xy_avg
    (get.c:281-307): the mean over the interior x-y plane at every level,
    ghosts included; get_z_avg (get.c:490-502): the dz-weighted mean over
    the interior levels."""
    g = grid(nx=8, ny=6, nz=11)
    ng, n2, d = g.ng, g.g.n2, g.dzbuf
    dt = 0.01
    for comp, val in ((0, 0.9), (1, -0.7), (2, 0.5)):
        f = [g.zeros(), g.zeros(), g.zeros()]
        f[comp][ng + 2, ng + 1, ng + 4] = val
        cfl = L.orc_cfl(dp(f[0]), dp(f[1]), dp(f[2]), g.ref, dt)
        h = (g.g.dx, g.g.dy, 0.5 * (d[ng + 4] + d[ng + 5]))[comp]
        assert abs(cfl - abs(val) * dt / h) <= 1e-15
    u, v, w = g.zeros(), g.zeros(), g.zeros()
    u[ng:-ng, ng:-ng, ng:-ng] = 0.3
    v[ng:-ng, ng:-ng, ng:-ng] = 0.3
    cfl = L.orc_cfl(dp(u), dp(v), dp(w), g.ref, dt)
    assert abs(cfl - 0.3 * dt / min(g.g.dx, g.g.dy)) <= 1e-15
    rng = np.random.default_rng(41)
    a = rng.standard_normal(g.shape)
    avg = np.zeros(n2)
    L.orc_xy_avg(dp(a), dp(avg), g.ref)
    assert np.abs(avg - a[ng:-ng, ng:-ng, :].mean(axis=(0, 1))).max() <= 1e-14
    vol = L.orc_z_avg(dp(avg), g.g.nz, ng, dp(g.dzbuf))
    assert abs(vol - (avg[ng:-ng] * d[ng:-ng]).sum() / d[ng:-ng].sum()) <= 1e-15


@pytest.mark.parametrize("order", [2, 4])
def test_pstat_log_on_poiseuille_profile(order):
    """The pstat record (chanfast.c:485-530): for u = 1.5 (1 - eta^2) on a
    uniform grid the bulk velocity tends to 1 as h^2 (midpoint rule), the
    centreline value of an odd nz is the profile's 1.5, the two wall
    gradients are opposite, and at 2nd order nu dU/dz at the walls tends
    to +-3 nu as h."""
    nu = 1e-3
    errb, errw = [], []
    for nz in (33, 65, 129):
        dz = np.full(nz, 2.0 / nz)
        cfg = rk_cfg(8, 8, nz, order, dz, dt=1e-9, nu=nu)
        zc = np.cumsum(dz) - 0.5 * dz
        o = ol.OracleRun(cfg)
        o.put("u", np.broadcast_to(1.5 * (1 - (zc - 1) ** 2), (8, 8, nz)).copy())
        o.put("v", np.zeros((8, 8, nz)))
        o.put("w", np.zeros((8, 8, nz)))
        log = o.steps(1)[0]
        assert abs(log[7] - 1.5) <= 1e-10
        assert abs(log[4] + log[5]) <= 1e-11 * nu
        errb.append(abs(log[6] - 1.0))
        errw.append(abs(log[4] - 3 * nu))
    assert min(errb[0] / errb[1], errb[1] / errb[2]) >= 3.5
    if order == 2:
        assert min(errw[0] / errw[1], errw[1] / errw[2]) >= 1.8
        assert errw[-1] <= 0.01 * 3 * nu


# ------------------------------------------------------------ statistics

def stat_fields(nx=12, ny=10, nz=7, seed=51):
    g = grid(nx=nx, ny=ny, nz=nz)
    rng = np.random.default_rng(seed)
    u, v = rng.uniform(size=g.shape), rng.uniform(size=g.shape)
    for a in (u, v):
        L.orc_periodic(dp(a), g.ref)
    return g, u, v


def test_moments_and_correlation_case15():
    """case15 (cases.c:4293-4424): the first central moment vanishes and the
    second equals the autocorrelation, per level (<= 1e-13); both also equal
    numpy's plane statistics over the interior points."""
    g, u, v = stat_fields()
    n2 = g.g.n2
    corr, m1, m2, m3, m4, cuv = (np.zeros(n2) for _ in range(6))
    L.orc_xy_corr(dp(u), dp(u), dp(corr), g.ref)
    L.orc_xy_corr(dp(u), dp(v), dp(cuv), g.ref)
    for en, m in ((1.0, m1), (2.0, m2), (3.0, m3), (4.0, m4)):
        L.orc_xy_moment(dp(u), dp(m), en, g.ref)
    assert np.abs(m1).max() < 1e-13 and np.abs(m2 - corr).max() < 1e-13
    I = u[3:-3, 3:-3, :]
    J = v[3:-3, 3:-3, :]
    du, dv = I - I.mean(axis=(0, 1)), J - J.mean(axis=(0, 1))
    assert np.allclose(cuv, (du * dv).mean(axis=(0, 1)), rtol=0, atol=1e-15)
    assert np.allclose(m3, (du ** 3).mean(axis=(0, 1)), rtol=0, atol=1e-15)
    assert np.allclose(m4, (du ** 4).mean(axis=(0, 1)), rtol=0, atol=1e-15)


@pytest.mark.parametrize("nx,ny", [(12, 10), (9, 10), (12, 7)])
def test_spectra_parseval_case16(nx, ny):
    """case16 (cases.c:4427-4732): the one-dimensional cross-spectra sum to
    the cross-correlation at every level -- the cospectrum counted twice
    except at kx = 0 and the Nyquist kx of an even length, the quadrature
    spectrum summing to zero (<= 1e-15 relative)."""
    g, u, v = stat_fields(nx=nx, ny=ny, nz=5, seed=52)
    n2 = g.g.n2

    def fft3d(a):   # FFT3D_FWD scaling (fft.c:161-202): 1/nx then 1/ny
        s = np.fft.fft(np.fft.rfft(a[3:-3, 3:-3, :], axis=0) / nx, axis=1) / ny
        return np.ascontiguousarray(np.stack([s.real, s.imag], -1))
    uh, vh = fft3d(u), fft3d(v)
    sx = np.zeros((nx // 2 + 1, n2, 2))
    sy = np.zeros((ny // 2 + 1, n2, 2))
    L.orc_xy_spec(dp(uh), dp(vh), dp(sx), dp(sy), nx, ny, n2)
    corr = np.zeros(n2)
    L.orc_xy_corr(dp(u), dp(v), dp(corr), g.ref)

    def total(spec, n):
        w = np.full(spec.shape[0], 2.0)
        w[0] = 1.0
        if n % 2 == 0:
            w[-1] = 1.0
        return (w[:, None] * spec[:, :, 0]).sum(0)
    scale = np.abs(corr).max()
    assert np.abs(total(sx, nx) - corr).max() <= 1e-14 * scale
    assert np.abs(total(sy, ny) - corr).max() <= 1e-14 * scale
