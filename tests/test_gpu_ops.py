"""GPU tests of the reference's operator layer (include/chanflow_ops.h):
every function with a reference name runs its HIP kernel on the padded
layout and is compared with the oracle's restatement of the same function
on the same inputs (<= 1e-13 relative; ghost levels and untouched entries
included), and the reference's own invariants are re-expressed on the GPU
operators: case1/8 (projection leaves div <= 1e-10, pressure.c through
cases.c:473-503), case2/4/9/11 (operator -> solve round trip <= 1e-9,
cases.c:560-865, :1112-1415, :2620-2926, :3172-3475) and the case0/case7
closed forms (cases.c:19-266, :2079-2327).  The pencil gather is bit-exact."""
import ctypes as C

import numpy as np
import pytest

import chanflow as cf
import oracle_lib as ol

pytestmark = pytest.mark.gpu

dp = ol.dptr
I = C.c_int
D = C.c_double
P = C.POINTER(C.c_double)
BCFN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double)
G6 = [I] * 6

SIGS = {
    "divergence": [P, P, P, P] + G6 + [D, D, P],
    "project": [P, P, P, P] + G6 + [D, D, P, D],
    "pressure_laplacian": [P, P] + G6 + [D, D, P, D],
    "helmholtz1d_operator_u": [P, P] + G6 + [P, D],
    "helmholtz1d_operator_w": [P, P] + G6 + [P, D],
    "solve_helmholtz1d_u": [P, P] + G6 + [P, D, D, D, I, I],
    "solve_helmholtz1d_w": [P, P] + G6 + [P, D, D, D, I, I],
    "add_laplacexy_operator_u": [P, P] + G6 + [D, D, D],
    "add_laplacexy_operator_w": [P, P] + G6 + [D, D, D],
    "helmholtz_operator_u": [P, P] + G6 + [D, D, P, D],
    "helmholtz_operator_w": [P, P] + G6 + [D, D, P, D],
    "apply_bc_u": [P] + G6 + [P, P, P, BCFN, BCFN, I, I],
    "apply_bc_w": [P] + G6 + [P, P, P, BCFN, BCFN, I, I],
    "pressure_poisson": [P, I, I, I, I, I, I, I, I, D, D, P, D],
}


def op(name, order):
    fn = getattr(cf.lib(), f"{name}{order}")
    fn.argtypes = SIGS[name]
    fn.restype = None
    return fn


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def tanh_dz(nz, beta=2.0, lz=2.0):
    j = np.arange(nz + 1)
    zf = 0.5 * lz * (1.0 + np.tanh(beta * (2.0 * j / nz - 1.0)) / np.tanh(beta))
    return np.diff(zf)


class Case:
    def __init__(self, nx=12, ny=10, nz=13, seed=0, dz=None):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.pg = ol.PaddedGrid(nx, ny, nz, 2 * np.pi, np.pi, tanh_dz(nz) if dz is None else dz)
        self.dz = self.pg.dzbuf.copy()
        self.dx, self.dy = self.pg.g.dx, self.pg.g.dy
        self.rng = np.random.default_rng(seed)
        self.g = (nx, ny, nz, 3, 3, 3)

    def rand(self):
        return self.rng.standard_normal(self.pg.shape)


@pytest.mark.parametrize("order", [2, 4])
def test_divergence_project_laplacian(order):
    c = Case(seed=1)
    u, v, w, p, d0 = c.rand(), c.rand(), c.rand(), c.rand(), c.rand()
    # divergence (order 4 also rewrites the w ghosts, in w)
    a = [x.copy() for x in (u, v, w, d0)]
    b = [x.copy() for x in (u, v, w, d0)]
    op("divergence", order)(dp(a[0]), dp(a[1]), dp(a[2]), dp(a[3]), *c.g, c.dx, c.dy, dp(c.dz))
    ol.lib().orc_divergence(dp(b[0]), dp(b[1]), dp(b[2]), dp(b[3]), c.pg.ref, order)
    assert rel(a[3], b[3]) <= 1e-13 and np.array_equal(a[2], b[2])
    # projection
    a = [x.copy() for x in (u, v, w)]
    b = [x.copy() for x in (u, v, w)]
    op("project", order)(dp(a[0]), dp(a[1]), dp(a[2]), dp(p), *c.g, c.dx, c.dy, dp(c.dz), 0.37)
    ol.lib().orc_project(dp(b[0]), dp(b[1]), dp(b[2]), dp(p), c.pg.ref, order, 0.37)
    for q in range(3):
        assert rel(a[q], b[q]) <= 1e-13, q
    # pressure Laplacian
    la, lb = d0.copy(), d0.copy()
    op("pressure_laplacian", order)(dp(p), dp(la), *c.g, c.dx, c.dy, dp(c.dz), 0.21)
    ol.lib().orc_pressure_laplacian(dp(p), dp(lb), c.pg.ref, order, 0.21)
    assert rel(la, lb) <= 1e-13


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("fam", ["u", "w"])
def test_z_operators_and_xy_laplacian(order, fam):
    c = Case(seed=2)
    q, h0 = c.rand(), c.rand()
    ha, hb = h0.copy(), h0.copy()
    op(f"helmholtz1d_operator_{fam}", order)(dp(q), dp(ha), *c.g, dp(c.dz), -0.013)
    (ol.lib().orc_helmholtz1d_cell if fam == "u" else ol.lib().orc_helmholtz1d_face)(dp(q), dp(hb), c.pg.ref, order, -0.013)
    assert rel(ha, hb) <= 1e-13
    ha, hb = h0.copy(), h0.copy()
    op(f"add_laplacexy_operator_{fam}", order)(dp(q), dp(ha), *c.g, c.dx, c.dy, 0.05)
    (ol.lib().orc_laplacexy_cell if fam == "u" else ol.lib().orc_laplacexy_face)(dp(q), dp(hb), c.pg.ref, order, 0.05)
    assert rel(ha, hb) <= 1e-13
    # helmholtz_operator = q + nudt (lap_xy + lap_z) on the interior rows it writes
    ha = h0.copy()
    op(f"helmholtz_operator_{fam}", order)(dp(q), dp(ha), *c.g, c.dx, c.dy, dp(c.dz), 0.02)
    hz, hxy = np.zeros_like(q), np.zeros_like(q)
    (ol.lib().orc_helmholtz1d_cell if fam == "u" else ol.lib().orc_helmholtz1d_face)(dp(q), dp(hz), c.pg.ref, order, 0.02)
    (ol.lib().orc_laplacexy_cell if fam == "u" else ol.lib().orc_laplacexy_face)(dp(q), dp(hxy), c.pg.ref, order, 0.02)
    lo, ng, n2 = (1 if order == 2 else 3), 3, c.pg.g.n2
    kt = n2 - ng - (1 if fam == "w" else 0)
    sl = (slice(lo, -lo), slice(lo, -lo), slice(ng, kt))
    assert rel(ha[sl], (hz + hxy)[sl]) <= 1e-13
    mask = np.ones(q.shape, bool)
    mask[sl] = False
    assert np.array_equal(ha[mask], h0[mask])


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("fam", ["u", "w"])
def test_apply_bc_and_solve(order, fam):
    c = Case(seed=3)
    q = c.rand()
    x = (np.arange(c.pg.g.n0) - 3) * c.dx
    y = (np.arange(c.pg.g.n1) - 3) * c.dy
    bcb, bct = (0.3, -0.7) if fam == "u" else (0.0, 0.0)
    fb, ft = BCFN(lambda a, b: bcb), BCFN(lambda a, b: bct)
    kind = 0 if fam == "u" else 20
    qa, qb = q.copy(), q.copy()
    op(f"apply_bc_{fam}", order)(dp(qa), *c.g, dp(x), dp(y), dp(c.dz), fb, ft, kind, kind)
    if fam == "u":
        ol.lib().orc_apply_bc_cell(dp(qb), c.pg.ref, order, bcb, bct, 0, 0)
    else:
        ol.lib().orc_apply_bc_face(dp(qb), c.pg.ref, order, bcb, bct)
    assert rel(qa, qb) <= 1e-13
    # per-column wall values reach their own column
    fx = BCFN(lambda a, b: np.sin(a) + 0.5 * np.cos(2 * b))
    qc = q.copy()
    op(f"apply_bc_{fam}", order)(dp(qc), *c.g, dp(x), dp(y), dp(c.dz), fx, ft, kind, kind)
    if order == 2 and fam == "u":
        want = np.sin(x)[:, None] + 0.5 * np.cos(2 * y)[None, :]
        assert np.abs(0.5 * (qc[:, :, 2] + qc[:, :, 3]) - want).max() <= 1e-14
    # the column solve, every level of u written
    h = c.rand()
    for tb, tt in ((0, 0), (1, 0)) if fam == "u" else ((20, 20),):
        ua, ub = c.rand(), np.zeros(q.shape)
        op(f"solve_helmholtz1d_{fam}", order)(dp(ua), dp(h), *c.g, dp(c.dz), -0.017, bcb, bct, tb, tt)
        if fam == "u":
            ol.lib().orc_solve1d_cell(dp(ub), dp(h), c.pg.ref, order, -0.017, bcb, bct, tb, tt)
        else:
            ol.lib().orc_solve1d_face(dp(ub), dp(h), c.pg.ref, order, -0.017, bcb, bct)
        assert rel(ua, ub) <= 1e-12, (tb, tt)


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("fam", ["u", "w"])
def test_roundtrip_cases_2_4_9_11(order, fam):
    """apply_bc -> helmholtz1d_operator -> solve_helmholtz1d gives the
    field back (cases.c:799-812 threshold 1e-9), all on the GPU."""
    c = Case(nx=8, ny=8, nz=24, seed=4)
    q = c.rand()
    x = (np.arange(c.pg.g.n0) - 3) * c.dx
    y = (np.arange(c.pg.g.n1) - 3) * c.dy
    bcb, bct = (0.25, -0.5) if fam == "u" else (0.0, 0.0)
    kind = 0 if fam == "u" else 20
    fb, ft = BCFN(lambda a, b: bcb), BCFN(lambda a, b: bct)
    op(f"apply_bc_{fam}", order)(dp(q), *c.g, dp(x), dp(y), dp(c.dz), fb, ft, kind, kind)
    h = np.zeros_like(q)
    op(f"helmholtz1d_operator_{fam}", order)(dp(q), dp(h), *c.g, dp(c.dz), 0.7)
    out = np.zeros_like(q)
    op(f"solve_helmholtz1d_{fam}", order)(dp(out), dp(h), *c.g, dp(c.dz), 0.7, bcb, bct, kind, kind)
    ng, n2 = 3, c.pg.g.n2
    ex = 1 if order == 2 else 3
    sl = slice(ng - ex, n2 - ng + ex - (1 if fam == "w" else 0))
    assert np.abs(out[:, :, sl] - q[:, :, sl]).max() <= 1e-9


def fft3d_fwd(a, pg):
    """FFT3D_FWD of chanfast.c:21-37 with fft.c scaling (1/nx, 1/ny), on the
    interior x-y of a padded field, every level: [kx][ky][n2] complex."""
    i = a[3:-3, 3:-3, :]
    return np.fft.fft(np.fft.rfft(i, axis=0) / i.shape[0], axis=1) / i.shape[1]


def fft3d_bwd(s, pg, nx):
    out = np.zeros(pg.shape)
    out[3:-3, 3:-3, :] = np.fft.irfft(np.fft.ifft(s, axis=1) * s.shape[1], n=nx, axis=0) * nx
    return out


@pytest.mark.parametrize("order", [2, 4])
def test_poisson_matches_oracle_with_ghosts(order):
    c = Case(nx=12, ny=10, nz=17, seed=5)
    nkx, n2 = c.nx // 2 + 1, c.pg.g.n2
    spec = c.rng.standard_normal((nkx, c.ny, n2, 2))
    a, b = spec.copy(), spec.copy()
    op("pressure_poisson", order)(dp(a), c.nx, nkx, 0, c.ny, c.ny, 0, c.nz, 3, c.dx, c.dy, dp(c.dz), 0.3)
    ol.lib().orc_poisson(dp(b), c.pg.ref, order, 0.3)
    assert rel(a, b) <= 1e-12


@pytest.mark.parametrize("order", [2, 4])
def test_projection_case1_on_gpu_operators(order):
    """case1/case8 (cases.c:426-503): divergence -> FFT -> Poisson ->
    inverse FFT -> projection with the GPU operators leaves |div| <= 1e-10,
    and the discrete Laplacian of p equals the divergence it came from."""
    c = Case(nx=16, ny=12, nz=17, seed=6)
    u, v, w = c.rand(), c.rand(), c.rand()
    x = (np.arange(c.pg.g.n0) - 3) * c.dx
    y = (np.arange(c.pg.g.n1) - 3) * c.dy
    zero = BCFN(lambda a, b: 0.0)
    for f in (u, v, w):
        ol.lib().orc_periodic(dp(f), c.pg.ref)
    op("apply_bc_w", order)(dp(w), *c.g, dp(x), dp(y), dp(c.dz), zero, zero, 20, 20)
    d = np.zeros_like(u)
    op("divergence", order)(dp(u), dp(v), dp(w), dp(d), *c.g, c.dx, c.dy, dp(c.dz))
    s = np.ascontiguousarray(np.stack([fft3d_fwd(d, c.pg).real, fft3d_fwd(d, c.pg).imag], -1))
    op("pressure_poisson", order)(dp(s), c.nx, c.nx // 2 + 1, 0, c.ny, c.ny, 0, c.nz, 3, c.dx, c.dy, dp(c.dz), 1.0)
    p = fft3d_bwd(s[..., 0] + 1j * s[..., 1], c.pg, c.nx)
    ol.lib().orc_periodic(dp(p), c.pg.ref)
    op("project", order)(dp(u), dp(v), dp(w), dp(p), *c.g, c.dx, c.dy, dp(c.dz), 1.0)
    for f in (u, v, w):
        ol.lib().orc_periodic(dp(f), c.pg.ref)
    op("apply_bc_w", order)(dp(w), *c.g, dp(x), dp(y), dp(c.dz), zero, zero, 20, 20)
    d2 = np.zeros_like(u)
    op("divergence", order)(dp(u), dp(v), dp(w), dp(d2), *c.g, c.dx, c.dy, dp(c.dz))
    assert np.abs(c.pg.interior(d2)).max() <= 1e-10
    lap = np.zeros_like(u)
    op("pressure_laplacian", order)(dp(p), dp(lap), *c.g, c.dx, c.dy, dp(c.dz), 1.0)
    assert np.abs(c.pg.interior(lap) - c.pg.interior(d)).max() <= 1e-10


@pytest.mark.parametrize("order,deg", [(2, 3), (4, 5)])
def test_laplacian_closed_form_case0_case7(order, deg):
    """case0 / case7: on uniform spacing the discrete Laplacian of a
    polynomial in x, y and z of degree 3 (2nd order) or 5 (4th order) is
    its exact Laplacian."""
    nz = 16
    c = Case(nx=12, ny=12, nz=nz, seed=7, dz=np.full(nz, 2.0 / nz))
    ng = 3
    xs = (np.arange(c.pg.g.n0) - ng) * c.dx
    ys = (np.arange(c.pg.g.n1) - ng) * c.dy
    zs = (np.arange(c.pg.g.n2) - ng + 0.5) * (2.0 / nz)
    pc = [c.rng.standard_normal(deg + 1) for _ in range(3)]
    val = lambda q, t: np.polynomial.polynomial.polyval(t, q)
    d2 = lambda q, t: np.polynomial.polynomial.polyval(t, np.polynomial.polynomial.polyder(q, 2))
    p = val(pc[0], xs)[:, None, None] + val(pc[1], ys)[None, :, None] + val(pc[2], zs)[None, None, :]
    want = d2(pc[0], xs)[:, None, None] + d2(pc[1], ys)[None, :, None] + d2(pc[2], zs)[None, None, :]
    lap = np.zeros(c.pg.shape)
    op("pressure_laplacian", order)(dp(np.ascontiguousarray(p)), dp(lap), *c.g, c.dx, c.dy, dp(c.dz), 1.0)
    sl = (slice(3, -3), slice(3, -3), slice(3, -3))
    assert np.abs(lap[sl] - want[sl]).max() <= 1e-9 * np.abs(want[sl]).max()


@pytest.mark.parametrize("dims", [(1, 1), (2, 3), (4, 2)])
def test_pencil_gather_bit_exact(dims):
    lib = cf.lib()
    lib.cf_pencil_gather.argtypes = [P, I, I, I, C.POINTER(I), C.POINTER(I), I, P]
    lib.cf_pencil_gather.restype = I
    nx, ny, nz = 17, 13, 11
    a = np.random.default_rng(8).standard_normal((nx, ny, nz))
    for c0 in range(dims[0]):
        for c1 in range(dims[1]):
            for pencil, pdim in enumerate(((1, 2, 3), (2, 1, 3), (2, 3, 1))):
                st, en, sz = (I * 3)(), (I * 3)(), (I * 3)()
                ol.lib().orc_partition(nx, ny, nz, (I * 3)(*pdim), (I * 2)(*dims), (I * 2)(c0, c1), st, en, sz)
                out = np.zeros(tuple(sz))
                assert lib.cf_pencil_gather(dp(a), nx, ny, nz, (I * 2)(*dims), (I * 2)(c0, c1), pencil, dp(out)) == 0
                want = a[st[0] - 1:en[0], st[1] - 1:en[1], st[2] - 1:en[2]]
                assert np.array_equal(out, want)
