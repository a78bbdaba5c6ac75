"""Pins the CPU oracle (oracle/) against known answers, closed forms,
invariants and brute force -- the checks the reference's own cases.c
makes (case0/1/2/4/6/14/16 semantics, decomp.c:103) plus dense solves on
tiny grids.  CPU only."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg

import oracle_lib as ol

L = ol.lib()
dp = ol.dptr


# ----------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 9, 12, 15, 16, 25, 30, 64, 96, 97, 128, 192, 256])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    z = np.ascontiguousarray(np.stack([x.real, x.imag], -1).ravel())
    L.orc_fft(dp(z), n, -1)
    y = z[0::2] + 1j * z[1::2]
    assert np.abs(y - np.fft.fft(x)).max() <= 1e-14 * max(1.0, np.abs(y).max()) * max(1, np.log2(n))
    z = np.ascontiguousarray(np.stack([x.real, x.imag], -1).ravel())
    L.orc_fft(dp(z), n, +1)
    y = z[0::2] + 1j * z[1::2]
    assert np.abs(y - n * np.fft.ifft(x)).max() <= 1e-14 * max(1.0, np.abs(y).max()) * max(1, np.log2(n))


@pytest.mark.parametrize("n", [2, 6, 15, 64, 96, 256])
def test_rfft_irfft_semantics(n):
    """r2c keeps X[0..n/2]; c2r ignores imag(X0) and imag(X[n/2]) (FFTW c2r)."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    X = np.zeros(2 * (n // 2 + 1))
    L.orc_rfft(dp(x), dp(X), n)
    Xc = X[0::2] + 1j * X[1::2]
    assert np.allclose(Xc, np.fft.rfft(x), rtol=0, atol=1e-13 * np.abs(Xc).max())
    X[1] += 5.0                                   # garbage in imag(X0)
    if n % 2 == 0:
        X[2 * (n // 2) + 1] += 3.0                # and imag(X[n/2])
    back = np.zeros(n)
    L.orc_irfft(dp(X), dp(back), n)
    assert np.allclose(back / n, x, rtol=0, atol=1e-14 * np.abs(x).max() * n)


def test_parseval_case16():
    """case16 semantics: sum of |u|^2 equals the sum over the spectrum."""
    n = 96
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    z = np.ascontiguousarray(np.stack([x.real, x.imag], -1).ravel())
    L.orc_fft(dp(z), n, -1)
    y = z[0::2] + 1j * z[1::2]
    assert abs(np.sum(np.abs(x) ** 2) - np.sum(np.abs(y) ** 2) / n) <= 1e-15 * n * np.sum(np.abs(x) ** 2)


# ------------------------------------------------------ banded solvers

def test_tridag_vs_dense():
    rng = np.random.default_rng(1)
    n = 40
    a, c = rng.standard_normal(n), rng.standard_normal(n)
    b = 3.0 + np.abs(a) + np.abs(c)
    r = rng.standard_normal(n)
    x = np.zeros(n)
    assert L.orc_tridag(dp(a), dp(b), dp(c), dp(r), dp(x), n) == 0
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    assert np.abs(x - np.linalg.solve(A, r)).max() <= 1e-14


def test_tridag_zero_pivot_reported():
    n = 3
    a = np.zeros(n); b = np.zeros(n); c = np.zeros(n); r = np.ones(n); x = np.zeros(n)
    assert L.orc_tridag(dp(a), dp(b), dp(c), dp(r), dp(x), n) != 0


def test_heptdag_vs_solve_banded():
    rng = np.random.default_rng(2)
    n = 37
    bands = rng.standard_normal((7, n))          # a b c d e f g: offsets -3..+3
    bands[3] = 8.0 + np.abs(bands).sum(0)
    h = rng.standard_normal(n)
    x = np.zeros(n)
    args = [np.ascontiguousarray(bands[q]) for q in range(7)]
    assert L.orc_heptdag(*[dp(a) for a in args], dp(h), dp(x), n) == 0
    ab = np.zeros((7, n))
    for q, off in enumerate(range(-3, 4)):       # row i holds column i + off
        for i in range(n):
            j = i + off
            if 0 <= j < n:
                ab[3 - off, j] = bands[q, i]
    ref = scipy.linalg.solve_banded((3, 3), ab, h)
    assert np.abs(x - ref).max() <= 1e-13


# ---------------------------------------------------------- index maps

def test_distribute_known_answer():
    """decomp.c:103: 17 points over 4 ranks -> (4, 4, 4, 5)."""
    st, en, sz = (C.c_int * 4)(), (C.c_int * 4)(), (C.c_int * 4)()
    L.orc_distribute(17, 4, st, en, sz)
    assert list(sz) == [4, 4, 4, 5]
    assert list(st) == [1, 5, 9, 13] and list(en) == [4, 8, 12, 17]


@pytest.mark.parametrize("n,p", [(1, 1), (7, 3), (16, 4), (193, 8), (5, 5), (256, 6)])
def test_distribute_brute_force(n, p):
    st, en, sz = (C.c_int * p)(), (C.c_int * p)(), (C.c_int * p)()
    L.orc_distribute(n, p, st, en, sz)
    sizes = list(sz)
    assert sum(sizes) == n and max(sizes) - min(sizes) <= 1
    assert sizes == sorted(sizes)                # larger blocks last
    assert st[0] == 1 and en[p - 1] == n
    for i in range(1, p):
        assert st[i] == en[i - 1] + 1


def test_partition_z_pencil():
    """Every rank's z-pencil tiles the global x-y plane exactly once."""
    nx, ny, nz, dims = 17, 10, 9, (3, 2)
    cover = np.zeros((nx, ny), int)
    for c0 in range(dims[0]):
        for c1 in range(dims[1]):
            lst, len_, lsz = (C.c_int * 3)(), (C.c_int * 3)(), (C.c_int * 3)()
            L.orc_partition(nx, ny, nz, (C.c_int * 3)(2, 3, 1), (C.c_int * 2)(*dims), (C.c_int * 2)(c0, c1),
                            lst, len_, lsz)
            assert lsz[2] == nz
            cover[lst[0] - 1:len_[0], lst[1] - 1:len_[1]] += 1
    assert (cover == 1).all()


# ----------------------------------------------------- grid helpers

def tanh_dz(nz, beta=2.0, lz=2.0):
    j = np.arange(nz + 1)
    zf = 0.5 * lz * (1.0 + np.tanh(beta * (2.0 * j / nz - 1.0)) / np.tanh(beta))
    return np.diff(zf)


def grid(nx=8, ny=6, nz=9, ng=3, dz=None, lx=2 * np.pi, ly=np.pi):
    if dz is None:
        dz = tanh_dz(nz)
    return ol.PaddedGrid(nx, ny, nz, lx, ly, dz, ng=ng)


def test_ghost_even_mirror():
    g = grid(nz=7)
    d = g.dzbuf
    assert np.array_equal(d[:3], d[3:6][::-1]) and np.array_equal(d[-3:], d[-6:-3][::-1])


# ------------------------------------------------- wall closures

@pytest.mark.parametrize("order", [2, 4])
def test_cell_dirichlet_closure(order):
    """The ghost values make the wall interpolant equal the wall value:
    0.5 (q_g + q_1) = bc (order 2, diffuseu.c:629-633) and the 4-point
    interpolant a2/2 (q_-2 + q_2) + (a1 + 2 a2)/2 (q_-1 + q_1) = bc (order 4)."""
    g = grid(nz=12)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(g.shape)
    L.orc_apply_bc_cell(dp(u), g.ref, order, 0.3, -0.7, 0, 0)
    ng, n2 = g.ng, g.g.n2
    a1, a2 = 1.125, -0.125 / 3.0
    for bc, lo, hi in ((0.3, ng - 1, ng), (-0.7, n2 - ng, n2 - ng - 1)):
        if order == 2:
            val = 0.5 * (u[:, :, lo] + u[:, :, hi])
        else:
            s = 1 if hi > lo else -1
            val = 0.5 * a2 * (u[:, :, lo - s] + u[:, :, hi + s]) + 0.5 * (a1 + 2 * a2) * (u[:, :, lo] + u[:, :, hi])
        assert np.abs(val - bc).max() <= 1e-14


@pytest.mark.parametrize("order", [2, 4])
def test_face_dirichlet_closure(order):
    g = grid(nz=12)
    rng = np.random.default_rng(4)
    w = rng.standard_normal(g.shape)
    L.orc_apply_bc_face(dp(w), g.ref, order, 0.0, 0.0)
    ng, n2 = g.ng, g.g.n2
    assert np.all(w[:, :, ng - 1] == 0.0) and np.all(w[:, :, n2 - ng - 1] == 0.0)
    if order == 4:   # odd reflection about the wall face (diffusew.c:714-740)
        assert np.abs(w[:, :, ng - 2] + w[:, :, ng]).max() <= 1e-15
        assert np.abs(w[:, :, n2 - ng] + w[:, :, n2 - ng - 2]).max() <= 1e-15


# ------------------------------------ Helmholtz operator <-> solver

@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("kind", ["cell", "face"])
def test_helmholtz_roundtrip(order, kind):
    """case2/4/9/11 semantics on the 1-D z operator: u -> apply BC ->
    h = (I + nudt d2/dz2) u -> solve -> u (reference tolerance 1e-9)."""
    g = grid(nx=6, ny=6, nz=20)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(g.shape)
    h = np.zeros(g.shape)
    out = np.zeros(g.shape)
    nudt = 0.7
    ng, n2 = g.ng, g.g.n2
    if kind == "cell":
        L.orc_apply_bc_cell(dp(u), g.ref, order, 0.2, -0.4, 0, 0)
        L.orc_helmholtz1d_cell(dp(u), dp(h), g.ref, order, nudt)
        L.orc_solve1d_cell(dp(out), dp(h), g.ref, order, nudt, 0.2, -0.4, 0, 0)
        sl = slice(ng - (1 if order == 2 else 3), n2 - ng + (1 if order == 2 else 3))
    else:
        L.orc_apply_bc_face(dp(u), g.ref, order, 0.0, 0.0)
        L.orc_helmholtz1d_face(dp(u), dp(h), g.ref, order, nudt)
        L.orc_solve1d_face(dp(out), dp(h), g.ref, order, nudt, 0.0, 0.0)
        sl = slice(ng - (1 if order == 2 else 3), n2 - ng - 1 + (1 if order == 2 else 3))
    err = np.abs(out[:, :, sl] - u[:, :, sl]).max()
    assert err <= 1e-9                                          # cases.c:799-812 threshold
    assert err <= 1e-13 * np.abs(u[:, :, sl]).max()             # round-off of the ghost values


def test_helmholtz_solve_vs_dense_cell2():
    """Brute force: assemble (I - nudt d2/dz2) with the Dirichlet ghost
    relation q_g = 2 bc - q_1 eliminated, solve densely, compare."""
    nz = 11
    dz = tanh_dz(nz)
    g = grid(nx=6, ny=6, nz=nz, dz=dz)
    ng = g.ng
    d = g.dzbuf
    nudt, bcb, bct = -0.013, 0.25, -0.5
    rng = np.random.default_rng(6)
    h = rng.standard_normal(g.shape)
    A = np.zeros((nz, nz))
    rhs = h[3, 3, ng:ng + nz].copy()
    for k in range(nz):
        K = k + ng
        hp, hm = 0.5 * (d[K + 1] + d[K]), 0.5 * (d[K - 1] + d[K])
        cp, cm = nudt / (hp * d[K]), nudt / (hm * d[K])
        A[k, k] = 1.0 - cp - cm
        if k + 1 < nz:
            A[k, k + 1] = cp
        else:                       # q_g = 2 bct - q_k
            A[k, k] -= cp
            rhs[k] -= cp * 2 * bct
        if k > 0:
            A[k, k - 1] = cm
        else:
            A[k, k] -= cm
            rhs[k] -= cm * 2 * bcb
    ref = np.linalg.solve(A, rhs)
    out = np.zeros(g.shape)
    L.orc_solve1d_cell(dp(out), dp(h), g.ref, 2, nudt, bcb, bct, 0, 0)
    assert np.abs(out[3, 3, ng:ng + nz] - ref).max() <= 1e-13
    assert abs(0.5 * (out[3, 3, ng - 1] + out[3, 3, ng]) - bcb) <= 1e-14


# ------------------------------------------------------- projection

def random_projected(order, nx=8, ny=6, nz=10, seed=7):
    g = grid(nx=nx, ny=ny, nz=nz)
    rng = np.random.default_rng(seed)
    u, v, w = (rng.standard_normal(g.shape) for _ in range(3))
    for a in (u, v, w):
        L.orc_periodic(dp(a), g.ref)
    L.orc_apply_bc_face(dp(w), g.ref, order, 0.0, 0.0)
    d = g.zeros()
    L.orc_divergence(dp(u), dp(v), dp(w), dp(d), g.ref, order)
    spec = np.zeros(((nx // 2 + 1) * ny * g.g.n2 * 2,))
    L.orc_fft3d_fwd(dp(d), dp(spec), g.ref)
    L.orc_poisson(dp(spec), g.ref, order, 1.0)
    p = g.zeros()
    L.orc_fft3d_bwd(dp(spec), dp(p), g.ref)
    L.orc_periodic(dp(p), g.ref)
    L.orc_project(dp(u), dp(v), dp(w), dp(p), g.ref, order, 1.0)
    for a in (u, v, w):
        L.orc_periodic(dp(a), g.ref)
    return g, u, v, w, p, d


@pytest.mark.parametrize("order", [2, 4])
def test_projection_case1(order):
    """case1/case8 (cases.c:473-503): after the projection |div| <= 1e-10
    and the discrete Laplacian of p equals the original divergence."""
    g, u, v, w, p, d0 = random_projected(order)
    d = g.zeros()
    L.orc_apply_bc_face(dp(w), g.ref, order, 0.0, 0.0)
    L.orc_divergence(dp(u), dp(v), dp(w), dp(d), g.ref, order)
    assert np.abs(g.interior(d)).max() <= 1e-12
    lap = g.zeros()
    L.orc_pressure_laplacian(dp(p), dp(lap), g.ref, order, 1.0)
    assert np.abs(g.interior(lap) - g.interior(d0)).max() <= 1e-12


def dense_laplacian_solve(nx, ny, nz, dx, dy, dzn, rhs):
    """Assemble the 2nd-order discrete Laplacian with periodic x, y,
    Neumann walls and the Dirichlet-mean bottom closure on the mean mode
    (pressure.c:331-339), solve densely."""
    N = nx * ny * nz
    idx = lambda i, j, k: ((i % nx) * ny + (j % ny)) * nz + k
    A = np.zeros((N, N))
    dzp = np.concatenate([[dzn[0]], dzn, [dzn[-1]]])
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                r = idx(i, j, k)
                vol = dx * dy * dzn[k]
                for di, dj, cf in ((1, 0, vol / dx ** 2), (-1, 0, vol / dx ** 2), (0, 1, vol / dy ** 2), (0, -1, vol / dy ** 2)):
                    A[r, idx(i + di, j + dj, k)] += cf
                    A[r, r] -= cf
                K = k + 1
                hp, hm = 0.5 * (dzp[K + 1] + dzp[K]), 0.5 * (dzp[K - 1] + dzp[K])
                cp, cm = vol / (hp * dzn[k]), vol / (hm * dzn[k])
                if k + 1 < nz:
                    A[r, idx(i, j, k + 1)] += cp
                    A[r, r] -= cp
                if k > 0:
                    A[r, idx(i, j, k - 1)] += cm
                    A[r, r] -= cm
    b = (rhs * (dx * dy * dzn)[None, None, :]).ravel()
    # fix the gauge of the mean mode as the solver does: p_g = -p_1 on the
    # plane average at the bottom, i.e. add -2 cm * mean(p_1) to the mean
    P = np.zeros((N, N))
    for i in range(nx):
        for j in range(ny):
            for i2 in range(nx):
                for j2 in range(ny):
                    P[idx(i, j, 0), idx(i2, j2, 0)] = 1.0 / (nx * ny)
    hm0 = 0.5 * (dzp[0] + dzp[1])
    cm0 = dx * dy * dzn[0] / (hm0 * dzn[0])
    A2 = A - 2.0 * cm0 * P
    return np.linalg.solve(A2, b).reshape(nx, ny, nz)


def test_poisson_vs_dense_bruteforce():
    nx, ny, nz = 4, 4, 5
    dzn = tanh_dz(nz)
    g = grid(nx=nx, ny=ny, nz=nz, dz=dzn)
    rng = np.random.default_rng(9)
    rhs = rng.standard_normal((nx, ny, nz))
    rhs -= rhs.mean()                      # compatible with the Neumann walls
    d = g.zeros()
    ng = g.ng
    d[ng:ng + nx, ng:ng + ny, ng:ng + nz] = rhs
    spec = np.zeros(((nx // 2 + 1) * ny * g.g.n2 * 2,))
    L.orc_fft3d_fwd(dp(d), dp(spec), g.ref)
    L.orc_poisson(dp(spec), g.ref, 2, 1.0)
    p = g.zeros()
    L.orc_fft3d_bwd(dp(spec), dp(p), g.ref)
    ref = dense_laplacian_solve(nx, ny, nz, g.g.dx, g.g.dy, dzn, rhs)
    assert np.abs(g.interior(p) - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_laplacian_exact_on_cubic_uniform():
    """case0 closed form: on a uniform grid the 2nd-order z Laplacian of a
    cubic is exact (6z for z^3); x-y parts vanish for a z-only field."""
    nz = 16
    lz = 2.0
    dzn = np.full(nz, lz / nz)
    g = grid(nx=4, ny=4, nz=nz, dz=dzn)
    ng = g.ng
    zc = (np.arange(-ng, nz + ng) + 0.5) * lz / nz
    p = np.broadcast_to(zc ** 3, g.shape).copy()
    lap = g.zeros()
    L.orc_pressure_laplacian(dp(p), dp(lap), g.ref, 2, 1.0)
    assert np.abs(g.interior(lap) - 6 * zc[ng:ng + nz][None, None, :]).max() <= 1e-11


# -------------------------------------------------------- advection

@pytest.mark.parametrize("order", [2, 4])
def test_advection_conservation_case6(order):
    """case6/case13 (cases.c:1939-1997): for a projected field with walls,
    the convective tendencies conserve u and v momentum and the kinetic
    energy of every component (Sanderse 2014 eq. 78, cases.c:1989).  The w
    momentum is not checked: its wall half-cell fluxes are never computed
    (the reference's own note on eq. 76)."""
    g, u, v, w, p, _ = random_projected(order, nx=12, ny=10, nz=12, seed=13)
    for a in (u, v):
        L.orc_apply_bc_cell(dp(a), g.ref, order, 0.0, 0.0, 0, 0)
    L.orc_apply_bc_face(dp(w), g.ref, order, 0.0, 0.0)
    mx, my, mz = g.zeros(), g.zeros(), g.zeros()
    fu, fv, fw = g.zeros(), g.zeros(), g.zeros()
    L.orc_massflux(0, dp(u), dp(v), dp(w), dp(mx), dp(my), dp(mz), g.ref, order)
    L.orc_advect(0, dp(u), dp(mx), dp(my), dp(mz), dp(fu), g.ref, order)
    L.orc_massflux(1, dp(u), dp(v), dp(w), dp(mx), dp(my), dp(mz), g.ref, order)
    L.orc_advect(1, dp(v), dp(mx), dp(my), dp(mz), dp(fv), g.ref, order)
    L.orc_massflux(2, dp(u), dp(v), dp(w), dp(mx), dp(my), dp(mz), g.ref, order)
    L.orc_advect(2, dp(w), dp(mx), dp(my), dp(mz), dp(fw), g.ref, order)
    ng, nz = g.ng, g.g.nz
    d = g.dzbuf
    vol = g.g.dx * g.g.dy * d[ng:ng + nz]
    volw = g.g.dx * g.g.dy * 0.5 * (d[ng:ng + nz - 1] + d[ng + 1:ng + nz])
    I = lambda a: g.interior(a)
    scale = np.abs(I(u) * I(fu) * vol).sum()
    msum = [(I(fu) * vol).sum(), (I(fv) * vol).sum()]
    ksum = [(I(u) * I(fu) * vol).sum(), (I(v) * I(fv) * vol).sum(),
            (I(w)[:, :, :nz - 1] * I(fw)[:, :, :nz - 1] * volw).sum()]
    assert max(abs(m) for m in msum) <= 1e-13 * scale
    assert max(abs(k) for k in ksum) <= 1e-13 * scale


def test_uniform_flow_has_no_tendency():
    """Galilean check: a uniform u with walls moving at the same speed
    (uframe = -U) produces zero advective and diffusive tendency."""
    g = grid(nx=8, ny=6, nz=9)
    U = 0.37
    u = np.full(g.shape, U)
    v, w = g.zeros(), g.zeros()
    L.orc_apply_bc_cell(dp(u), g.ref, 2, U, U, 0, 0)
    mx, my, mz, fu = g.zeros(), g.zeros(), g.zeros(), g.zeros()
    L.orc_massflux(0, dp(u), dp(v), dp(w), dp(mx), dp(my), dp(mz), g.ref, 2)
    L.orc_advect(0, dp(u), dp(mx), dp(my), dp(mz), dp(fu), g.ref, 2)
    L.orc_laplacexy_cell(dp(u), dp(fu), g.ref, 2, 0.01)
    assert np.abs(g.interior(fu)).max() <= 1e-13


# ------------------------------------------------------- full step

def small_run(order=2, nx=16, ny=12, nz=17, steps=4, seed=1234):
    nzw = nz
    dz = tanh_dz(nzw)
    cfg = dict(nx=nx, ny=ny, nz=nz, order=order, scalar=0, lx=2 * np.pi, ly=np.pi,
               dt=1e-3, want_cfl=0.5, cflmode=0, nu=1 / 2800, prandtl=1.0, dpdx=0.0,
               want_uvolavg=1.0, uvolavgmode=1, want_cvolavg=0.0, cvolavgmode=0,
               uframe=0.0, vframe=0.0, betg_x=0.0, betg_y=0.0, betg_z=0.0, dz=dz)
    rng = np.random.default_rng(seed)
    zc = np.cumsum(dz) - 0.5 * dz
    eta = zc - 1.0
    o = ol.OracleRun(cfg)
    o.put("u", 1.5 * (1 - eta ** 2)[None, None, :] + 0.1 * rng.uniform(-1, 1, (nx, ny, nz)))
    o.put("v", 0.1 * rng.uniform(-1, 1, (nx, ny, nz)))
    wi = 0.1 * rng.uniform(-1, 1, (nx, ny, nz))
    wi[:, :, -1] = 0.0
    o.put("w", wi)
    o.project_ic(1.0)
    return o, cfg


def test_step_invariants_order2():
    o, cfg = small_run()
    log = o.steps(5)
    assert np.allclose(log[:, 6], 1.0, rtol=0, atol=1e-13)   # bulk velocity held at Ub
    assert o.max_div() <= 1e-13
    assert np.all(np.isfinite(log))


def test_step_restart_bit_identical():
    """RK3 is self-starting (zeta_0 = 0, rk3.c:6): a run restarted from the
    io fields after 2 steps matches 4 uninterrupted steps bit for bit."""
    a, cfg = small_run()
    a.steps(4)
    b, _ = small_run()
    b.steps(2)
    mid = {n: b.get(n) for n in "uvw"}
    c = ol.OracleRun(cfg)
    for n in "uvw":
        c.put(n, mid[n])
    c.steps(2)
    for n in "uvw":
        assert np.array_equal(a.get(n), c.get(n))


def general_run(order, scalar, nx=12, ny=12, nz=17, seed=4321, cvolavgmode=0, betg_z=0.0):
    dz = tanh_dz(nz)
    cfg = dict(nx=nx, ny=ny, nz=nz, order=order, scalar=scalar, lx=2 * np.pi, ly=np.pi,
               dt=5e-4, want_cfl=0.5, cflmode=0, nu=1 / 3300, prandtl=0.71, dpdx=0.0,
               want_uvolavg=1.0, uvolavgmode=1, want_cvolavg=0.1, cvolavgmode=cvolavgmode,
               uframe=0.0, vframe=0.0, betg_x=0.0, betg_y=0.0, betg_z=betg_z, dz=dz)
    rng = np.random.default_rng(seed)
    zc = np.cumsum(dz) - 0.5 * dz
    o = ol.OracleRun(cfg)
    o.put("u", 1.5 * (1 - (zc - 1.0) ** 2)[None, None, :] + 0.1 * rng.uniform(-1, 1, (nx, ny, nz)))
    o.put("v", 0.1 * rng.uniform(-1, 1, (nx, ny, nz)))
    wi = 0.1 * rng.uniform(-1, 1, (nx, ny, nz))
    wi[:, :, -1] = 0.0
    o.put("w", wi)
    if scalar:
        o.put("c", (zc / 2.0 - 0.5)[None, None, :] + 0.05 * rng.uniform(-1, 1, (nx, ny, nz)))
    o.project_ic(1.0)
    return o, cfg


@pytest.mark.parametrize("order,scalar", [(4, 0), (2, 1), (4, 1)])
def test_step_invariants_general(order, scalar):
    o, cfg = general_run(order, scalar)
    c0 = o.get("c") if scalar else None
    log = o.steps(4)
    assert np.all(np.isfinite(log))
    assert np.allclose(log[:, 6], 1.0, rtol=0, atol=1e-13)      # bulk velocity held at Ub
    assert o.max_div() <= 1e-12
    if scalar:
        c = o.get("c")
        # QUICK transport + diffusion of a field bounded by the wall values
        # stays within a few percent of its initial range
        assert c.min() >= c0.min() - 0.05 and c.max() <= c0.max() + 0.05


def test_scalar_mean_constraint_order4():
    """cvolavgmode = 1 (chanfast.c:572-617): the bulk scalar is held at
    the target every step."""
    o, cfg = general_run(4, 1, cvolavgmode=1)
    log = o.steps(3)
    assert np.allclose(log[:, 11], 0.1, rtol=0, atol=1e-13)


def test_step_restart_bit_identical_order4_scalar():
    a, cfg = general_run(4, 1)
    a.steps(4)
    b, _ = general_run(4, 1)
    b.steps(2)
    mid = {n: b.get(n) for n in "uvwc"}
    c = ol.OracleRun(cfg)
    for n in "uvwc":
        c.put(n, mid[n])
    c.steps(2)
    for n in "uvwc":
        assert np.array_equal(a.get(n), c.get(n)), n
