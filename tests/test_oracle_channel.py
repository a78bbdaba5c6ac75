"""Pins the oracle's fully implicit driver (channel2 / channel4,
channel.c:71-1968) against what the mathematics fixes:

- the 3-D Helmholtz operators (diffuseu.c:362-457, diffusew.c:366-465)
  equal the already-pinned z operator plus the x-y Laplacian;
- the spectral Helmholtz solves (diffuseu.c:460-614, diffusew.c:466-622)
  invert the 3-D operator exactly on fields that meet the wall conditions
  (the modified wavenumbers diagonalise the discrete x-y second
  differences), for homogeneous walls and for the inhomogeneous real part
  of mode (0, 0);
- the RK3 step with every diffusion term Crank-Nicolson: a spanwise
  Fourier mode times a z eigenvector of the discrete operator decays by
  prod_s (1 + nu alpha_s dt L) / (1 - nu beta_s dt L), L = mu_z - lam / dy^2;
- bulk velocity held, divergence at round-off.

CPU only.
"""
import numpy as np
import pytest

import oracle_lib as ol
from test_oracle_math import RK_A, RK_B, grid, lam_xy, operator_z, rk_cfg, tanh_dz

L = ol.lib()
dp = ol.dptr


def random_field(g, seed, kind, order, bcb=0.0, bct=0.0):
    rng = np.random.default_rng(seed)
    q = g.zeros()
    ng = g.ng
    q[ng:-ng, ng:-ng, ng:-ng] = rng.uniform(-1, 1, (g.g.nx, g.g.ny, g.g.nz))
    if kind == "face":
        q[:, :, g.g.n2 - ng - 1] = 0.0
    L.orc_periodic(dp(q), g.ref)
    if kind == "cell":
        L.orc_apply_bc_cell(dp(q), g.ref, order, bcb, bct, 0, 0)
    else:
        L.orc_apply_bc_face(dp(q), g.ref, order, bcb, bct)
    return q


@pytest.mark.parametrize("kind", ["cell", "face"])
@pytest.mark.parametrize("order", [2, 4])
def test_helmholtz3d_is_z_operator_plus_xy_laplacian(kind, order):
    g = grid(nx=10, ny=8, nz=13)
    q = random_field(g, 1, kind, order)
    nudt = -0.37
    h3, hz, hxy = g.zeros(), g.zeros(), g.zeros()
    (L.orc_helmholtz3d_cell if kind == "cell" else L.orc_helmholtz3d_face)(dp(q), dp(h3), g.ref, order, nudt)
    (L.orc_helmholtz1d_cell if kind == "cell" else L.orc_helmholtz1d_face)(dp(q), dp(hz), g.ref, order, nudt)
    (L.orc_laplacexy_cell if kind == "cell" else L.orc_laplacexy_face)(dp(q), dp(hxy), g.ref, order, nudt)
    ng = g.ng
    kt = g.g.n2 - ng - (1 if kind == "face" else 0)
    sl = (slice(ng, -ng), slice(ng, -ng), slice(ng, kt))
    assert np.abs(h3[sl] - (hz[sl] + hxy[sl])).max() <= 1e-12 * np.abs(h3[sl]).max()


@pytest.mark.parametrize("kind", ["cell", "face"])
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("bc", [(0.0, 0.0), (0.7, -0.4)])
def test_spectral_solve_inverts_the_3d_operator(kind, order, bc):
    """q meeting the walls (q = bc there) -> h = q + nudt lap q (3-D) ->
    FFT3D_FWD -> spectral solve(nudt, bc) -> FFT3D_BWD gives q back."""
    if kind == "face" and bc != (0.0, 0.0):
        bc = (0.0, 0.0)          # w walls are 0 in every driver; the face solve is homogeneous there
    nx, ny, nz = 12, 10, 15
    g = grid(nx=nx, ny=ny, nz=nz)
    q = random_field(g, 3, kind, order, *bc)
    nudt = -0.21
    h = g.zeros()
    (L.orc_helmholtz3d_cell if kind == "cell" else L.orc_helmholtz3d_face)(dp(q), dp(h), g.ref, order, nudt)
    n2 = g.g.n2
    spec = np.zeros((nx // 2 + 1, ny, n2, 2))
    L.orc_fft3d_fwd(dp(h), dp(spec), g.ref)
    if kind == "cell":
        L.orc_solve_helmholtz_cell(dp(spec), g.ref, order, nudt, bc[0], bc[1], 0, 0)
    else:
        L.orc_solve_helmholtz_face(dp(spec), g.ref, order, nudt, bc[0], bc[1])
    back = g.zeros()
    L.orc_fft3d_bwd(dp(spec), dp(back), g.ref)
    ng = g.ng
    kt = n2 - ng - (1 if kind == "face" else 0)
    sl = (slice(ng, -ng), slice(ng, -ng), slice(ng, kt))
    assert np.abs(back[sl] - q[sl]).max() <= 1e-12
    # the 1-D z solve (no horizontal term) does not invert the 3-D operator,
    # so the check above sees the horizontal terms
    if kind == "cell" and bc == (0.0, 0.0):
        hz = g.zeros()
        L.orc_solve1d_cell(dp(hz), dp(h), g.ref, order, nudt, 0.0, 0.0, 0, 0)
        assert np.abs(hz[sl] - q[sl]).max() > 1e-3


@pytest.mark.parametrize("order", [2, 4])
def test_channel_rk3_spanwise_mode_decays_by_the_crank_nicolson_factor(order):
    nx, ny, nz = 6, 12, 15
    dz = tanh_dz(nz)
    cfg = rk_cfg(nx, ny, nz, order, dz, driver=1)
    g = grid(nx=nx, ny=ny, nz=nz, dz=dz)
    Lz = operator_z(g, "cell", order)
    ev, V = np.linalg.eig(Lz)
    real = np.where((np.abs(ev.imag) < 1e-9 * np.abs(ev).max()) & (ev.real < 0))[0]
    pick = real[np.argsort(-ev.real[real])][2]
    mu = ev.real[pick]
    phi = V[:, pick].real
    phi /= np.abs(phi).max()
    theta = 2 * np.pi / ny
    u0 = np.sin(theta * np.arange(ny))[None, :, None] * phi[None, None, :] + np.zeros((nx, ny, nz))
    o = ol.OracleRun(cfg)
    o.put("u", u0)
    o.put("v", np.zeros((nx, ny, nz)))
    o.put("w", np.zeros((nx, ny, nz)))
    nsteps = 4
    o.steps(nsteps)
    nu, dt = cfg["nu"], cfg["dt"]
    Lall = mu - lam_xy(order, theta) / g.g.dy ** 2
    a = 1.0
    for _ in range(nsteps):
        for s in range(3):
            a *= (1 + nu * RK_A[s] * dt * Lall) / (1 - nu * RK_B[s] * dt * Lall)
    got = o.get("u")
    assert abs(a - 1.0) > 1e-2
    assert np.abs(got - a * u0).max() <= 1e-13
    # FFT round-off in the fully implicit solves leaves v, w at round-off, not exact zeros
    assert np.abs(o.get("v")).max() <= 1e-14 and np.abs(o.get("w")).max() <= 1e-14


@pytest.mark.parametrize("order,scalar", [(2, 0), (4, 1)])
def test_channel_step_holds_bulk_and_divergence(order, scalar):
    nx, ny, nz = 16, 12, 17
    dz = tanh_dz(nz)
    cfg = rk_cfg(nx, ny, nz, order, dz, driver=1, nu=1 / 300, dt=2e-3, uvolavgmode=1, want_uvolavg=1.0,
                 scalar=scalar, prandtl=0.71, cvolavgmode=scalar, want_cvolavg=0.1)
    rng = np.random.default_rng(5)
    zc = np.cumsum(dz) - 0.5 * dz
    o = ol.OracleRun(cfg)
    o.put("u", 1.5 * (1 - (zc - 1) ** 2)[None, None, :] + 0.1 * rng.uniform(-1, 1, (nx, ny, nz)))
    o.put("v", 0.1 * rng.uniform(-1, 1, (nx, ny, nz)))
    w = 0.1 * rng.uniform(-1, 1, (nx, ny, nz))
    w[:, :, -1] = 0
    o.put("w", w)
    if scalar:
        o.put("c", (zc / 2 - 0.5)[None, None, :] + 0.05 * rng.uniform(-1, 1, (nx, ny, nz)))
    o.project_ic(1.0)
    log = o.steps(6)
    assert np.allclose(log[:, 6], 1.0, rtol=0, atol=1e-13)
    if scalar:
        assert np.allclose(log[:, 11], 0.1, rtol=0, atol=1e-13)
    assert o.max_div() <= 1e-12
