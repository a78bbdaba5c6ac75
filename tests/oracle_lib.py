"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / the timed CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# ORACLE_LIB_PATH selects another build of the same sources (e.g. one with a
# deliberately planted mistake, to check that the pinning tests catch it)
ORACLE_LIB = os.environ.get("ORACLE_LIB_PATH") or os.path.join(ORACLE_DIR, "liboracle.so")
LOG_N = 13


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ng", C.c_int),
                ("n0", C.c_int), ("n1", C.c_int), ("n2", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.POINTER(C.c_double))]


class Cfg(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ng", C.c_int),
                ("order", C.c_int), ("scalar", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double),
                ("dt", C.c_double), ("want_cfl", C.c_double), ("cflmode", C.c_int),
                ("nu", C.c_double), ("prandtl", C.c_double),
                ("dpdx", C.c_double), ("want_uvolavg", C.c_double), ("uvolavgmode", C.c_int),
                ("want_cvolavg", C.c_double), ("cvolavgmode", C.c_int),
                ("uframe", C.c_double), ("vframe", C.c_double),
                ("betg_x", C.c_double), ("betg_y", C.c_double), ("betg_z", C.c_double),
                ("driver", C.c_int)]


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        L = C.CDLL(ORACLE_LIB)
        dp = C.POINTER(C.c_double)
        P = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Cfg), dp, C.POINTER(P)]
        L.orc_destroy.argtypes = [P]
        L.orc_put.argtypes = [P, C.c_int, dp]
        L.orc_get.argtypes = [P, C.c_int, dp]
        L.orc_steps.argtypes = [P, C.c_int, dp]
        L.orc_project_ic.argtypes = [P, C.c_double]
        L.orc_max_div.argtypes = [P]
        L.orc_max_div.restype = C.c_double
        L.orc_run_grid.argtypes = [P]
        L.orc_run_grid.restype = C.POINTER(Grid)
        L.orc_run_dt.argtypes = [P]
        L.orc_run_dt.restype = C.c_double
        L.orc_tridag.argtypes = [dp] * 5 + [C.c_int]
        L.orc_heptdag.argtypes = [dp] * 9 + [C.c_int]
        L.orc_fft.argtypes = [dp, C.c_int, C.c_int]
        L.orc_rfft.argtypes = [dp, dp, C.c_int]
        L.orc_irfft.argtypes = [dp, dp, C.c_int]
        L.orc_distribute.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_partition.argtypes = [C.c_int, C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 6
        G = C.POINTER(Grid)
        for name, args in {
            "orc_grid_init": [G, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, dp, dp],
            "orc_periodic": [dp, G],
            "orc_apply_bc_cell": [dp, G, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int],
            "orc_apply_bc_face": [dp, G, C.c_int, C.c_double, C.c_double],
            "orc_helmholtz1d_cell": [dp, dp, G, C.c_int, C.c_double],
            "orc_helmholtz1d_face": [dp, dp, G, C.c_int, C.c_double],
            "orc_solve1d_cell": [dp, dp, G, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int],
            "orc_solve1d_face": [dp, dp, G, C.c_int, C.c_double, C.c_double, C.c_double],
            "orc_laplacexy_cell": [dp, dp, G, C.c_int, C.c_double],
            "orc_laplacexy_face": [dp, dp, G, C.c_int, C.c_double],
            "orc_massflux": [C.c_int, dp, dp, dp, dp, dp, dp, G, C.c_int],
            "orc_advect": [C.c_int, dp, dp, dp, dp, dp, G, C.c_int],
            "orc_divergence": [dp, dp, dp, dp, G, C.c_int],
            "orc_project": [dp, dp, dp, dp, G, C.c_int, C.c_double],
            "orc_pressure_laplacian": [dp, dp, G, C.c_int, C.c_double],
            "orc_poisson": [dp, G, C.c_int, C.c_double],
            "orc_helmholtz3d_cell": [dp, dp, G, C.c_int, C.c_double],
            "orc_helmholtz3d_face": [dp, dp, G, C.c_int, C.c_double],
            "orc_solve_helmholtz_cell": [dp, G, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int],
            "orc_solve_helmholtz_face": [dp, G, C.c_int, C.c_double, C.c_double, C.c_double],
            "orc_fft3d_fwd": [dp, dp, G],
            "orc_fft3d_bwd": [dp, dp, G],
            "orc_xy_avg": [dp, dp, G],
            "orc_flux_quick": [dp, dp, dp, dp, dp, dp, dp, G],
            "orc_advect_scalar": [dp, dp, dp, dp, G],
            "orc_xy_corr": [dp, dp, dp, G],
            "orc_xy_moment": [dp, dp, C.c_double, G],
            "orc_xy_spec": [dp, dp, dp, dp, C.c_int, C.c_int, C.c_int],
            "orc_add_buoyancy": [dp, dp, dp, dp, C.c_double, C.c_double, C.c_double, G, C.c_int],
        }.items():
            getattr(L, name).argtypes = args
        L.orc_cfl.argtypes = [dp, dp, dp, G, C.c_double]
        L.orc_cfl.restype = C.c_double
        L.orc_z_avg.argtypes = [dp, C.c_int, C.c_int, dp]
        L.orc_z_avg.restype = C.c_double
        _L = L
    return _L


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class PaddedGrid:
    """A padded [n0][n1][n2] grid (reference layout) for operator tests."""

    def __init__(self, nx, ny, nz, lx, ly, dz_interior, ng=3):
        self.dzbuf = np.zeros(nz + 2 * ng)
        self.g = Grid()
        dzi = np.ascontiguousarray(dz_interior, dtype=np.float64)
        lib().orc_grid_init(C.byref(self.g), nx, ny, nz, ng, lx, ly, dptr(self.dzbuf), dptr(dzi))
        self.shape = (self.g.n0, self.g.n1, self.g.n2)
        self.ng = ng

    def zeros(self):
        return np.zeros(self.shape)

    def interior(self, a):
        ng = self.ng
        return a[ng:-ng, ng:-ng, ng:-ng]

    @property
    def ref(self):
        return C.byref(self.g)


class OracleRun:
    """The oracle's chanfast2/4 time loop on one process."""

    def __init__(self, cfg, ng=3):
        """cfg: a chanflow.Config-like object (nx, ny, nz, order, ... , dz via dz_interior)."""
        from_dz = np.ascontiguousarray(cfg["dz"], dtype=np.float64)
        c = Cfg()
        for f, _ in Cfg._fields_:
            if f == "ng":
                c.ng = ng
            elif f == "driver":
                c.driver = int(cfg.get("driver", 0))     # 0 chanfast (default), 1 channel
            else:
                setattr(c, f, cfg[f])
        self.shape = (cfg["nx"], cfg["ny"], cfg["nz"])
        self._h = C.c_void_p()
        st = lib().orc_create(C.byref(c), dptr(from_dz), C.byref(self._h))
        if st != 0:
            raise RuntimeError(f"orc_create failed ({st})")

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def put(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape
        lib().orc_put(self._h, "uvwcp".index(name), dptr(a))

    def get(self, name):
        out = np.zeros(self.shape)
        lib().orc_get(self._h, "uvwcp".index(name), dptr(out))
        return out

    def steps(self, n):
        log = np.zeros((n, LOG_N))
        lib().orc_steps(self._h, n, dptr(log))
        return log

    def project_ic(self, dt=1.0):
        lib().orc_project_ic(self._h, dt)

    def max_div(self):
        return lib().orc_max_div(self._h)


def oracle_cfg(ccfg, dz):
    """Oracle parameter dict from a chanflow.Config and its dz."""
    keys = [f for f, _ in Cfg._fields_ if f != "ng"]
    d = {k: getattr(ccfg, k, 0) for k in keys}
    d["dz"] = dz
    return d
