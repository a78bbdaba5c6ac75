"""chanflow -- Python mirror of the chanflow C ABI (include/chanflow.h).

The product is the C-ABI shared library ``libchanflow.so`` (HIP kernels for
gfx950 plus host C/C++).  This module binds it with ctypes for tests, the
benchmark and interactive use.  It never falls back to anything else: if
the library is missing or a GPU call fails, it raises.

Axis names follow the reference solver: x streamwise, y spanwise, z
wall-normal.  Fields cross the boundary in the reference's restart-file
layout (io.c:211-296): float64 C-order [nx][ny][nz].
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field, fields

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libchanflow.so")

CF_OK, CF_ERR_ARG, CF_ERR_UNSUPPORTED, CF_ERR_NOMEM, CF_ERR_DEVICE, CF_ERR_NUMERIC, CF_ERR_STATE = range(7)
CF_FP64, CF_FP32 = 0, 1
CF_DRIVER_CHANFAST, CF_DRIVER_CHANNEL = 0, 1
STRETCH = {"cosine": 0, "tanh": 1, "uniform": 2, "user": 3}
FIELDS = {"u": 0, "v": 1, "w": 2, "c": 3, "p": 4}
LOG_N = 13
LOG_COLUMNS = ("step", "dt", "cfl", "mdpdx", "nu_dudz_bot", "nu_dudz_top", "ubulk", "ucl",
               "dcdt", "kap_dcdz_bot", "kap_dcdz_top", "cbulk", "ccl")


class ChanflowError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"chanflow status {status}: {msg}")
        self.status = status


class _CConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("order", C.c_int), ("scalar", C.c_int), ("precision", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("stretch", C.c_int), ("stretch_beta", C.c_double), ("dz_user", C.POINTER(C.c_double)),
        ("dt", C.c_double), ("want_cfl", C.c_double), ("cflmode", C.c_int),
        ("nu", C.c_double), ("prandtl", C.c_double),
        ("dpdx", C.c_double), ("want_uvolavg", C.c_double), ("uvolavgmode", C.c_int),
        ("want_cvolavg", C.c_double), ("cvolavgmode", C.c_int),
        ("uframe", C.c_double), ("vframe", C.c_double),
        ("betg_x", C.c_double), ("betg_y", C.c_double), ("betg_z", C.c_double),
        ("device", C.c_int),
        ("driver", C.c_int),
    ]


@dataclass
class Config:
    """Run configuration; mirrors cf_config (chanfast2/4 parameters)."""
    nx: int = 64
    ny: int = 64
    nz: int = 65
    order: int = 2
    scalar: int = 0
    precision: int = CF_FP64
    lx: float = 2 * np.pi
    ly: float = np.pi
    lz: float = 2.0
    stretch: str = "cosine"
    stretch_beta: float = 2.0
    dz_user: object = None
    dt: float = 1e-3
    want_cfl: float = 0.5
    cflmode: int = 0
    nu: float = 1.0 / 2800.0
    prandtl: float = 1.0
    dpdx: float = 0.0
    want_uvolavg: float = 1.0
    uvolavgmode: int = 1
    want_cvolavg: float = 0.0
    cvolavgmode: int = 0
    uframe: float = 0.0
    vframe: float = 0.0
    betg_x: float = 0.0
    betg_y: float = 0.0
    betg_z: float = 0.0
    device: int = 0
    driver: int = 0          # 0 chanfast (x-y diffusion explicit), 1 channel (all diffusion implicit)

    def to_c(self):
        c = _CConfig()
        keep = None
        for f in fields(self):
            val = getattr(self, f.name)
            if f.name == "stretch":
                c.stretch = STRETCH[val] if isinstance(val, str) else int(val)
            elif f.name == "dz_user":
                if val is not None:
                    keep = np.ascontiguousarray(val, dtype=np.float64)
                    c.dz_user = keep.ctypes.data_as(C.POINTER(C.c_double))
            else:
                setattr(c, f.name, val)
        return c, keep

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)


_lib = None


def lib():
    """Load libchanflow.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"chanflow: {LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.cf_last_error.restype = C.c_char_p
    L.cf_version.restype = C.c_char_p
    sig = {
        "cf_config_default": (None, [C.POINTER(_CConfig)]),
        "cf_grid_dz": (C.c_int, [C.POINTER(_CConfig), dp]),
        "cf_grid_z": (C.c_int, [C.POINTER(_CConfig), C.c_int, dp, dp]),
        "cf_io_write_ascii": (C.c_int, [C.c_char_p, dp, C.c_int]),
        "cf_initial_fields": (C.c_int, [C.POINTER(_CConfig), C.c_uint64, C.c_double, C.c_double, dp, dp, dp, dp]),
        "cf_io_write": (C.c_int, [C.c_char_p, dp, C.c_size_t]),
        "cf_io_read": (C.c_int, [C.c_char_p, dp, C.c_size_t]),
        "cf_io_append_ascii": (C.c_int, [C.c_char_p, dp, C.c_int]),
        "cf_run_create": (C.c_int, [C.POINTER(_CConfig), C.POINTER(P)]),
        "cf_run_destroy": (C.c_int, [P]),
        "cf_run_upload": (C.c_int, [P, C.c_int, dp]),
        "cf_run_download": (C.c_int, [P, C.c_int, dp]),
        "cf_run_project": (C.c_int, [P, C.c_double]),
        "cf_run_step": (C.c_int, [P, C.c_int]),
        "cf_run_sync": (C.c_int, [P]),
        "cf_run_log": (C.c_int, [P, dp, C.c_int, C.POINTER(C.c_int)]),
        "cf_run_max_divergence": (C.c_int, [P, dp]),
        "cf_run_info": (C.c_int, [P, dp, C.POINTER(C.c_int)]),
        "cf_run_footprint": (C.c_int, [C.POINTER(_CConfig), C.POINTER(C.c_size_t)]),
        "cf_run_set_profiling": (C.c_int, [P, C.c_int]),
        "cf_run_kernel_time": (C.c_int, [P, C.c_char_p, dp, C.POINTER(C.c_int)]),
        "cf_run_xy_stats": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_double, dp]),
        "cf_op_transpose_roundtrip": (C.c_int, [P, dp, dp, dp]),
        "cf_op_fft3d_fwd": (C.c_int, [P, dp, dp]),
        "cf_op_fft3d_bwd": (C.c_int, [P, dp, dp]),
        "cf_op_poisson": (C.c_int, [P, dp, C.c_double]),
        "cf_batch_create": (C.c_int, [C.POINTER(_CConfig), C.c_int, C.POINTER(P)]),
        "cf_batch_destroy": (C.c_int, [P]),
        "cf_batch_upload": (C.c_int, [P, C.c_int, C.c_int, dp]),
        "cf_batch_download": (C.c_int, [P, C.c_int, C.c_int, dp]),
        "cf_batch_project": (C.c_int, [P, C.c_double]),
        "cf_batch_step": (C.c_int, [P, C.c_int]),
        "cf_batch_sync": (C.c_int, [P]),
        "cf_batch_log": (C.c_int, [P, C.c_int, dp, C.c_int, C.POINTER(C.c_int)]),
        "cf_batch_max_divergence": (C.c_int, [P, C.c_int, dp]),
        "cf_batch_set_profiling": (C.c_int, [P, C.c_int]),
        "cf_batch_kernel_time": (C.c_int, [P, C.c_char_p, dp, C.POINTER(C.c_int)]),
        "cf_batch_tuning": (C.c_int, [P, C.c_char_p, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(st):
    if st != CF_OK:
        raise ChanflowError(st, lib().cf_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def version():
    return lib().cf_version().decode()


def grid_dz(cfg: Config) -> np.ndarray:
    c, keep = cfg.to_c()
    out = np.zeros(cfg.nz)
    _check(lib().cf_grid_dz(C.byref(c), _dptr(out)))
    return out


def grid_z(cfg: Config, nghost=3):
    """zp, zw of get_z_at_p / get_z_at_w (get.c:458-488): cell-centre and
    face heights of the padded grid (nz + 2 nghost levels)."""
    c, keep = cfg.to_c()
    n2 = cfg.nz + 2 * nghost
    zp, zw = np.zeros(n2), np.zeros(n2)
    _check(lib().cf_grid_z(C.byref(c), int(nghost), _dptr(zp), _dptr(zw)))
    del keep
    return zp, zw


def initial_fields(cfg: Config, seed=1234, amp=0.1, camp=0.05):
    """Seeded laminar + noise fields (unprojected) in io layout."""
    c, keep = cfg.to_c()
    out = [np.zeros(cfg.shape) for _ in range(4)]
    _check(lib().cf_initial_fields(C.byref(c), C.c_uint64(seed), amp, camp, *[_dptr(a) for a in out]))
    return dict(zip("uvwc", out))


def footprint(cfg: Config) -> int:
    c, keep = cfg.to_c()
    n = C.c_size_t(0)
    _check(lib().cf_run_footprint(C.byref(c), C.byref(n)))
    return n.value


def io_write(path, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    _check(lib().cf_io_write(path.encode(), _dptr(a), a.size))


def io_read(path, shape):
    a = np.zeros(shape)
    _check(lib().cf_io_read(path.encode(), _dptr(a), a.size))
    return a


class Run:
    """One device-resident channel run (a cf_run handle)."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        c, self._keep = cfg.to_c()
        h = C.c_void_p()
        _check(lib().cf_run_create(C.byref(c), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().cf_run_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.cfg.shape:
            raise ValueError(f"field {name} has shape {a.shape}, expected {self.cfg.shape}")
        _check(lib().cf_run_upload(self._h, FIELDS[name], _dptr(a)))

    def download(self, name):
        out = np.zeros(self.cfg.shape)
        _check(lib().cf_run_download(self._h, FIELDS[name], _dptr(out)))
        return out

    def project(self, dt=1.0):
        _check(lib().cf_run_project(self._h, dt))

    def step(self, n=1):
        _check(lib().cf_run_step(self._h, int(n)))

    def sync(self):
        _check(lib().cf_run_sync(self._h))

    def log(self, max_steps=1 << 20):
        n = C.c_int(0)
        _check(lib().cf_run_log(self._h, None, max_steps, C.byref(n)))
        out = np.zeros((max(n.value, 1), LOG_N))
        _check(lib().cf_run_log(self._h, _dptr(out), n.value, C.byref(n)))
        return out[: n.value]

    def max_divergence(self):
        x = C.c_double(0)
        _check(lib().cf_run_max_divergence(self._h, C.byref(x)))
        return x.value

    def info(self):
        dt = C.c_double(0)
        n = C.c_int(0)
        _check(lib().cf_run_info(self._h, C.byref(dt), C.byref(n)))
        return dt.value, n.value

    def set_profiling(self, on=True):
        _check(lib().cf_run_set_profiling(self._h, 1 if on else 0))

    def kernel_time(self, name):
        ms = C.c_double(0)
        n = C.c_int(0)
        _check(lib().cf_run_kernel_time(self._h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    STATS = {"avg": 0, "corr": 1, "moment": 2, "spec": 3}

    def xy_stats(self, kind, a, b=None, en=0.0):
        """Plane statistics of the current state at the interior levels
        (cf_run_xy_stats): "avg" / "moment" (power en) of field a, "corr" of a
        and b -> array [nz]; "spec" of a and b -> (specx [nx/2+1][nz],
        specy [ny/2+1][nz]) complex."""
        k = self.STATS[kind]
        fa, fb = FIELDS[a], FIELDS[b if b is not None else a]
        nz = self.cfg.nz
        if k == 3:
            nsx, nsy = (self.cfg.nx // 2 + 1) * nz * 2, (self.cfg.ny // 2 + 1) * nz * 2
            out = np.zeros(nsx + nsy)
            _check(lib().cf_run_xy_stats(self._h, k, fa, fb, float(en), _dptr(out)))
            sx = out[:nsx].reshape(-1, nz, 2)
            sy = out[nsx:].reshape(-1, nz, 2)
            return sx[..., 0] + 1j * sx[..., 1], sy[..., 0] + 1j * sy[..., 1]
        out = np.zeros(nz)
        _check(lib().cf_run_xy_stats(self._h, k, fa, fb, float(en), _dptr(out)))
        return out

    # operator-level entry points
    def transpose_roundtrip(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        dev = np.zeros(self.cfg.nx * self.cfg.ny * self.cfg.nz)
        back = np.zeros(self.cfg.shape)
        _check(lib().cf_op_transpose_roundtrip(self._h, _dptr(a), _dptr(dev), _dptr(back)))
        return dev.reshape(self.cfg.nz, self.cfg.ny, self.cfg.nx), back

    def fft3d_fwd(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        out = np.zeros((self.cfg.nx // 2 + 1, self.cfg.ny, self.cfg.nz, 2))
        _check(lib().cf_op_fft3d_fwd(self._h, _dptr(a), _dptr(out)))
        return out[..., 0] + 1j * out[..., 1]

    def fft3d_bwd(self, spec):
        s = np.ascontiguousarray(np.stack([spec.real, spec.imag], -1), dtype=np.float64)
        out = np.zeros(self.cfg.shape)
        _check(lib().cf_op_fft3d_bwd(self._h, _dptr(s), _dptr(out)))
        return out

    def poisson(self, spec, dt):
        s = np.ascontiguousarray(np.stack([spec.real, spec.imag], -1), dtype=np.float64)
        _check(lib().cf_op_poisson(self._h, _dptr(s), dt))
        return s[..., 0] + 1j * s[..., 1]


class Batch:
    """Independent runs of one order / scalar setting / precision stepped
    together (a cf_batch handle): every substep stage is one persistent
    launch over all runs; launch shapes are autotuned at creation."""

    def __init__(self, cfgs):
        self.cfgs = list(cfgs)
        arr = (_CConfig * len(self.cfgs))()
        self._keep = []
        for q, cfg in enumerate(self.cfgs):
            c, keep = cfg.to_c()
            arr[q] = c
            self._keep.append(keep)
        h = C.c_void_p()
        _check(lib().cf_batch_create(arr, len(self.cfgs), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().cf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.cfgs)

    def upload(self, run, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.cfgs[run].shape:
            raise ValueError(f"field {name} has shape {a.shape}, expected {self.cfgs[run].shape}")
        _check(lib().cf_batch_upload(self._h, run, FIELDS[name], _dptr(a)))

    def download(self, run, name):
        out = np.zeros(self.cfgs[run].shape)
        _check(lib().cf_batch_download(self._h, run, FIELDS[name], _dptr(out)))
        return out

    def project(self, dt=1.0):
        _check(lib().cf_batch_project(self._h, dt))

    def step(self, n=1):
        _check(lib().cf_batch_step(self._h, int(n)))

    def sync(self):
        _check(lib().cf_batch_sync(self._h))

    def log(self, run, max_steps=1 << 20):
        n = C.c_int(0)
        _check(lib().cf_batch_log(self._h, run, None, max_steps, C.byref(n)))
        out = np.zeros((max(n.value, 1), LOG_N))
        _check(lib().cf_batch_log(self._h, run, _dptr(out), n.value, C.byref(n)))
        return out[: n.value]

    def max_divergence(self, run):
        x = C.c_double(0)
        _check(lib().cf_batch_max_divergence(self._h, run, C.byref(x)))
        return x.value

    def set_profiling(self, on=True):
        _check(lib().cf_batch_set_profiling(self._h, 1 if on else 0))

    def kernel_time(self, name):
        ms = C.c_double(0)
        n = C.c_int(0)
        _check(lib().cf_batch_kernel_time(self._h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def tuning(self):
        buf = C.create_string_buffer(16384)
        _check(lib().cf_batch_tuning(self._h, buf, len(buf)))
        return buf.value.decode().splitlines()
