"""GPU parity of the fully implicit driver (channel2 / channel4,
channel.c:71-1968): the x-y diffusion moves from the explicit tendency into
the 3-D Helmholtz operator of the right-hand side (helmholtz_operator_u/w,
diffuseu.c:362-457, diffusew.c:366-465), and every field's implicit update
is a spectral solve -- FFT3D_FWD, per (kx, ky) the wall-reduced banded z
system with the modified wavenumbers on its diagonal
(solve_helmholtz_u/w, diffuseu.c:460-614, diffusew.c:466-622), FFT3D_BWD.

FP64 fields to <= 1e-12 relative L2 against the CPU oracle's channel
driver, per-step logs as for chanfast, divergence at round-off.
"""
import numpy as np
import pytest

import chanflow as cf
from test_gpu_general import general_case, run_and_compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order,scalar,shape,steps", [
    (2, 0, (32, 24, 33), 50),
    (2, 1, (24, 16, 25), 30),
    (4, 0, (24, 20, 25), 30),
    (4, 1, (32, 24, 33), 50),
])
def test_channel_driver_matches_oracle(order, scalar, shape, steps):
    cfg, orc, init, _, names = general_case(*shape, order, scalar, driver=cf.CF_DRIVER_CHANNEL, dt=1e-3,
                                            nu=1.0 / 1000.0)
    run_and_compare(cfg, orc, init, names, steps)


def test_channel_driver_with_adaptive_dt():
    """channel4 with the reference's main.c settings (cflmode 1, constant
    bulk u in the moving frame, constant bulk c), 40 steps."""
    cfg, orc, init, _, names = general_case(24, 16, 25, 4, 1, driver=cf.CF_DRIVER_CHANNEL, dt=2e-3, cflmode=1,
                                            want_cfl=1.0, nu=1.0 / 180.0, prandtl=1.0, dpdx=-1.0,
                                            want_uvolavg=1.0, uvolavgmode=1, want_cvolavg=0.0, cvolavgmode=1,
                                            uframe=1.0, stretch="cosine")
    run_and_compare(cfg, orc, init, names, 40)


def test_channel_driver_differs_from_chanfast():
    """The two drivers are different schemes: the same run stepped by each
    differs well above round-off (the check above is not vacuous)."""
    out = []
    for drv in (cf.CF_DRIVER_CHANFAST, cf.CF_DRIVER_CHANNEL):
        cfg = cf.Config(nx=24, ny=16, nz=25, order=2, stretch="tanh", dt=2e-3, nu=1.0 / 100.0, driver=drv)
        f = cf.initial_fields(cfg, seed=3)
        with cf.Run(cfg) as r:
            for n in "uvw":
                r.upload(n, f[n])
            r.project(1.0)
            r.step(10)
            r.sync()
            out.append(r.download("u"))
    d = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[1])
    assert 1e-9 < d < 1e-2


def test_channel_driver_rejected_in_batches_and_fp32():
    cfg = cf.Config(nx=16, ny=12, nz=17, driver=cf.CF_DRIVER_CHANNEL)
    with pytest.raises(cf.ChanflowError):
        cf.Batch([cfg, cfg])
    with pytest.raises(cf.ChanflowError):
        cf.Run(cf.Config(nx=16, ny=12, nz=17, driver=cf.CF_DRIVER_CHANNEL, precision=cf.CF_FP32))
