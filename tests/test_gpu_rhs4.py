"""GPU parity of the order-4 explicit stage on column tiles (k_rhs4_uv and
k_rhs4_wc, taken when nx and ny are multiples of the 32 x 8 tile): against
the oracle to <= 1e-12 relative L2, on grids chosen to cross the kernels'
edges -- tiles whose halos wrap onto themselves (nx = 32, ny = 8), several
z-chunks per column (chunk boundaries inside the wall-ghost reach and in
the interior), the scalar with buoyancy on all three components, and the
fully implicit driver.
"""
import numpy as np
import pytest

import chanflow as cf
from test_gpu_general import general_case, run_and_compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,scalar,steps", [
    ((32, 8, 17), 0, 20),        # one tile, halos wrap onto the tile itself
    ((64, 16, 33), 1, 100),      # 2 x 2 tiles, several z-chunks
    ((32, 16, 65), 1, 20),       # more z-chunks, boundaries away from the walls
    ((32, 8, 13), 1, 20),        # chunks that hold a wall's whole ghost reach
])
def test_order4_tiles_match_oracle(shape, scalar, steps):
    kw = dict(betg_x=0.3, betg_y=-0.2, betg_z=0.5) if scalar else {}
    cfg, orc, init, f, names = general_case(*shape, 4, scalar, nu=1.0 / 2000.0, dt=5e-4, **kw)
    run_and_compare(cfg, orc, init, names, steps)


def test_order4_tiles_fully_implicit_driver():
    cfg, orc, init, _, names = general_case(64, 16, 33, 4, 1, driver=cf.CF_DRIVER_CHANNEL, dt=1e-3,
                                            nu=1.0 / 1000.0)
    run_and_compare(cfg, orc, init, names, 30)


def test_order4_tiles_and_per_field_kernels_agree():
    """The same flow on a grid the tiles divide (nx = 64) and on one they
    do not (nx = 48, the per-field kernels), both against the oracle, so a
    difference between the two launch paths cannot hide behind either."""
    for nx in (64, 48):
        cfg, orc, init, f, names = general_case(nx, 16, 33, 4, 1, nu=1.0 / 2000.0, dt=5e-4)
        run_and_compare(cfg, orc, init, names, 10)
