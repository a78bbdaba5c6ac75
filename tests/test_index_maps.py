"""The pencil index maps of the product (decomp.c:18-186 names and
arguments, exported by libchanflow.so): the known answer of decomp.c:103,
bit-exact agreement with the oracle's restatement, and exact tiling of the
global grid by every pencil family.  CPU only (host integer code)."""
import ctypes as C

import numpy as np
import pytest

import chanflow as cf
import oracle_lib as ol

IA = C.c_int * 3


class Plan(C.Structure):
    _fields_ = [(n, C.c_int * 3) for n in ("xst", "xen", "xsz", "yst", "yen", "ysz", "zst", "zen", "zsz")] + \
               [(n, C.POINTER(C.c_int)) for n in ("x1st", "x1en", "x1dist", "y1st", "y1en", "y1dist",
                                                  "y2st", "y2en", "y2dist", "z2st", "z2en", "z2dist")]


def L():
    lib = cf.lib()
    lib.create_decomp_plan.restype = C.POINTER(Plan)
    lib.create_decomp_plan.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int]
    lib.destroy_decomp_plan.argtypes = [C.POINTER(Plan)]
    return lib


def distribute(lib, n, p):
    st, en, sz = (C.c_int * p)(), (C.c_int * p)(), (C.c_int * p)()
    lib.distribute(n, p, st, en, sz)
    return list(st), list(en), list(sz)


def test_distribute_known_answer():
    """decomp.c:103: 17 points over 4 processors are (4, 4, 4, 5)."""
    st, en, sz = distribute(L(), 17, 4)
    assert sz == [4, 4, 4, 5] and st == [1, 5, 9, 13] and en == [4, 8, 12, 17]


@pytest.mark.parametrize("n,p", [(1, 1), (7, 3), (16, 4), (193, 8), (5, 5), (256, 6), (65, 7), (3, 3)])
def test_distribute_matches_oracle(n, p):
    lib = L()
    mine = distribute(lib, n, p)
    st, en, sz = (C.c_int * p)(), (C.c_int * p)(), (C.c_int * p)()
    ol.lib().orc_distribute(n, p, st, en, sz)
    assert mine == (list(st), list(en), list(sz))


@pytest.mark.parametrize("dims", [(1, 1), (2, 3), (4, 2), (3, 5)])
def test_plans_tile_the_grid(dims):
    """create_decomp_plan (decomp.c:121-168): for every rank the x, y and z
    pencils are the oracle's partitions (pdim {1,2,3}, {2,1,3}, {2,3,1}),
    and each pencil family covers the global grid exactly once."""
    lib = L()
    nx, ny, nz = 17, 13, 11
    cover = {q: np.zeros((nx, ny, nz), int) for q in "xyz"}
    for c0 in range(dims[0]):
        for c1 in range(dims[1]):
            d, c = (C.c_int * 2)(*dims), (C.c_int * 2)(c0, c1)
            p = lib.create_decomp_plan(d, c, nx, ny, nz).contents
            for q, pdim in (("x", (1, 2, 3)), ("y", (2, 1, 3)), ("z", (2, 3, 1))):
                st, en, sz = list(getattr(p, q + "st")), list(getattr(p, q + "en")), list(getattr(p, q + "sz"))
                ost, oen, osz = IA(), IA(), IA()
                ol.lib().orc_partition(nx, ny, nz, IA(*pdim), (C.c_int * 2)(*dims), (C.c_int * 2)(c0, c1), ost, oen, osz)
                assert (st, en, sz) == (list(ost), list(oen), list(osz))
                assert all(e - s + 1 == z for s, e, z in zip(st, en, sz))
                cover[q][st[0] - 1:en[0], st[1] - 1:en[1], st[2] - 1:en[2]] += 1
            assert [p.x1dist[r] for r in range(dims[0])] == distribute(lib, nx, dims[0])[2]
            assert [p.z2dist[r] for r in range(dims[1])] == distribute(lib, nz, dims[1])[2]
            lib.destroy_decomp_plan(C.pointer(p))
    for q in "xyz":
        assert (cover[q] == 1).all(), q
