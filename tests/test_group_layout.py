"""The grouped symmetric SpMV's host layout (opendeformer_amd/csrc/od_groups.cpp), CPU only.

The layout stores each upper block of a group's rows once; a block between two rows of the
same group feeds the column row through an LDS deposit, and the group's crossing list brings
the lower blocks whose upper owner lies outside.  tests/cpp/group_shim.cpp replays both device
products on the host with the kernels' rounding (fused multiply-adds, the 16-lane slot tree):
the grouped product must equal the slot product bit for bit, every deposit must be written
exactly once, every halo (the outside rows whose x a group stages in LDS) must hold distinct
outside rows in order, and both must match the oracle's SpMDV (core/numerMethod.h:134-262) and a dense
product.  The 4-colour tiled numbering must make most upper blocks internal (the property the
product's bytes rest on), and meshes with rows of more than 16 blocks must be refused.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from opendeformer_amd import mesh as M
from opendeformer_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def gs(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("gs") / "libgs.so")
    csrc = os.path.join(ROOT, "opendeformer_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "cpp", "group_shim.cpp"), os.path.join(csrc, "od_setup.cpp"),
                    os.path.join(csrc, "od_groups.cpp"), "-o", out], check=True)
    L = C.CDLL(out)
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    L.gs_build.argtypes = [C.c_int, dp, C.c_int, ip, C.c_int, ip, C.POINTER(C.c_longlong)]
    L.gs_error.restype = C.c_char_p
    L.gs_products.argtypes = [dp, dp, dp, dp]
    L.gs_products.restype = C.c_longlong
    L.gs_chunk_groups.argtypes = [ip]
    return L


def build(gs, w):
    X = np.ascontiguousarray(w.X, np.float64)
    T = np.ascontiguousarray(w.tets, np.int32)
    con = np.ascontiguousarray(np.sort(w.fixed_dofs), np.int32)
    st = np.zeros(8, np.int64)
    rc = gs.gs_build(len(X), O._d(X), len(T), O._i(T), len(con), O._i(con), st.ctypes.data_as(C.POINTER(C.c_longlong)))
    assert rc == 0, gs.gs_error().decode()
    keys = ("ok", "groups", "max_deps", "n_up", "n_x", "internal", "crossing", "rows")
    return dict(zip(keys, (int(v) for v in st)))


def products(gs, st, nvals, seed):
    rng = np.random.default_rng(seed)
    vals = np.ascontiguousarray(rng.standard_normal(nvals))
    x = np.ascontiguousarray(rng.standard_normal(3 * st["rows"]))
    ys, yg = np.zeros_like(x), np.zeros_like(x)
    bad = gs.gs_products(O._d(vals), O._d(x), O._d(ys), O._d(yg))
    return vals, x, ys, yg, bad


CASES = [("small_natural", lambda: W.box_small("natural")),
         ("small_color4t", lambda: W.box_small("color4t")),
         ("mid_color4", lambda: W.box_mid_neo("color4")),
         ("mid_color4t", lambda: W.box_mid_neo("color4t")),
         ("large_half_color4t", lambda: W.box_1m_dof("color4t", scale=0.5)),
         ("batch8_color4t", lambda: W.batch(8, seed=4, tet_lo=500, tet_hi=20000, ordering="color4t"))]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_grouped_product_equals_slot_product_bitwise(gs, name, make):
    w = make()
    st = build(gs, w)
    assert st["ok"] == 1
    nvals = gs.gs_nvalues()
    vals, x, ys, yg, bad = products(gs, st, nvals, seed=7)
    assert bad == 0, f"{bad} deposits not written exactly once or halo slots out of place"
    assert np.array_equal(ys.view(np.uint64), yg.view(np.uint64))
    if st["rows"] <= 30000:
        s = O.System(w.X, w.tets, w.fixed_dofs)
        yo = s.spmdv(vals, x)
        np.testing.assert_allclose(ys, yo, rtol=0, atol=1e-12 * np.abs(yo).max())
        if st["rows"] <= 1000:
            D = s.dense(vals)
            np.testing.assert_allclose(ys, D @ x, rtol=0, atol=1e-12 * np.abs(yo).max())


def test_tiled_colour_numbering_keeps_most_blocks_internal(gs):
    """A full 4x8x8 tile keeps about 3/4 of its rows' upper blocks inside its group (per edge
    direction 1 - 1/4 along i and 1 - 1/8 along j and k); the boundary layers of partial tiles
    keep fewer.  At half the 1M-DOF box (21 nodes across: 2 full tiles and a 5-node layer) the
    tiled 4-colour numbering keeps 64% inside (73% at full size), so the grouped product stores
    each off-diagonal block about 1.3 times rather than twice; the untiled 4-colour numbering,
    whose chunks are strips, keeps few inside."""
    st = build(gs, W.box_1m_dof("color4t", scale=0.5))
    frac = st["internal"] / (st["internal"] + st["crossing"])
    assert frac > 0.6, frac
    st4 = build(gs, W.box_1m_dof("color4", scale=0.5))
    frac4 = st4["internal"] / (st4["internal"] + st4["crossing"])
    assert frac4 < frac - 0.2, (frac4, frac)


def test_tiled_colour_numbering_is_proper_and_colour_grouped():
    m = M.box_mesh(24, 9, 10, 1.0, 0.4, 0.4, seed=3)
    fixed = M.plane_fixed_nodes(m, axis=0, value=0.0)
    perm = M.tiled_color4_permutation(m, fixed)
    assert np.array_equal(np.sort(perm), np.arange(m.n_nodes))
    r = M.relabel(m, perm)
    # the free nodes come first; their colours are non-decreasing in the new numbering
    n = np.arange(m.n_nodes)
    nx, ny, nz = m.shape
    k, j, i = n % (nz + 1), (n // (nz + 1)) % (ny + 1), n // ((nz + 1) * (ny + 1))
    color_new = np.empty(m.n_nodes, np.int64)
    color_new[perm] = (i + j + k) % 4
    nfree = m.n_nodes - len(fixed)
    assert np.all(np.diff(color_new[:nfree]) >= 0)
    assert np.array_equal(np.sort(perm[fixed]), np.arange(nfree, m.n_nodes))
    # proper: no edge joins two nodes of one colour
    E = np.concatenate([r.tets[:, [a, b]] for a in range(4) for b in range(a + 1, 4)])
    assert np.all(color_new[E[:, 0]] != color_new[E[:, 1]])


def test_rows_longer_than_16_blocks_are_refused(gs):
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(5)
    X = rng.random((400, 3))
    tets = Delaunay(X).simplices.astype(np.int32)
    st = np.zeros(8, np.int64)
    con = np.zeros(0, np.int32)
    rc = gs.gs_build(len(X), O._d(np.ascontiguousarray(X)), len(tets), O._i(np.ascontiguousarray(tets)), 0, O._i(con),
                     st.ctypes.data_as(C.POINTER(C.c_longlong)))
    assert rc == 0 and st[0] == 0
