"""The oracle's precompute against the reference's own code (CPU).

Golden vectors in tests/golden/ were written by tests/golden/make_golden.py from the
reference's CTriangleInfo / CUnstructuredGrid / CConfig (oracle/_ref/libref.so).  The
oracle restates src/triangleinfo.cpp and src/gauss.cpp with the same operation order and
must reproduce them bit for bit.  When oracle/_ref is built (a container with the
reference checkout), the comparison is also run live on fresh random triangles.
"""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def oracle_info_rows(P, x, y):
    lib = po.load()
    sz = lib.orc_info_size(P)
    out = np.zeros((x.shape[0], sz))
    for e in range(x.shape[0]):
        xe = np.ascontiguousarray(x[e])
        ye = np.ascontiguousarray(y[e])
        row = np.zeros(sz)
        lib.orc_triangle_info(P, xe.ctypes.data_as(po.D), ye.ctypes.data_as(po.D), row.ctypes.data_as(po.D))
        out[e] = row
    return out


def test_precompute_bitwise_equal_to_reference_golden():
    d = np.load(os.path.join(GOLD, "tri_info_p2.npz"))
    got = oracle_info_rows(2, d["x"], d["y"])
    assert got.shape == d["info"].shape
    assert np.array_equal(got, d["info"]), np.abs(got - d["info"]).max()


def test_precompute_bitwise_equal_to_reference_live():
    ref = po.load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    rng = np.random.default_rng(123)
    x = rng.uniform(-10, 10, (300, 3))
    y = rng.uniform(-10, 10, (300, 3))
    cr = (x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (y[:, 1] - y[:, 0]) * (x[:, 2] - x[:, 0])
    sw = cr < 0
    x[sw, 1], x[sw, 2] = x[sw, 2].copy(), x[sw, 1].copy()
    y[sw, 1], y[sw, 2] = y[sw, 2].copy(), y[sw, 1].copy()
    x, y = np.ascontiguousarray(x), np.ascontiguousarray(y)
    out = np.zeros((300, po.load().orc_info_size(2)))
    ref.ref_triangle_info(300, x.ctypes.data_as(po.D), y.ctypes.data_as(po.D), out.ctypes.data_as(po.D))
    assert np.array_equal(oracle_info_rows(2, x, y), out)


def test_p1_is_truncated_p2():
    """P1 keeps the first three modes of the P2 construction (same quadrature)."""
    d = np.load(os.path.join(GOLD, "tri_info_p2.npz"))
    x, y = d["x"][:20], d["y"][:20]
    for e in range(20):
        i1 = po.triangle_info(1, x[e], y[e])
        i2 = po.triangle_info(2, x[e], y[e])
        assert np.array_equal(i1["mass"], i2["mass"][:3])
        assert np.array_equal(i1["vol_bf"], i2["vol_bf"][:3])
        assert np.array_equal(i1["edge_bf"], i2["edge_bf"][:3])
        assert np.array_equal(i1["coeff"][:, :3], i2["coeff"][:3, :3])


def test_naca_ghosts_and_shared_edges_match_reference():
    """Per-edge boundary tags reproduce the reference's ghost flags; shared edges match."""
    d = np.load(os.path.join(GOLD, "naca_grid.npz"))
    tri, nb, ed = d["tri"], d["neigh"], d["edge"]
    E = tri.shape[0]
    rn, rse, rflag, rbary = d["ref_neighbour"], d["ref_shared_edge"], d["ref_flag"], d["ref_bary"]
    assert int(d["ref_counts"][0]) == E
    # interior shared edges: the neighbour's local edge with the same edge id
    for e in range(0, E, 7):
        for k in range(3):
            n = nb[e, k]
            if n >= 0:
                se = int(np.nonzero(ed[n] == ed[e, k])[0][0])
                assert rse[k, e] == se
                assert rn[k, e] == n
    # every boundary element has exactly one boundary edge in this mesh, so the
    # reference's one-ghost-per-element rule covers every boundary edge
    nbnd = (nb < 0).sum(axis=1)
    assert nbnd.max() == 1
    vx, vy = d["vertices"][:, 0], d["vertices"][:, 1]
    for e in np.nonzero(nbnd)[0]:
        k = int(np.nonzero(nb[e] < 0)[0][0])
        g = rn[k, e]
        # the ghost keeps the interior edge index; the interior sees the ghost's edge 0
        assert g >= E and rn[0, g] == e and rse[0, g] == k and rse[k, e] == 0
        a, b, c = tri[e, k], tri[e, (k + 1) % 3], tri[e, (k + 2) % 3]
        gx, gy = vx[b] + vx[c] - vx[a], vy[b] + vy[c] - vy[a]
        bx, by = (gx + vx[b] + vx[c]) / 3.0, (gy + vy[b] + vy[c]) / 3.0
        assert abs(bx - rbary[0, g]) < 1e-12 and abs(by - rbary[1, g]) < 1e-12
        far = bx < -15 or bx > 15 or by < -15 or by > 15
        assert rflag[g] == (1 if far else 3)


def test_config_golden_values_are_plain_data():
    with open(os.path.join(GOLD, "config_values.json")) as f:
        d = json.load(f)
    m = dict(zip(d["main.conf"]["keys"], d["main.conf"]["values"]))
    assert m["cfl"] == "0.18" and m["reductionthreads"] == "512"
    c = dict(zip(d["constructed"]["keys"], d["constructed"]["values"]))
    assert c == {"gamma": "1.4", "cfl": "0.25", "mach": "0.8", "title": "my", "pref": "", "rhoref": "2.5",
                 "time": "3"}


def test_reference_reader_reads_the_written_mesh(tmp_path):
    """The reference's own grid set-up (CUnstructuredGrid, run through oracle/_ref) reads a
    mesh written by rkdg_grid_write and builds the same neighbours, shared edges, ghost
    flags and ghost barycentres as from its original NACA0012 files."""
    ref = po.load_ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) absent")
    from rkdg_heter_calculate_amd import Grid
    from test_host import write_reference_mesh
    d = np.load(os.path.join(GOLD, "naca_grid.npz"))
    src = tmp_path / "src"
    src.mkdir()
    conf, part = write_reference_mesh(src, d)
    g = Grid.read(conf, part, box=15.0)
    # one rank: every element in part 0 (with parts beyond 0 and one rank the reference reads
    # a zero-length array, src/rkdgadvance.cpp:55 and src/unstructuredgrid.cpp:57-59)
    g.partition = np.zeros(g.n_elements, np.int32)
    run = tmp_path / "run"
    (run / "input").mkdir(parents=True)
    g.write(run / "input")  # the reference reads input/ElemDepart.dat from its working directory
    counts = np.zeros(2, np.int32)
    rc = ref.ref_grid(str(run).encode(), str(run / "input" / "mesh.conf").encode(), counts.ctypes.data_as(po.I),
                      None, None, None, None)
    assert rc == 0
    cn = int(counts[1])
    assert int(counts[0]) == int(d["ref_counts"][0]) and cn == d["ref_flag"].shape[0]
    nb = np.zeros(3 * cn, np.int32)
    se = np.zeros(3 * cn, np.int32)
    fl = np.zeros(cn, np.int32)
    ba = np.zeros(2 * cn)
    rc = ref.ref_grid(str(run).encode(), str(run / "input" / "mesh.conf").encode(), counts.ctypes.data_as(po.I),
                      nb.ctypes.data_as(po.I), se.ctypes.data_as(po.I), fl.ctypes.data_as(po.I),
                      ba.ctypes.data_as(po.D))
    assert rc == 0
    assert np.array_equal(nb.reshape(3, cn), d["ref_neighbour"])
    # the reference sets a ghost cell's shared edge 0 only (rows 1, 2 stay uninitialised)
    E = int(counts[0])
    se = se.reshape(3, cn)
    assert np.array_equal(se[:, :E], d["ref_shared_edge"][:, :E])
    assert np.array_equal(se[0, E:], d["ref_shared_edge"][0, E:])
    assert np.array_equal(fl, d["ref_flag"])
    assert np.allclose(ba.reshape(2, cn), d["ref_bary"], rtol=0, atol=1e-12)
