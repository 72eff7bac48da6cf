"""Writes the golden fixtures of tests/golden/ from the reference's own host code.

Runs only in a container that has the reference checkout: it calls the reference's
CTriangleInfo / CUnstructuredGrid / CConfig through oracle/_ref/libref.so (compiled by
oracle/Makefile from the reference sources, in place) and reads the reference's mesh
input files.  The fixtures hold data only (inputs and the reference's outputs):

  tri_info_p2.npz     random triangles and the reference's per-element precompute
  naca_grid.npz       the NACA0012 mesh (input/previous/*) and the reference's grid set-up
                      (neighbours with ghost cells, shared edges, boundary flags, barycentres)
  config_values.json  values the reference's CConfig parses from its own config files and
                      from a few constructed lines

Usage: python tests/golden/make_golden.py [reference multi_gpuRkdg dir]
"""
import ctypes
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import pyoracle as po  # noqa: E402

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def random_triangles(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-3, 3, (n, 3))
    y = rng.uniform(-3, 3, (n, 3))
    # a few slivers and a few far-from-origin elements
    x[: n // 8, 2] = x[: n // 8, 0] + 1e-2 * rng.uniform(0.5, 1, n // 8)
    x[n // 8: n // 4] += 1e3
    cr = (x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (y[:, 1] - y[:, 0]) * (x[:, 2] - x[:, 0])
    keep = np.abs(cr) > 1e-6
    x, y, cr = x[keep], y[keep], cr[keep]
    sw = cr < 0
    x[sw, 1], x[sw, 2] = x[sw, 2].copy(), x[sw, 1].copy()
    y[sw, 1], y[sw, 2] = y[sw, 2].copy(), y[sw, 1].copy()
    return np.ascontiguousarray(x), np.ascontiguousarray(y)


def tri_info(ref, out_dir):
    x, y = random_triangles(256, 20240601)
    n = x.shape[0]
    sz = po.load().orc_info_size(2)
    out = np.zeros((n, sz))
    ref.ref_triangle_info(n, x.ctypes.data_as(D), y.ctypes.data_as(D), out.ctypes.data_as(D))
    np.savez_compressed(os.path.join(out_dir, "tri_info_p2.npz"), x=x, y=y, info=out)
    print("tri_info_p2.npz", out.shape)


def read_table(path, ncol, dtype, skip_header):
    rows = []
    with open(path, encoding="latin-1") as f:
        lines = f.read().splitlines()
    for ln in lines[1 if skip_header else 0:]:
        t = ln.split()
        if len(t) >= ncol:
            rows.append([dtype(v) for v in t[:ncol]])
    return np.array(rows, dtype=np.float64 if dtype is float else np.int32)


def naca_grid(ref, refdir, out_dir):
    src = os.path.join(refdir, "input", "previous")
    vert = read_table(os.path.join(src, "Vertice.dat"), 2, float, True)
    tri = read_table(os.path.join(src, "ElemVertice.dat"), 3, int, True)
    nb = read_table(os.path.join(src, "ElemNeigh.dat"), 3, int, True)
    ed = read_table(os.path.join(src, "ElemEdge.dat"), 3, int, True)
    ev = read_table(os.path.join(src, "EdgeVertice.dat"), 2, int, True)
    part = read_table(os.path.join(src, "ElemDepart.dat"), 1, int, False)[:, 0]
    E = tri.shape[0]
    tmp = tempfile.mkdtemp(prefix="rkdg_golden_")
    try:
        os.makedirs(os.path.join(tmp, "input"))
        for name in ("Vertice.dat", "ElemVertice.dat", "ElemEdge.dat", "EdgeVertice.dat"):
            shutil.copy(os.path.join(src, name), os.path.join(tmp, "input", name))
        # the reference's neighbour reader does not skip the header line
        # (src/unstructuredgrid.cpp:220): feed it a header-less copy
        with open(os.path.join(src, "ElemNeigh.dat"), encoding="latin-1") as f:
            lines = f.read().splitlines()[1:]
        with open(os.path.join(tmp, "input", "ElemNeigh.dat"), "w") as f:
            f.write("\n".join(lines) + "\n")
        with open(os.path.join(tmp, "input", "ElemDepart.dat"), "w") as f:
            f.write("\n".join("0" for _ in range(E)) + "\n")
        shutil.copy(os.path.join(src, "mesh.conf"), os.path.join(tmp, "mesh.conf"))
        counts = np.zeros(2, np.int32)
        rc = ref.ref_grid(tmp.encode(), b"mesh.conf", counts.ctypes.data_as(I), None, None, None, None)
        assert rc == 0, rc
        cn = int(counts[1])
        neigh = np.zeros(3 * cn, np.int32)
        se = np.zeros(3 * cn, np.int32)
        flag = np.zeros(cn, np.int32)
        bary = np.zeros(2 * cn)
        rc = ref.ref_grid(tmp.encode(), b"mesh.conf", counts.ctypes.data_as(I), neigh.ctypes.data_as(I),
                          se.ctypes.data_as(I), flag.ctypes.data_as(I), bary.ctypes.data_as(D))
        assert rc == 0, rc
    finally:
        shutil.rmtree(tmp)
    np.savez_compressed(os.path.join(out_dir, "naca_grid.npz"), vertices=vert, tri=tri, neigh=nb, edge=ed,
                        edge_vertices=ev, partition=part, ref_counts=counts, ref_neighbour=neigh.reshape(3, cn),
                        ref_shared_edge=se.reshape(3, cn), ref_flag=flag, ref_bary=bary.reshape(2, cn))
    print("naca_grid.npz", E, "elements,", cn, "cells")


def config_values(ref, refdir, out_dir):
    main_keys = ["title", "gamma", "alpha", "mach", "cfl", "rhoref", "pref", "time", "gridconf", "logfile",
                 "solutionfile", "residualfile", "threadsperblock", "reductionthreads", "loghistory",
                 "printinterval"]
    mesh_keys = ["trianglenum", "verticenum", "edgenum", "trivertice_filename", "trineigh_filename",
                 "triedge_filename", "vertice_filename", "edgevertice_filename", "output_filename"]
    constructed = ("# comment line\n"
                   "   gamma=1.4   # leading blanks and '='\n"
                   "cfl\t0.25;trailing\n"
                   "mach 0.8,0.9\n"
                   "title   my case name # blanks end a value\n"
                   "pref\n"
                   "rhoref =  2.5\n"
                   "time 3#no blank before the hash\n")
    res = {}
    tmp = tempfile.mkdtemp(prefix="rkdg_conf_")
    try:
        def stripped(path):
            # keep the key/value structure of the reference's config file, drop comment text
            out = []
            with open(path, "rb") as f:
                for raw in f.read().decode("latin-1").splitlines():
                    t = raw.lstrip(" \t")
                    if not t or t.startswith("#"):
                        out.append("#")
                        continue
                    h = raw.find("#")
                    line = raw if h < 0 else raw[:h] + "# c"
                    out.append("".join(ch if ord(ch) < 128 else "X" for ch in line))
            return "\n".join(out) + "\n"

        files = {}
        for name, src, keys in (("main.conf", os.path.join(refdir, "input", "main.conf"), main_keys),
                                ("mesh.conf", os.path.join(refdir, "input", "previous", "mesh.conf"), mesh_keys)):
            text = stripped(src)
            dst = os.path.join(tmp, name)
            with open(dst, "w") as f:
                f.write(text)
            files[name] = (dst, keys, text)
        cpath = os.path.join(tmp, "constructed.conf")
        with open(cpath, "w") as f:
            f.write(constructed)
        files["constructed"] = (cpath, ["gamma", "cfl", "mach", "title", "pref", "rhoref", "time"], constructed)
        for name, (path, keys, text) in files.items():
            arr = (ctypes.c_char_p * len(keys))(*[k.encode() for k in keys])
            buf = ctypes.create_string_buffer(1 << 16)
            rc = ref.ref_config(path.encode(), arr, len(keys), buf, len(buf))
            assert rc == 0
            vals = buf.value.decode("latin-1").split("\n")[: len(keys)]
            res[name] = {"keys": keys, "values": vals, "text": text}
    finally:
        shutil.rmtree(tmp)
    with open(os.path.join(out_dir, "config_values.json"), "w") as f:
        json.dump(res, f, indent=1)
    print("config_values.json")


def main():
    refdir = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/multi_gpuRkdg"
    ref = po.load_ref()
    if ref is None:
        sys.exit("oracle/_ref/libref.so is not built (needs the reference checkout)")
    tri_info(ref, HERE)
    naca_grid(ref, refdir, HERE)
    config_values(ref, refdir, HERE)


if __name__ == "__main__":
    main()
