// Driver around the reference's own host code (test infrastructure only).
//
// Compiled together with the reference sources, read in place from the reference
// checkout by oracle/Makefile, into oracle/_ref/libref.so.  It exposes the reference's
// per-element precompute (CTriangleInfo, src/triangleinfo.cpp:153-624), its grid
// set-up (CUnstructuredGrid, src/unstructuredgrid.cpp:14-80,126-162,502-643,729-760,
// 934-976) and its config parser (CConfig, src/config.cpp:25-112) through a C ABI so
// that tests/golden/make_golden.py can write golden vectors and the CPU tests can
// compare the oracle with the reference directly.  Nothing here computes anything
// itself.

#include <unistd.h>
#include <cstring>
#include <string>
#include "triangleinfo.h"
#include "unstructuredgrid.h"
#include "config.h"

extern "C" {

// Same output layout as orc_triangle_info at P2 (Np=6, Nqv=7, Nqe=3), n elements back
// to back.  x, y: [n][3].
int ref_triangle_info(int n, const double* x, const double* y, double* out) {
    CTriangleInfo T;
    T.allocateMemory(n);
    for (int e = 0; e < n; ++e) {
        double xx[3] = {x[3 * e], x[3 * e + 1], x[3 * e + 2]};
        double yy[3] = {y[3 * e], y[3 * e + 1], y[3 * e + 2]};
        T.initInformation(e, xx, yy);
    }
    const int Np = BASIS_FUNCTIONS, Nqv = VOLUME_GPOINTS, Ne = TRIANGLE_EDGES * EDGE_GPOINTS;
    double* o = out;
    for (int e = 0; e < n; ++e) {
        *o++ = T.area[e];
        *o++ = T.perimeter[e];
        for (int k = 0; k < 6; ++k) *o++ = T.outer_normal_vector[e + k * n];
        *o++ = T.barycenter[e].getX();
        *o++ = T.barycenter[e].getY();
        for (int i = 0; i < Np; ++i)
            for (int c = 0; c < BASIS_FUNCTION_COEFFS; ++c) *o++ = T.basis_fun_coeff[e * Np * BASIS_FUNCTION_COEFFS + i * BASIS_FUNCTION_COEFFS + c];
        for (int i = 0; i < Np; ++i) *o++ = T.mass_coeff[e + i * n];
        for (int q = 0; q < Nqv; ++q) *o++ = T.vol_gauss_weight[e + q * n];
        for (int p = 0; p < Ne; ++p) *o++ = T.edge_gauss_weight[e + p * n];
        for (int i = 0; i < Np; ++i)
            for (int q = 0; q < Nqv; ++q) *o++ = T.vol_bf_value[e + (i * Nqv + q) * n];
        for (int i = 0; i < 2 * Np; ++i)
            for (int q = 0; q < Nqv; ++q) *o++ = T.vol_bdf_value[e + (i * Nqv + q) * n];
        for (int i = 0; i < Np; ++i)
            for (int p = 0; p < Ne; ++p) *o++ = T.edge_bf_value[e + (i * Ne + p) * n];
    }
    return 0;
}

// Runs the reference grid set-up for one rank in directory `dir` (which must hold
// input/ElemDepart.dat and the files named by `meshconf`).  Outputs, for the first
// call with out arrays NULL: counts[0] = local triangles, counts[1] = local cells.
// With arrays: neighbour[3][cells], shared_edge[3][cells], flag[cells],
// bary[2][cells] (barycentres, ghost cells included).
int ref_grid(const char* dir, const char* meshconf, int* counts, int* neighbour, int* shared_edge,
             int* flag, double* bary) {
    char old[4096];
    if (!getcwd(old, sizeof old)) return -1;
    if (chdir(dir) != 0) return -2;
    int rc = 0;
    {
        CUnstructuredGrid grid;
        grid.config_file = meshconf;
        grid.initializeGrid(0, 1);
        grid.initializeGridNext();
        grid.triangle_infos.allocateMemory(grid.getLocalCellNumber());
        grid.initializeTriangleInfos();
        grid.markLocalBoundaryTriangles();
        int tn = grid.getLocalTriangleNumber(), cn = grid.getLocalCellNumber();
        counts[0] = tn;
        counts[1] = cn;
        if (neighbour) {
            for (int i = 0; i < 3 * cn; ++i) {
                neighbour[i] = grid.local_tri_neighbour[i];
                shared_edge[i] = grid.local_tri_sharedEdge[i];
            }
            for (int i = 0; i < cn; ++i) {
                flag[i] = grid.local_tri_flag[i];
                bary[i] = grid.triangle_infos.barycenter[i].getX();
                bary[cn + i] = grid.triangle_infos.barycenter[i].getY();
            }
        }
    }
    if (chdir(old) != 0) rc = -3;
    return rc;
}

// Parses `file` with the reference's CConfig for the given keys; values are written
// '\n'-separated into out.
int ref_config(const char* file, const char** keys, int nkeys, char* out, int outlen) {
    std::string* items = new std::string[nkeys];
    for (int i = 0; i < nkeys; ++i) items[i] = keys[i];
    CConfig conf(file, items, nkeys);
    conf.parseConfigFile();
    std::string all;
    for (int i = 0; i < nkeys; ++i) {
        all += conf.config_items[items[i]];
        all += "\n";
    }
    delete[] items;
    if ((int)all.size() + 1 > outlen) return -1;
    std::memcpy(out, all.c_str(), all.size() + 1);
    return 0;
}

}  // extern "C"
