// Internal declarations shared by the host C++ and the HIP kernels of librkdg.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>
#include "../../include/rkdg.h"

namespace rkdg {

constexpr int MAXNP = 10;   // P3
constexpr int MAXQV = 12;
constexpr int MAXQE = 4;
constexpr int NVAR = 4;
// Host geometry of one element: xi_x, xi_y, eta_x, eta_y, area, l0, l1, l2, n0x, n0y, n1x,
// n1y, n2x, n2y.  The device keeps the first GEO_ROWS; the kernels derive the edge lengths
// and outward normals from the inverse Jacobian and the area.
constexpr int GEO_HOST = 14;
constexpr int GEO_ROWS = 5;
constexpr int LIM_ROWS = 7;   // 3 x (a1, a2), h^2
constexpr int NBR_ROWS = 3;   // neighbour codes (and LDS slots) across edges 0, 1, 2

// Device arrays of R rows per element are stored tile-major, [ld / 256][R][256]: the 256
// elements of a layout tile keep all their rows in one contiguous block, so a tile's state
// is one sequential stream of HBM (24 rows x 2 KiB = 48 KiB at P2) rather than R streams
// ld elements apart.  Measured on MI355X: 7.2 TB/s for reading such blocks versus 5.6 TB/s
// for the same bytes as rows of an [R][ld] array; 5.6 versus 4.7 TB/s read plus write.
constexpr int LAYOUT_TILE = 256;

// Integer tuning switch from the environment (RKDG_*), or dflt when unset.
inline int env_int(const char* name, int dflt) {
    const char* s = std::getenv(name);
    return s ? std::atoi(s) : dflt;
}
inline size_t tile_index(int r, int R, long long p) {
    return ((size_t)(p / LAYOUT_TILE) * R + r) * LAYOUT_TILE + (size_t)(p % LAYOUT_TILE);
}
// [R][ld] host image -> tile-major device image (ld a multiple of LAYOUT_TILE)
template <class T>
std::vector<T> to_tile_major(const std::vector<T>& s, int R, long long ld) {
    std::vector<T> t(s.size());
    for (int r = 0; r < R; ++r)
        for (long long p = 0; p < ld; ++p) t[tile_index(r, R, p)] = s[(size_t)r * ld + p];
    return t;
}

inline int np_of(int P) { return (P + 1) * (P + 2) / 2; }
inline int nqv_of(int P) { return P <= 2 ? 7 : 12; }
inline int nqe_of(int P) { return P <= 2 ? 3 : 4; }

// Reference-element operators for one degree.  Reference triangle (0,0),(1,0),(0,1);
// xi = lambda_1, eta = lambda_2.  The basis psi is orthonormal in the mean over the
// triangle (psi_0 = 1), hierarchical by degree.
struct DevTables {
    double Vv[MAXQV][MAXNP];       // psi_i at the volume points
    double Dxi[MAXQV][MAXNP];      // d psi_i / d xi
    double Deta[MAXQV][MAXNP];     // d psi_i / d eta
    double wv[MAXQV];              // volume weights as fractions of the area
    double Ve[3][MAXQE][MAXNP];    // psi_i at edge k point j (edge k: vertex k+1 -> k+2)
    double we[MAXQE];              // edge weights per half length
    double minv[MAXNP];            // 1 / (quadrature mass per unit area)
    double Vmid[3][3];             // P1 modes at the edge midpoints (limiter)
    double recon[2][3];            // modes 1,2 from midpoint deviations (limiter)
    // the edge rule applied to a constant flux, sum_j we_j psi_i(edge k point j): the BF16
    // precision forms' edge corrections (for a constant flux the volume rule's
    // A (F.xi sum_q w_q d psi_i/d xi + F.eta sum_q w_q d psi_i/d eta) equals
    // sum_k l_k/2 F.n_k Se[k][i], so the precision forms leave both out)
    double Se[3][MAXNP];
};

// Device-side step state: wave-speed maxima and times per step parity.
struct DevState {
    unsigned long long ws[2];      // max (|u|+c) P/A as double bits (non-negative)
    double t[2];
    double dt[2];
    int err_pos;                   // smallest device position with non-physical mean (INT_MAX: none)
    int pad;
    long long nsteps;              // steps advanced (end-time no-ops excluded)
};

// Minimal fork-join helper for host loops (the library carries no OpenMP runtime).
void parallel_for(int n, const std::function<void(int, int)>& body);

// Host-side basis/geometry helpers (basis.cpp)
struct RefBasis {
    int P, Np, Nqv, Nqe;
    long double C[MAXNP][MAXNP];   // psi_k = sum_m C[k][m] * mono_m(xi, eta)
    DevTables tab;
};
const RefBasis& ref_basis(int P);
double psi_value(const RefBasis& B, int i, double xi, double eta);
void psi_grad(const RefBasis& B, int i, double xi, double eta, double* gx, double* ge);

// collapsed 6x6 Gauss rule on the reference triangle; weights sum to 1
struct TriRule { int n; std::vector<double> xi, eta, w; };
const TriRule& projection_rule();
const std::vector<double>& projection_psi(int P);  // psi_j at the rule's points, [q][j]

// Element geometry rows for one element (GEO_HOST doubles)
void element_geometry(const double* x, const double* y, double* g);
// True for a non-degenerate counter-clockwise element (positive Jacobian determinant)
inline bool element_ccw(const double* g) { return g[4] > 0 && g[0] * g[3] - g[1] * g[2] > 0; }

// The reference's per-element Gram-Schmidt basis (P <= 3), evaluated in physical
// coordinates.  Built from monomials 1, X, Y, X^2, XY, Y^2, X^3, ... with
// X = (x - xc)/sqrt(A), Y = (y - yc)/sqrt(A).
struct ElemBasis {
    int Np;
    double xc, yc, isq;
    double c[MAXNP][MAXNP];  // coefficients over the scaled monomials
    double value(int i, double x, double y) const;
};
void element_basis(int P, const double* x, const double* y, ElemBasis& out);

// Conversions for one element between coefficient sets (U: [4][Np] for this element)
void element_to_native(int P, const double* x, const double* y, const double* c, double* d);
void native_to_element(int P, const double* x, const double* y, const double* d, double* c);

// Initial-condition states
void ic_state(int kind, const double* p, int np, double x, double y, double* U);
// L2 projection onto the native basis of one element
void project_native(int P, const double* x, const double* y, int kind, const double* p, int np, double* d /*[4][Np]*/);

// Limiter geometry of one element (LIM_ROWS doubles + packed pairs)
void limiter_geometry(const double* x, const double* y, const double nbx[3], const double nby[3],
                      double* lim, int* pairs);

// Mesh helpers (grid.cpp)
// se[3e+k]: the neighbour's local index of edge k of element e, matched by vertex ids, or
// (with vertex coordinates) across a periodic seam by the translated edge that points back.
int build_shared_edges(int E, const int32_t* tri, const int32_t* nbr, int32_t* se, std::string& err,
                       const double* vx = nullptr, const double* vy = nullptr);
// Morton (Z-order) permutation of points (bx, by): order[p] = index of the p-th point.
void morton_order(int n, const double* bx, const double* by, int32_t* order);
// Staged-form tables for tiles of B consecutive device positions: per tile the sorted
// out-of-tile neighbour positions (hids[hoff[t] .. hoff[t+1])) and each element's
// neighbours re-encoded as slots of the tile's LDS image (slot < B: own tile, else
// B + index into the tile's halo).  nbr is [3][ld] packed (pos << 2) | edge or -tag.
// Returns the largest halo.
int build_halo(const int* nbr, long long ld, int E, int B, std::vector<int>& hoff, std::vector<int>& hids,
               std::vector<int>& nloc);
int max_lds_per_block(int device);

// Config parser (io.cpp)
bool config_parse(const char* path, std::vector<std::pair<std::string, std::string>>& items);

}  // namespace rkdg

// ---- kernel launchers (kernels.hip) ---------------------------------------------------
namespace rkdg {

struct StageArgs {
    const double* U;       // state the residual is evaluated on
    const double* U0;      // step-start state (read by stages 1, 2)
    double* Uout;          // stage result
    const double* geo;     // [GEO_ROWS][ld]
    const int* nbr;        // [3][ld] packed: (pos << 2) | shared edge, or -tag for a boundary
    const unsigned char* dtflag;
    DevState* st;
    double* res_part;      // stage 2: residual maxima per tile [ntiles][2] (NULL: not kept)
    const int* tile_mesh;  // batching: mesh id per workgroup tile (NULL: single mesh)
    DevState* st_batch;    // batching: per-mesh states
    const double* cfl_batch;
    long long ld;
    int E;
    int stage;
    int slot;
    int dt_mode;           // 0 device CFL, 1 host dt, 2 fixed
    int xcd_remap;
    int ntiles;            // workgroup tiles covering E
    int persistent;        // >0: grid capped at this many workgroups, each looping over tiles
    double dt_host;
    double cfl, gamma, t_end;
    double fs[4];
    // staged form: neighbour slots in the tile's LDS image and the per-tile halo lists
    const int* nloc;       // [3][ld] packed (slot << 2) | shared edge, or -tag for a boundary
    const int* hoff;       // [ntiles + 1] offsets into hids
    const int* hids;       // halo element positions, slot blockDim.x + h of tile t is hids[hoff[t] + h]
    int lds_stride;        // >= block + largest halo
    int staged;
    int pipe;              // staged form, persistent: next tile's image loaded during the update
    int stagger;           // persistent grids: the second half of the workgroups starts this
                           // many 8128-cycle sleeps late, so that the two workgroups sharing a
                           // CU alternate their memory and arithmetic phases
};

enum { DTM_DEVICE = 0, DTM_HOST = 1, DTM_FIXED = 2 };

struct LimArgs {
    double* U;
    const double* lim;     // [LIM_ROWS][ld]
    const int* pairs;      // [ld] packed neighbour pairs
    const int* nbr;
    int* flags;
    long long ld;
    int E;
    double M;
    double gamma;
    // end time: a launch after the end time is a no-op like its stage (its state is final
    // and its flags are the last stage's), decided from the step-start time st->t[slot]
    const DevState* st;
    int slot;
    double t_end;
};

// Staged form: tiles of staged_block(P) elements whose LDS image has rows of
// staged_stride(P) doubles (own tile, then at most stride - block halo elements).  The
// stride is a compile-time constant so that every row is an immediate LDS offset (below
// 64 KiB: 4 Np rows of stride doubles).
constexpr int staged_block(int P) { return P <= 2 ? 256 : 128; }
constexpr int staged_stride(int P) { return P <= 2 ? 352 : 200; }
// Precision form: 256-element staged tiles at every degree; at P3 its LDS image holds FP32
// modes (FP32 form) or BF16 deviations (BF16 form) and FP64 cell means (k_stage_pz; the bound
// below is the FP32 form's, the larger), at P1/P2 the FP64 image of the FP64 form.
constexpr int px_block(int P) { return 256; }
constexpr int px_stride(int P) { return 352; }
constexpr long long px_lds_bytes(int P) {
    return P <= 2 ? 4LL * ((P + 1) * (P + 2) / 2) * px_stride(P) * 8
                  : 4LL * ((P + 1) * (P + 2) / 2) * px_stride(P) * 4 + 4LL * px_stride(P) * 8;
}

int upload_tables(int P, const DevTables& t);  // per current device
int launch_stage(int P, int flux, int prec, int block, const StageArgs& a, void* stream);
int launch_wavespeed(int P, const StageArgs& a, int block, void* stream);
int launch_state_init(DevState* st, int slot, double t, int n, void* stream);
int launch_ws_clear(DevState* st, int slot, int n, void* stream);
// Allocation-free step (rkdg_step_noalloc): its scratch, one static device array per GPU
// holding a DevState in the first 256 bytes and then NOALLOC_MAX_TILES residual pairs.
constexpr int NOALLOC_MAX_TILES = RKDG_NOALLOC_MAX_TILES;
constexpr int NOALLOC_MAX_DEVICES = 64;
constexpr size_t NOALLOC_SCRATCH_BYTES = 256 + sizeof(double) * 2 * (size_t)NOALLOC_MAX_TILES;
// The scratch's device address on device dev (dev must be current); resolved once per
// device with hipGetSymbolAddress.  Returns a hipError_t.
int noalloc_scratch(int dev, void** out);
int launch_res_reduce(const double* part, int ntiles, double* out, void* stream);
int launch_limiter(int P, const LimArgs& a, int block, void* stream);
int launch_stream_copy(const void* src, void* dst, long long bytes, void* stream);
const char* hip_error_string(int code);

}  // namespace rkdg
