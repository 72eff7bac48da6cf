// Batching form: many independent meshes of uneven size advance together.
//
// All instances are concatenated into one tile-major layout; each instance starts on a
// 256-element boundary, so a layout tile and a workgroup tile (64, 128 or 256 elements)
// never span two instances, an instance's state is one contiguous block, and a workgroup
// tile carries a single instance id (tile_mesh).  Every instance keeps its own
// DevState, so its CFL dt is reduced and consumed on the device without any host
// synchronisation.  One launch per RK stage advances every instance; the stage kernels
// run persistently (grid capped at a few workgroups per CU, each looping over tiles),
// and two full steps are captured once in a hipGraph and replayed.  The workgroup size
// is chosen by timing (rkdg_batch_autotune).  Padding elements are valid dummy cells
// (unit triangle, transmissive sides, uniform state at rest) excluded from the CFL
// reduction.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "internal.h"

using namespace rkdg;

struct rkdg_batch {
    int dev = 0;
    hipStream_t stream = nullptr;
    rkdg_params prm{};
    int n = 0, P = 2, Np = 6, block = 256, persistent = 0, xcd_remap = 1;
    long long ld = 0;           // total padded elements
    std::vector<int> E, off;    // per instance
    std::vector<std::vector<int32_t>> order;  // per instance: caller element at device position p
    std::vector<std::vector<double>> ex, ey;
    int* nloc = nullptr;        // staged form tables (P2)
    int* hoff = nullptr;
    int* hids = nullptr;
    bool staged_ok = false;     // the halo of every tile fits the staged LDS image
    int staged = 0;             // staged form in use (chosen by the autotuner)
    double* U[3] = {nullptr, nullptr, nullptr};
    double* geo = nullptr;
    int* nbr = nullptr;
    unsigned char* dtflag = nullptr;
    DevState* st = nullptr;
    double* cfl = nullptr;
    int* tile_mesh = nullptr;
    int ntiles = 0;
    long long steps = 0;
    hipGraphExec_t graph2 = nullptr;
    size_t bytes = 0;
    std::string err;
};

static thread_local std::string g_berr;

static int bfail(rkdg_batch* b, int code, const std::string& m) {
    if (b) b->err = m;
    g_berr = m;
    return code;
}

#define BCHK(b, call)                                                                          \
    do {                                                                                       \
        hipError_t _e = (call);                                                                \
        if (_e != hipSuccess) return bfail(b, RKDG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

// set-up: a failed HIP call releases the half-built batch before it reports
#define BCREATECHK(call)                                                                       \
    do {                                                                                       \
        hipError_t _e = (call);                                                                \
        if (_e != hipSuccess) {                                                                \
            std::string _m = std::string(#call) + ": " + hipGetErrorString(_e);                \
            rkdg_batch_destroy(b);                                                             \
            return bfail(nullptr, _e == hipErrorOutOfMemory ? RKDG_ERR_NOMEM : RKDG_ERR_HIP, _m); \
        }                                                                                      \
    } while (0)

static int upload_tiles(rkdg_batch* b) {
    b->ntiles = (int)(b->ld / b->block);
    std::vector<int> tm(b->ntiles);
    for (int i = 0; i < b->n; ++i) {
        long long pad = (long long)(b->E[i] + 255) / 256 * 256;
        for (long long t = b->off[i] / b->block; t < (b->off[i] + pad) / b->block; ++t)
            tm[t] = (i << 1) | (t * b->block == b->off[i] ? 1 : 0);
    }
    // kernels and graph replays in flight read the old tile table
    BCHK(b, hipStreamSynchronize(b->stream));
    if (b->tile_mesh) hipFree(b->tile_mesh);
    b->tile_mesh = nullptr;
    BCHK(b, hipMalloc(&b->tile_mesh, sizeof(int) * tm.size()));
    BCHK(b, hipMemcpy(b->tile_mesh, tm.data(), sizeof(int) * tm.size(), hipMemcpyHostToDevice));
    if (b->graph2) { hipGraphExecDestroy(b->graph2); b->graph2 = nullptr; }
    return RKDG_OK;
}

static StageArgs bargs(rkdg_batch* b, int stage, int slot) {
    StageArgs a{};
    const int u_in = stage, u_out = stage == 2 ? 0 : stage + 1;
    a.U = b->U[u_in];
    a.U0 = b->U[0];
    a.Uout = b->U[u_out];
    a.geo = b->geo;
    a.nbr = b->nbr;
    a.nloc = b->nloc;
    a.hoff = b->hoff;
    a.hids = b->hids;
    a.staged = b->staged;
    a.lds_stride = b->prm.precision != RKDG_PREC_FP64 ? px_stride(b->P) : staged_stride(b->P);
    a.dtflag = b->dtflag;
    a.st = b->st;
    a.tile_mesh = b->tile_mesh;
    a.st_batch = b->st;
    a.cfl_batch = b->cfl;
    a.ld = b->ld;
    a.E = (int)b->ld;
    a.stage = stage;
    a.slot = slot;
    a.dt_mode = b->prm.dt_policy == RKDG_DT_FIXED ? DTM_FIXED : DTM_DEVICE;
    a.dt_host = b->prm.dt_fixed;
    a.xcd_remap = b->xcd_remap;
    a.ntiles = b->ntiles;
    a.persistent = b->persistent;
    a.cfl = b->prm.cfl;
    a.gamma = b->prm.gamma;
    a.t_end = 0.0;
    for (int v = 0; v < 4; ++v) a.fs[v] = b->prm.freestream[v];
    return a;
}

static int bstep(rkdg_batch* b, int slot) {
    for (int s = 0; s < 3; ++s) {
        int rc = launch_stage(b->P, b->prm.flux, b->prm.precision, b->block, bargs(b, s, slot), b->stream);
        if (rc) return bfail(b, RKDG_ERR_HIP, std::string("batch stage: ") + hip_error_string(rc));
    }
    return RKDG_OK;
}

static int bwavespeed_all(rkdg_batch* b) {
    int rc = launch_state_init(b->st, 0, 0.0, b->n, b->stream);
    if (rc) return bfail(b, RKDG_ERR_HIP, hip_error_string(rc));
    StageArgs a = bargs(b, 0, 0);
    a.U = b->U[0];
    rc = launch_wavespeed(b->P, a, b->block, b->stream);
    if (rc) return bfail(b, RKDG_ERR_HIP, hip_error_string(rc));
    b->steps = 0;
    return RKDG_OK;
}

extern "C" {

int rkdg_batch_create(int32_t n, const rkdg_mesh* meshes, const rkdg_params* params, rkdg_batch** out) {
    if (n < 1 || !meshes || !params || !out) return bfail(nullptr, RKDG_ERR_ARG, "batch arguments");
    const rkdg_params& p0 = params[0];
    if (p0.degree < 1 || p0.degree > 3) return bfail(nullptr, RKDG_ERR_ARG, "degree");
    if (p0.limiter != RKDG_LIMITER_NONE) return bfail(nullptr, RKDG_ERR_ARG, "the batching form runs without limiter");
    if (p0.precision != RKDG_PREC_FP64 && p0.precision != RKDG_PREC_FP32_MFMA && p0.precision != RKDG_PREC_BF16_MFMA)
        return bfail(nullptr, RKDG_ERR_ARG, "precision: RKDG_PREC_FP64, RKDG_PREC_FP32_MFMA or RKDG_PREC_BF16_MFMA");
    if (p0.block != 0 && (p0.block < 64 || p0.block > 256 || p0.block % 64 != 0 || 256 % p0.block != 0))
        return bfail(nullptr, RKDG_ERR_ARG, "block: 0 (library default) or 64, 128, 256");
    // The instances share one launch, so everything but the mesh, the initial state and
    // the CFL number must agree; a field the batch cannot honour per instance is rejected
    // rather than taken from instance 0.
    auto dt_subset = [](const rkdg_params& p) {  // the reference dt subset is on (R >= 128)
        return p.dt_policy == RKDG_DT_REFERENCE && (p.reduction_threads > 0 ? p.reduction_threads : 512) >= 128;
    };
    for (int i = 1; i < n; ++i) {
        const rkdg_params& pi = params[i];
        if (pi.degree != p0.degree || pi.flux != p0.flux || pi.gamma != p0.gamma ||
            pi.dt_policy != p0.dt_policy || pi.precision != p0.precision)
            return bfail(nullptr, RKDG_ERR_ARG, "instances must share degree, flux, gamma, dt policy, precision");
        if (pi.limiter != p0.limiter || pi.device != p0.device || pi.block != p0.block)
            return bfail(nullptr, RKDG_ERR_ARG, "instances must share limiter, device and block");
        if (p0.dt_policy == RKDG_DT_FIXED && pi.dt_fixed != p0.dt_fixed)
            return bfail(nullptr, RKDG_ERR_ARG, "instances must share dt_fixed");
        if (dt_subset(pi) != dt_subset(p0))
            return bfail(nullptr, RKDG_ERR_ARG, "instances must share the dt subset rule (reduction_threads)");
    }
    // the far-field state is one per launch: instances with far-field edges must agree on it
    int ff_src = 0;  // the instance whose free stream the launch uses
    {
        int first_ff = -1;
        for (int i = 0; i < n; ++i) {
            const rkdg_mesh& m = meshes[i];
            bool ff = false;
            if (m.boundary && m.neighbour && m.n_elements > 0)
                for (long long k = 0; k < 3LL * m.n_elements && !ff; ++k)
                    ff = m.neighbour[k] < 0 && m.boundary[k] == RKDG_BC_FARFIELD;
            if (!ff) continue;
            if (first_ff < 0) { first_ff = i; continue; }
            for (int v = 0; v < 4; ++v)
                if (params[i].freestream[v] != params[first_ff].freestream[v])
                    return bfail(nullptr, RKDG_ERR_ARG,
                                 "instances with far-field edges must share the free stream (instances " +
                                     std::to_string(first_ff) + " and " + std::to_string(i) + ")");
        }
        if (first_ff > 0) ff_src = first_ff;
    }
    for (int i = 0; i < n; ++i)
        if (meshes[i].n_elements < 1 || !meshes[i].tri || !meshes[i].neighbour || !meshes[i].vx || !meshes[i].vy)
            return bfail(nullptr, RKDG_ERR_ARG, "mesh " + std::to_string(i));
    auto b = new rkdg_batch();
    b->prm = p0;
    for (int v = 0; v < 4; ++v) b->prm.freestream[v] = params[ff_src].freestream[v];
    b->n = n;
    b->P = p0.degree;
    b->Np = np_of(b->P);
    b->block = p0.block > 0 ? p0.block : 256;
    if (p0.precision != RKDG_PREC_FP64) b->block = 64;  // 64 elements on 256 threads
    b->xcd_remap = 1;
    b->E.resize(n);
    b->off.resize(n);
    b->order.resize(n);
    b->ex.resize(n);
    b->ey.resize(n);
    long long tot = 0;
    for (int i = 0; i < n; ++i) {
        b->E[i] = meshes[i].n_elements;
        b->off[i] = (int)tot;
        tot += (long long)(b->E[i] + 255) / 256 * 256;
    }
    if (tot > (1LL << 29)) { delete b; return bfail(nullptr, RKDG_ERR_ARG, "too many elements"); }
    b->ld = tot;
    const long long ld = tot;
    std::vector<double> hgeo((size_t)GEO_ROWS * ld);
    std::vector<int> hnbr((size_t)3 * ld, -RKDG_BC_OUTFLOW);
    std::vector<unsigned char> hflag(ld, 0);
    std::vector<double> hU((size_t)4 * b->Np * ld, 0.0);
    // dummy padding cell: unit right triangle, fluid at rest
    {
        double x[3] = {0, 1, 0}, y[3] = {0, 0, 1}, g[GEO_HOST];
        element_geometry(x, y, g);
        for (long long p = 0; p < ld; ++p) {
            for (int r = 0; r < GEO_ROWS; ++r) hgeo[(size_t)r * ld + p] = g[r];
            hU[(size_t)(0 * b->Np) * ld + p] = 1.0;
            hU[(size_t)(3 * b->Np) * ld + p] = 1.0 / (p0.gamma - 1.0);
        }
    }
    std::string merr;
    for (int i = 0; i < n && merr.empty(); ++i) {
        const rkdg_mesh& m = meshes[i];
        const int E = b->E[i], off = b->off[i];
        b->ex[i].resize((size_t)3 * E);
        b->ey[i].resize((size_t)3 * E);
        std::vector<int32_t> se((size_t)3 * E);
        if (m.shared_edge) std::memcpy(se.data(), m.shared_edge, sizeof(int32_t) * 3 * E);
        else if (build_shared_edges(E, m.tri, m.neighbour, se.data(), merr, m.vx, m.vy)) break;
        for (int e = 0; e < E; ++e)
            for (int k = 0; k < 3; ++k) {
                int v = m.tri[3 * e + k];
                if (v < 0 || v >= m.n_vertices) { merr = "vertex id out of range"; break; }
                b->ex[i][3 * e + k] = m.vx[v];
                b->ey[i][3 * e + k] = m.vy[v];
            }
        if (!merr.empty()) break;
        // device order: the caller's, or the Z-order of the barycentres (compact tiles)
        std::vector<int32_t>& ord = b->order[i];
        std::vector<int32_t> pos(E);
        ord.resize(E);
        if (m.order) {
            for (int p = 0; p < E; ++p) ord[p] = m.order[p];
        } else {
            std::vector<double> bx(E), by(E);
            for (int e = 0; e < E; ++e) {
                bx[e] = (b->ex[i][3 * e] + b->ex[i][3 * e + 1] + b->ex[i][3 * e + 2]) / 3;
                by[e] = (b->ey[i][3 * e] + b->ey[i][3 * e + 1] + b->ey[i][3 * e + 2]) / 3;
            }
            morton_order(E, bx.data(), by.data(), ord.data());
        }
        {
            std::vector<char> seen(E, 0);
            for (int p = 0; p < E; ++p) {
                const int e = ord[p];
                if (e < 0 || e >= E || seen[e]) { merr = "order is not a permutation"; break; }
                seen[e] = 1;
                pos[e] = p;
            }
        }
        if (!merr.empty()) break;
        for (int p = 0; p < E; ++p) {
            const int e = ord[p];
            double g[GEO_HOST];
            element_geometry(&b->ex[i][3 * e], &b->ey[i][3 * e], g);
            if (!element_ccw(g)) { merr = "degenerate or clockwise element"; break; }
            for (int r = 0; r < GEO_ROWS; ++r) hgeo[(size_t)r * ld + off + p] = g[r];
            for (int k = 0; k < 3; ++k) {
                int nb = m.neighbour[3 * e + k];
                if (nb >= E) { merr = "neighbour id out of range"; break; }
                if (nb >= 0) {
                    const int sk = se[3 * e + k];
                    if (sk < 0 || sk > 2 || m.neighbour[3 * nb + sk] != e) {
                        merr = "neighbour relation not symmetric at element " + std::to_string(e);
                        break;
                    }
                    hnbr[(size_t)k * ld + off + p] = ((off + pos[nb]) << 2) | sk;
                } else {
                    const int t = m.boundary ? m.boundary[3 * e + k] : RKDG_BC_OUTFLOW;
                    if (t != RKDG_BC_FARFIELD && t != RKDG_BC_OUTFLOW && t != RKDG_BC_REFLECTION) {
                        merr = "unknown boundary tag";
                        break;
                    }
                    hnbr[(size_t)k * ld + off + p] = -t;
                }
            }
            if (!merr.empty()) break;
            int id = m.dt_id ? m.dt_id[e] : e;
            bool inc = true;
            if (p0.dt_policy == RKDG_DT_REFERENCE) {
                int R = p0.reduction_threads > 0 ? p0.reduction_threads : 512;
                if (R >= 128) inc = (id % 128) < 64;
            }
            hflag[off + p] = inc;
        }
    }
    if (!merr.empty()) { delete b; return bfail(nullptr, RKDG_ERR_MESH, merr); }
    std::vector<double> hcfl(n);
    for (int i = 0; i < n; ++i) hcfl[i] = params[i].cfl;

    b->dev = p0.device;
    BCREATECHK(hipSetDevice(b->dev));
    BCREATECHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    size_t ub = sizeof(double) * hU.size();
    // the kernels address every array with 32-bit byte offsets
    if (ub >= (1ull << 32) || sizeof(double) * hgeo.size() >= (1ull << 32)) {
        rkdg_batch_destroy(b);
        return bfail(nullptr, RKDG_ERR_ARG, "batch too large (state or geometry array above 4 GiB)");
    }
    // host images are [rows][ld]; the device layout is tile-major
    {
        const std::vector<double> tU = to_tile_major(hU, 4 * b->Np, ld);
        for (int k = 0; k < 3; ++k) {
            BCREATECHK(hipMalloc(&b->U[k], ub));
            BCREATECHK(hipMemcpy(b->U[k], tU.data(), ub, hipMemcpyHostToDevice));
        }
    }
    BCREATECHK(hipMalloc(&b->geo, sizeof(double) * hgeo.size()));
    BCREATECHK(hipMalloc(&b->nbr, sizeof(int) * hnbr.size()));
    BCREATECHK(hipMalloc(&b->dtflag, hflag.size()));
    BCREATECHK(hipMalloc(&b->st, sizeof(DevState) * n));
    BCREATECHK(hipMalloc(&b->cfl, sizeof(double) * n));
    BCREATECHK(hipMemcpy(b->geo, to_tile_major(hgeo, GEO_ROWS, ld).data(), sizeof(double) * hgeo.size(), hipMemcpyHostToDevice));
    BCREATECHK(hipMemcpy(b->nbr, to_tile_major(hnbr, NBR_ROWS, ld).data(), sizeof(int) * hnbr.size(), hipMemcpyHostToDevice));
    BCREATECHK(hipMemcpy(b->dtflag, hflag.data(), hflag.size(), hipMemcpyHostToDevice));
    BCREATECHK(hipMemcpy(b->cfl, hcfl.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    b->bytes = 3 * ub + sizeof(double) * hgeo.size() + sizeof(int) * hnbr.size() + hflag.size();
    const bool prec_form = p0.precision != RKDG_PREC_FP64;
    if (b->P <= 2 || prec_form) {
        // staged-form tables for 256-element tiles (instances start on 256-element
        // boundaries, so no tile spans two instances): the FP64 form's at P1/P2, the
        // precision form's at every degree (at P3 its FP32-image tiles, k_stage_pz)
        std::vector<int> hoff, hids, hnloc;
        const int B = prec_form ? px_block(b->P) : staged_block(b->P);
        const int stride = prec_form ? px_stride(b->P) : staged_stride(b->P);
        const long long lds = prec_form ? px_lds_bytes(b->P) : 4LL * b->Np * stride * (long long)sizeof(double);
        const int hmax = build_halo(hnbr.data(), ld, (int)ld, B, hoff, hids, hnloc);
        if (B + hmax <= stride && lds <= max_lds_per_block(b->dev)) {
            BCREATECHK(hipMalloc(&b->nloc, sizeof(int) * hnloc.size()));
            BCREATECHK(hipMalloc(&b->hoff, sizeof(int) * hoff.size()));
            BCREATECHK(hipMalloc(&b->hids, sizeof(int) * std::max<size_t>(1, hids.size())));
            BCREATECHK(hipMemcpy(b->nloc, to_tile_major(hnloc, NBR_ROWS, ld).data(), sizeof(int) * hnloc.size(),
                              hipMemcpyHostToDevice));
            BCREATECHK(hipMemcpy(b->hoff, hoff.data(), sizeof(int) * hoff.size(), hipMemcpyHostToDevice));
            if (!hids.empty())
                BCREATECHK(hipMemcpy(b->hids, hids.data(), sizeof(int) * hids.size(), hipMemcpyHostToDevice));
            b->bytes += sizeof(int) * (hnloc.size() + hoff.size() + hids.size());
            b->staged_ok = true;
        }
    }
    // precision form: staged tiles, one lane per element, when every halo fits (else four
    // lanes per element on 64-element tiles)
    if (p0.precision == RKDG_PREC_BF16_MFMA && !b->staged_ok) {
        rkdg_batch_destroy(b);
        return bfail(nullptr, RKDG_ERR_ARG, "RKDG_PREC_BF16_MFMA: a tile's halo does not fit the staged LDS image");
    }
    if (prec_form && b->staged_ok) {
        b->block = px_block(b->P);
        b->staged = 1;
    }
    // FP64: the global-gather tiles until rkdg_batch_autotune picks; RKDG_BATCH_STAGED=1
    // starts from the staged tiles instead (when every halo fits)
    if (p0.precision == RKDG_PREC_FP64 && b->staged_ok && env_int("RKDG_BATCH_STAGED", 0)) {
        b->block = staged_block(b->P);
        b->staged = 1;
    }
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->dev) == hipSuccess && prop.multiProcessorCount > 0) cus = prop.multiProcessorCount;
    b->persistent = cus * 8;
    int rc = upload_tables(b->P, ref_basis(b->P).tab);
    if (rc) { rkdg_batch_destroy(b); return bfail(nullptr, RKDG_ERR_HIP, "table upload"); }
    rc = upload_tiles(b);
    if (rc) { std::string m2 = b->err; rkdg_batch_destroy(b); return bfail(nullptr, rc, m2); }
    rc = bwavespeed_all(b);
    if (rc) { std::string m2 = b->err; rkdg_batch_destroy(b); return bfail(nullptr, rc, m2); }
    BCREATECHK(hipStreamSynchronize(b->stream));
    *out = b;
    return RKDG_OK;
}

const char* rkdg_batch_last_error(rkdg_batch* b) { return b ? b->err.c_str() : g_berr.c_str(); }

int rkdg_batch_destroy(rkdg_batch* b) {
    if (!b) return RKDG_OK;
    hipSetDevice(b->dev);
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->graph2) hipGraphExecDestroy(b->graph2);
    for (int k = 0; k < 3; ++k) hipFree(b->U[k]);
    hipFree(b->geo); hipFree(b->nbr); hipFree(b->dtflag); hipFree(b->st); hipFree(b->cfl); hipFree(b->tile_mesh);
    hipFree(b->nloc); hipFree(b->hoff); hipFree(b->hids);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return RKDG_OK;
}

int rkdg_batch_set_state(rkdg_batch* b, int32_t i, int32_t basis, const double* U) {
    if (!b || i < 0 || i >= b->n || !U) return bfail(b, RKDG_ERR_ARG, "batch_set_state");
    BCHK(b, hipSetDevice(b->dev));
    const int E = b->E[i], Np = b->Np, P = b->P;
    // the instance's layout tiles are one contiguous block of the tile-major state; its
    // padding columns keep their rest state, so the block is read, updated and written back
    const size_t blk = (size_t)(E + LAYOUT_TILE - 1) / LAYOUT_TILE * LAYOUT_TILE * 4 * Np;
    double* dev = b->U[0] + (size_t)b->off[i] * 4 * Np;
    std::vector<double> h(blk);
    BCHK(b, hipMemcpyAsync(h.data(), dev, sizeof(double) * blk, hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    parallel_for(E, [&](int a0, int a1) {
        double in[4 * MAXNP], o[4 * MAXNP];
        for (int p = a0; p < a1; ++p) {
            const int e = b->order[i][p];
            for (int v = 0; v < 4; ++v)
                for (int k = 0; k < Np; ++k) in[v * Np + k] = U[((size_t)v * Np + k) * E + e];
            if (basis == RKDG_BASIS_ELEMENT) element_to_native(P, &b->ex[i][3 * e], &b->ey[i][3 * e], in, o);
            else std::memcpy(o, in, sizeof(double) * 4 * Np);
            for (int r = 0; r < 4 * Np; ++r) h[tile_index(r, 4 * Np, p)] = o[r];
        }
    });
    BCHK(b, hipMemcpyAsync(dev, h.data(), sizeof(double) * blk, hipMemcpyHostToDevice, b->stream));
    int rc = bwavespeed_all(b);
    if (rc) return rc;
    BCHK(b, hipStreamSynchronize(b->stream));
    return RKDG_OK;
}

int rkdg_batch_get_state(rkdg_batch* b, int32_t i, int32_t basis, double* U) {
    if (!b || i < 0 || i >= b->n || !U) return bfail(b, RKDG_ERR_ARG, "batch_get_state");
    BCHK(b, hipSetDevice(b->dev));
    const int E = b->E[i], Np = b->Np, P = b->P;
    const size_t blk = (size_t)(E + LAYOUT_TILE - 1) / LAYOUT_TILE * LAYOUT_TILE * 4 * Np;
    std::vector<double> h(blk);
    BCHK(b, hipMemcpyAsync(h.data(), b->U[0] + (size_t)b->off[i] * 4 * Np, sizeof(double) * blk, hipMemcpyDeviceToHost,
                           b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    parallel_for(E, [&](int a0, int a1) {
        double in[4 * MAXNP], o[4 * MAXNP];
        for (int p = a0; p < a1; ++p) {
            const int e = b->order[i][p];
            for (int r = 0; r < 4 * Np; ++r) in[r] = h[tile_index(r, 4 * Np, p)];
            if (basis == RKDG_BASIS_ELEMENT) native_to_element(P, &b->ex[i][3 * e], &b->ey[i][3 * e], in, o);
            else std::memcpy(o, in, sizeof(double) * 4 * Np);
            for (int v = 0; v < 4; ++v)
                for (int k = 0; k < Np; ++k) U[((size_t)v * Np + k) * E + e] = o[v * Np + k];
        }
    });
    return RKDG_OK;
}

int rkdg_batch_project_ic(rkdg_batch* b, int32_t i, int32_t kind, const double* p, int32_t np) {
    if (!b || i < 0 || i >= b->n || !p) return bfail(b, RKDG_ERR_ARG, "batch_project_ic");
    int need = kind == RKDG_IC_FREESTREAM ? 5 : kind == RKDG_IC_VORTEX ? 6 : kind == RKDG_IC_RIEMANN4 ? 19 : -1;
    if (need < 0 || np < need) return bfail(b, RKDG_ERR_ARG, "initial condition");
    const int E = b->E[i], Np = b->Np, P = b->P;
    std::vector<double> U((size_t)4 * Np * E, 0.0);
    parallel_for(E, [&](int a0, int a1) {
        double d[4 * MAXNP];
        for (int e = a0; e < a1; ++e) {
            if (kind == RKDG_IC_FREESTREAM) {
                double s[4];
                ic_state(kind, p, np, 0, 0, s);
                for (int r = 0; r < 4 * Np; ++r) d[r] = 0;
                for (int v = 0; v < 4; ++v) d[v * Np] = s[v];
            } else {
                project_native(P, &b->ex[i][3 * e], &b->ey[i][3 * e], kind, p, np, d);
            }
            for (int v = 0; v < 4; ++v)
                for (int k = 0; k < Np; ++k) U[((size_t)v * Np + k) * E + e] = d[v * Np + k];
        }
    });
    return rkdg_batch_set_state(b, i, RKDG_BASIS_NATIVE, U.data());
}

int rkdg_batch_advance(rkdg_batch* b, int32_t nsteps, int32_t use_graph) {
    if (!b || nsteps < 0) return bfail(b, RKDG_ERR_ARG, "batch_advance");
    BCHK(b, hipSetDevice(b->dev));
    int done = 0;
    if (use_graph && (b->steps & 1) && nsteps > 0) {
        int rc = bstep(b, 1);
        if (rc) return rc;
        b->steps++;
        done++;
    }
    if (use_graph && nsteps - done >= 2) {
        if (!b->graph2) {
            hipGraph_t g;
            BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = bstep(b, 0);
            if (!rc) rc = bstep(b, 1);
            hipError_t e2 = hipStreamEndCapture(b->stream, &g);
            if (rc) return rc;
            if (e2 != hipSuccess) return bfail(b, RKDG_ERR_HIP, std::string("capture: ") + hipGetErrorString(e2));
            hipError_t e3 = hipGraphInstantiate(&b->graph2, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e3 != hipSuccess) { b->graph2 = nullptr; return bfail(b, RKDG_ERR_HIP, hipGetErrorString(e3)); }
        }
        while (nsteps - done >= 2) {
            BCHK(b, hipGraphLaunch(b->graph2, b->stream));
            b->steps += 2;
            done += 2;
        }
    }
    while (done < nsteps) {
        int rc = bstep(b, (int)(b->steps & 1));
        if (rc) return rc;
        b->steps++;
        done++;
    }
    return RKDG_OK;
}

int rkdg_batch_sync(rkdg_batch* b) {
    if (!b) return RKDG_ERR_ARG;
    BCHK(b, hipSetDevice(b->dev));
    BCHK(b, hipStreamSynchronize(b->stream));
    std::vector<DevState> h(b->n);
    BCHK(b, hipMemcpy(h.data(), b->st, sizeof(DevState) * b->n, hipMemcpyDeviceToHost));
    for (int i = 0; i < b->n; ++i)
        if (h[i].err_pos != INT_MAX)
            return bfail(b, RKDG_ERR_NONPHYSICAL, "non-physical state in instance " + std::to_string(i));
    return RKDG_OK;
}

int rkdg_batch_time(rkdg_batch* b, int32_t i, double* t) {
    if (!b || i < 0 || i >= b->n || !t) return bfail(b, RKDG_ERR_ARG, "batch_time");
    BCHK(b, hipSetDevice(b->dev));
    DevState h;
    BCHK(b, hipMemcpyAsync(&h, b->st + i, sizeof h, hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    *t = h.t[b->steps & 1];
    return RKDG_OK;
}

}  // extern "C"

// Workgroups per CU of the device (256 CUs on MI355X).
static int device_cus(int dev) {
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) cus = prop.multiProcessorCount;
    return cus;
}

// Checks one tile configuration against the batch's form: staged 256-element tiles (when
// every halo fits the LDS image), or global-gather tiles of 64, 128 or 256 elements (the
// precision form's global-gather kernel takes 64-element tiles only); per_cu 0 launches
// one workgroup per tile, else a persistent grid of per_cu 256-thread workgroups per CU.
static bool tiles_valid(const rkdg_batch* b, int staged, int block, int per_cu) {
    if (per_cu != 0 && per_cu != 4 && per_cu != 8) return false;
    if (staged) {
        if (b->prm.precision != RKDG_PREC_FP64) return b->staged_ok && block == px_block(b->P);
        return b->staged_ok && b->P <= 2 && block == staged_block(b->P);
    }
    if (b->prm.precision == RKDG_PREC_BF16_MFMA) return false;  // staged tiles only
    if (b->prm.precision == RKDG_PREC_FP32_MFMA) return block == 64;
    return block == 64 || block == 128 || block == 256;
}

static int apply_tiles(rkdg_batch* b, int staged, int block, int per_cu) {
    b->staged = staged;
    b->block = block;
    b->persistent = per_cu ? device_cus(b->dev) * per_cu * 256 / block : 0;
    return upload_tiles(b);
}

extern "C" {

int rkdg_batch_set_tiles(rkdg_batch* b, int32_t staged, int32_t block, int32_t per_cu) {
    if (!b) return bfail(nullptr, RKDG_ERR_ARG, "batch_set_tiles");
    if (!tiles_valid(b, staged, block, per_cu))
        return bfail(b, RKDG_ERR_ARG, "tile configuration not available for this batch");
    BCHK(b, hipSetDevice(b->dev));
    return apply_tiles(b, staged ? 1 : 0, block, per_cu);
}

int rkdg_batch_get_tiles(rkdg_batch* b, int32_t out[3]) {
    if (!b || !out) return bfail(b, RKDG_ERR_ARG, "batch_get_tiles");
    const int cus = device_cus(b->dev);
    out[0] = b->staged;
    out[1] = b->block;
    out[2] = b->persistent ? b->persistent * b->block / (256 * cus) : 0;
    return RKDG_OK;
}

int rkdg_batch_autotune(rkdg_batch* b, int32_t* best_block) {
    if (!b) return RKDG_ERR_ARG;
    BCHK(b, hipSetDevice(b->dev));
    const int cands[3] = {64, 128, 256};
    const int pers[3] = {0, 4, 8};  // 0: one workgroup per tile; else workgroups per CU
    float best = 1e30f;
    int bb = b->block, bs = b->staged;
    int bp = 0;
    {
        int32_t cur[3];
        rkdg_batch_get_tiles(b, cur);
        bp = cur[2];
    }
    hipEvent_t e0, e1;
    BCHK(b, hipEventCreate(&e0));
    BCHK(b, hipEventCreate(&e1));
    auto finish = [&](int code, const std::string& m) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return bfail(b, code, m);
    };
    for (int st = 0; st <= 1; ++st)
        for (int c : cands)
            for (int pw : pers) {
                if (!tiles_valid(b, st, c, pw)) continue;
                int rc = apply_tiles(b, st, c, pw);
                if (rc) return finish(rc, b->err);
                // stage 1 reads buffers 1 and 0 and writes buffer 2: repeatable without
                // changing the state
                StageArgs a = bargs(b, 1, (int)(b->steps & 1));
                int lrc = launch_stage(b->P, b->prm.flux, b->prm.precision, b->block, a, b->stream);
                hipError_t he = hipEventRecord(e0, b->stream);
                for (int r = 0; r < 10 && !lrc; ++r) lrc = launch_stage(b->P, b->prm.flux, b->prm.precision, b->block, a, b->stream);
                if (lrc) return finish(RKDG_ERR_HIP, std::string("autotune stage: ") + hip_error_string(lrc));
                if (he == hipSuccess) he = hipEventRecord(e1, b->stream);
                if (he == hipSuccess) he = hipEventSynchronize(e1);
                float ms = 0;
                if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
                if (he != hipSuccess) return finish(RKDG_ERR_HIP, std::string("autotune timing: ") + hipGetErrorString(he));
                if (ms < best) { best = ms; bb = c; bp = pw; bs = st; }
            }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    int rc = apply_tiles(b, bs, bb, bp);
    if (rc) return rc;
    if (best_block) *best_block = bb;
    return RKDG_OK;
}

int rkdg_batch_info(rkdg_batch* b, int64_t out[4]) {
    if (!b || !out) return RKDG_ERR_ARG;
    long long tot = 0;
    for (int e : b->E) tot += e;
    out[0] = b->n; out[1] = tot; out[2] = b->block; out[3] = (int64_t)b->bytes;
    return RKDG_OK;
}

}  // extern "C"
