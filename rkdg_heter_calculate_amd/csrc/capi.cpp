// C ABI of librkdg: context set-up, device layout, state exchange and the time loop.
//
// Device layout (one context, E elements padded to ld = multiple of 256):
//   U[3]    three FP64 state buffers [4 vars][Np modes][ld]: the step-start state and
//           the two intermediate SSP-RK3 stages (neighbours read the previous stage while
//           the next is written, so stages cannot update in place)
//   geo     [5][ld] inverse Jacobian and area (edge lengths and normals follow from them)
//   nbr     [3][ld] int32 packed (device position << 2 | neighbour's edge), or -tag
//   dtflag  [ld] uint8, 1 if the element takes part in the CFL reduction
//   lim     [7][ld], pairs [ld] (TVB limiter geometry, only with the limiter)
//   st      DevState (wave-speed maxima, time, dt per step parity)
#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include "internal.h"

namespace rkdg {

void parallel_for(int n, const std::function<void(int, int)>& body) {
    unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min<unsigned>(hw ? hw : 1, 16);
    if (const char* s = std::getenv("RKDG_HOST_THREADS")) nt = std::max(1, std::atoi(s));
    if (n < 4096 || nt <= 1) { body(0, n); return; }
    std::vector<std::thread> th;
    int chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        int a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back(body, a, b);
    }
    for (auto& x : th) x.join();
}

int build_halo(const int* nbr, long long ld, int E, int B, std::vector<int>& hoff, std::vector<int>& hids,
               std::vector<int>& nloc) {
    const int ntiles = (E + B - 1) / B;
    hoff.assign(ntiles + 1, 0);
    hids.clear();
    nloc.assign((size_t)3 * ld, 0);
    int hmax = 0;
    std::vector<int> h;
    for (int t = 0; t < ntiles; ++t) {
        const int lo = t * B, hi = std::min(E, lo + B);
        h.clear();
        for (int k = 0; k < 3; ++k)
            for (int p = lo; p < hi; ++p) {
                const int w = nbr[(size_t)k * ld + p];
                if (w >= 0 && ((w >> 2) < lo || (w >> 2) >= hi)) h.push_back(w >> 2);
            }
        std::sort(h.begin(), h.end());
        h.erase(std::unique(h.begin(), h.end()), h.end());
        for (int k = 0; k < 3; ++k)
            for (int p = lo; p < hi; ++p) {
                const int w = nbr[(size_t)k * ld + p];
                int v = w;
                if (w >= 0) {
                    const int n = w >> 2;
                    const int slot = (n >= lo && n < hi) ? n - lo
                                                         : B + (int)(std::lower_bound(h.begin(), h.end(), n) - h.begin());
                    v = (slot << 2) | (w & 3);
                }
                nloc[(size_t)k * ld + p] = v;
            }
        hids.insert(hids.end(), h.begin(), h.end());
        hoff[t + 1] = (int)hids.size();
        hmax = std::max(hmax, (int)h.size());
    }
    for (int k = 0; k < 3; ++k)
        for (long long p = E; p < ld; ++p) nloc[(size_t)k * ld + p] = nbr[(size_t)k * ld + p];
    return hmax;
}

int max_lds_per_block(int device) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) return 65536;
    return v;
}

}  // namespace rkdg

using namespace rkdg;

struct rkdg_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    rkdg_params prm{};
    int P = 2, Np = 6, Nqv = 7, Nqe = 3, E = 0, block = 256, ntiles = 0, xcd_remap = 1;
    long long ld = 0;
    std::vector<int32_t> order, pos;  // order[p] = caller id at device position p
    std::vector<double> ex, ey;       // [E][3] element vertex coordinates (caller order)
    double* U[3] = {nullptr, nullptr, nullptr};
    double* geo = nullptr;
    int* nbr = nullptr;
    int* nloc = nullptr;        // staged form: neighbours as LDS slots
    int* hoff = nullptr;        // staged form: per-tile halo offsets
    int* hids = nullptr;        // staged form: halo element positions
    int lds_stride = 0;         // staged form: doubles per LDS row (block + largest halo)
    int staged = 0;
    int persistent = 0;         // workgroups of the persistent grid (0: one per tile)
    int stagger = 0;            // start offset of half of a persistent grid (8128-cycle sleeps)
    int pipe = 0;               // staged form: persistent, next image overlapped with the update
    unsigned char* dtflag = nullptr;
    double* lim = nullptr;
    int* pairs = nullptr;
    int* flags = nullptr;
    DevState* st = nullptr;
    double* res_part = nullptr;  // residual maxima per tile [ntiles][2], written by stage 2
    void* scratch = nullptr;     // this device's static scratch of rkdg_step_noalloc (not owned)
    int scratch_rc = 0;          // hipError_t of resolving it
    int cur = 0;            // buffer holding the latest state
    int next_stage = 0;     // rkdg_stage protocol
    long long steps = 0;    // steps launched (host count; parity of the step slot)
    bool have_state = false;
    size_t bytes = 0;
    hipGraphExec_t graph2 = nullptr;  // two steps (slot 0 then 1)
    std::string err;
};

static thread_local std::string g_err;

static int fail(rkdg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    g_err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                     \
    do {                                                                                      \
        hipError_t _e = (call);                                                               \
        if (_e != hipSuccess) return fail(ctx, RKDG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

static int hip_code(hipError_t e) { return e == hipErrorOutOfMemory ? RKDG_ERR_NOMEM : RKDG_ERR_HIP; }

// set-up: a failed HIP call releases the half-built context before it reports
#define CREATECHK(call)                                                                       \
    do {                                                                                      \
        hipError_t _e = (call);                                                               \
        if (_e != hipSuccess) {                                                               \
            std::string _m = std::string(#call) + ": " + hipGetErrorString(_e);               \
            rkdg_destroy(c);                                                                  \
            return fail(nullptr, hip_code(_e), _m);                                           \
        }                                                                                     \
    } while (0)

static StageArgs make_args(rkdg_ctx* c, int stage, int slot, int dt_mode, double dt, double t_end) {
    StageArgs a{};
    const int u_in = stage == 0 ? 0 : stage;     // stage s reads buffer s (0, 1, 2)
    const int u_out = stage == 2 ? 0 : stage + 1;
    a.U = c->U[u_in];
    a.U0 = c->U[0];
    a.Uout = c->U[u_out];
    a.geo = c->geo;
    a.nbr = c->nbr;
    a.nloc = c->nloc;
    a.hoff = c->hoff;
    a.hids = c->hids;
    a.lds_stride = c->lds_stride;
    a.staged = c->staged;
    a.persistent = c->persistent;
    a.stagger = c->stagger;
    a.pipe = c->pipe;
    a.dtflag = c->dtflag;
    a.st = c->st;
    a.res_part = c->res_part;
    a.ld = c->ld;
    a.E = c->E;
    a.stage = stage;
    a.slot = slot;
    a.dt_mode = dt_mode;
    a.dt_host = dt;
    a.xcd_remap = c->xcd_remap;
    a.ntiles = c->ntiles;
    a.cfl = c->prm.cfl;
    a.gamma = c->prm.gamma;
    a.t_end = t_end;
    for (int v = 0; v < 4; ++v) a.fs[v] = c->prm.freestream[v];
    return a;
}

// The stage kernel of a and, with the TVB limiter, the limiter pass over its output.
static int launch_stage_and_limiter(rkdg_ctx* c, const StageArgs& a, hipStream_t stream) {
    int rc = launch_stage(c->P, c->prm.flux, c->prm.precision, c->block, a, stream);
    if (rc) return fail(c, RKDG_ERR_HIP, std::string("stage kernel: ") + hip_error_string(rc));
    if (c->prm.limiter == RKDG_LIMITER_TVB) {
        LimArgs la{};
        la.U = a.Uout; la.lim = c->lim; la.pairs = c->pairs; la.nbr = c->nbr; la.flags = c->flags;
        la.ld = c->ld; la.E = c->E; la.M = c->prm.tvb_M; la.gamma = c->prm.gamma;
        la.st = a.st; la.slot = a.slot; la.t_end = a.t_end;
        rc = launch_limiter(c->P, la, c->block, stream);
        if (rc) return fail(c, RKDG_ERR_HIP, std::string("limiter kernel: ") + hip_error_string(rc));
    }
    return RKDG_OK;
}

static int launch_one_stage(rkdg_ctx* c, int s, int slot, int dt_mode, double dt, double t_end) {
    return launch_stage_and_limiter(c, make_args(c, s, slot, dt_mode, dt, t_end), c->stream);
}

static int launch_step(rkdg_ctx* c, int slot, double t_end) {
    int dt_mode = c->prm.dt_policy == RKDG_DT_FIXED ? DTM_FIXED : DTM_DEVICE;
    for (int s = 0; s < 3; ++s) {
        int rc = launch_one_stage(c, s, slot, dt_mode, c->prm.dt_fixed, t_end);
        if (rc) return rc;
    }
    return RKDG_OK;
}

static int refresh_wavespeed(rkdg_ctx* c, int slot, double t) {
    int rc = launch_state_init(c->st, slot, t, 1, c->stream);
    if (rc) return fail(c, RKDG_ERR_HIP, hip_error_string(rc));
    StageArgs a = make_args(c, 0, slot, DTM_DEVICE, 0.0, 0.0);
    a.U = c->U[c->cur];
    rc = launch_wavespeed(c->P, a, c->block, c->stream);
    if (rc) return fail(c, RKDG_ERR_HIP, hip_error_string(rc));
    return RKDG_OK;
}

static int check_physical(rkdg_ctx* c) {
    DevState h;
    HIPCHK(c, hipMemcpyAsync(&h, c->st, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (h.err_pos != INT_MAX) {
        int id = (h.err_pos >= 0 && h.err_pos < c->E) ? c->order[h.err_pos] : h.err_pos;
        return fail(c, RKDG_ERR_NONPHYSICAL, "non-positive density or pressure in element " + std::to_string(id));
    }
    return RKDG_OK;
}

extern "C" {

void rkdg_default_params(rkdg_params* p) {
    std::memset(p, 0, sizeof *p);
    p->degree = 2;
    p->flux = RKDG_FLUX_LF;
    p->limiter = RKDG_LIMITER_NONE;
    p->dt_policy = RKDG_DT_REFERENCE;
    p->reduction_threads = 512;
    p->precision = RKDG_PREC_FP64;
    p->device = 0;
    p->block = 0;
    p->gamma = 1.4;
    p->cfl = 0.18;
    p->tvb_M = 50.0;
    p->dt_fixed = 0.0;
    // free stream of src/rkdgadvance.cpp:31-34 with the defaults mach 0.4, alpha 5 deg
    double g = 1.4, ut = std::sqrt(g * 1.0 / 1.0) * 0.4, al = 5.0 * std::atan(1.0) * 4 / 180;
    p->freestream[0] = 1.0;
    p->freestream[1] = ut * std::cos(al);
    p->freestream[2] = ut * std::sin(al);
    p->freestream[3] = 0.5 * ut * ut + 1.0 / (g - 1);
}

const char* rkdg_strerror(int32_t code) {
    switch (code) {
        case RKDG_OK: return "ok";
        case RKDG_ERR_ARG: return "invalid argument";
        case RKDG_ERR_HIP: return "HIP runtime error";
        case RKDG_ERR_NONPHYSICAL: return "non-physical state";
        case RKDG_ERR_MESH: return "inconsistent mesh";
        case RKDG_ERR_IO: return "I/O error";
        case RKDG_ERR_STATE: return "call out of order";
        case RKDG_ERR_NOMEM: return "out of memory";
        default: return "unknown error";
    }
}

const char* rkdg_last_error(rkdg_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int rkdg_create(const rkdg_mesh* m, const rkdg_params* prm, rkdg_ctx** out) {
    if (!m || !prm || !out || m->n_elements < 1 || !m->vx || !m->vy || !m->tri || !m->neighbour)
        return fail(nullptr, RKDG_ERR_ARG, "mesh arrays missing");
    if (prm->degree < 1 || prm->degree > 3) return fail(nullptr, RKDG_ERR_ARG, "degree must be 1, 2 or 3");
    if (prm->flux != RKDG_FLUX_LF && prm->flux != RKDG_FLUX_ROE) return fail(nullptr, RKDG_ERR_ARG, "flux");
    if (prm->precision != RKDG_PREC_FP64 && prm->precision != RKDG_PREC_FP32_MFMA && prm->precision != RKDG_PREC_BF16_MFMA)
        return fail(nullptr, RKDG_ERR_ARG, "precision: RKDG_PREC_FP64, RKDG_PREC_FP32_MFMA or RKDG_PREC_BF16_MFMA");
    if (!(prm->cfl > 0) && prm->dt_policy != RKDG_DT_FIXED) return fail(nullptr, RKDG_ERR_ARG, "cfl");
    if (prm->dt_policy == RKDG_DT_FIXED && !(prm->dt_fixed > 0)) return fail(nullptr, RKDG_ERR_ARG, "dt_fixed");
    if (m->n_elements > (1 << 29) - 1) return fail(nullptr, RKDG_ERR_ARG, "too many elements");
    auto c = new rkdg_ctx();
    c->prm = *prm;
    c->P = prm->degree; c->Np = np_of(c->P); c->Nqv = nqv_of(c->P); c->Nqe = nqe_of(c->P);
    c->E = m->n_elements;
    // P3: 64-element tiles (measured 3.5% faster than 128 and 10% faster than 256 at 2.1M)
    c->block = prm->block > 0 ? prm->block : env_int("RKDG_BLOCK", prm->degree >= 3 ? 64 : 256);
    // precision form: staged tiles of 256 elements (lane per element), else 64 elements on
    // 256 threads (four lanes per element)
    const bool bf16 = prm->precision == RKDG_PREC_BF16_MFMA;
    const bool px = bf16 || (prm->precision == RKDG_PREC_FP32_MFMA && env_int("RKDG_PX", 1));
    if (prm->precision != RKDG_PREC_FP64) c->block = px ? px_block(c->P) : 64;
    if (c->block % 64 || c->block < 64 || c->block > 256) { delete c; return fail(nullptr, RKDG_ERR_ARG, "block must be 64..256, multiple of 64"); }
    c->xcd_remap = env_int("RKDG_XCD_REMAP", 1);
    c->ntiles = (c->E + c->block - 1) / c->block;
    c->ld = (long long)((c->E + 255) / 256) * 256;
    const int E = c->E;
    const long long ld = c->ld;
    // the kernels address every array with 32-bit byte offsets
    if (sizeof(double) * (size_t)4 * c->Np * ld >= (1ull << 32) || sizeof(double) * (size_t)GEO_ROWS * ld >= (1ull << 32)) {
        delete c;
        return fail(nullptr, RKDG_ERR_ARG, "mesh too large for one context (state or geometry array above 4 GiB)");
    }

    // locality permutation
    c->order.resize(E);
    c->pos.resize(E);
    if (m->order) {
        for (int p = 0; p < E; ++p) c->order[p] = m->order[p];
    } else if (env_int("RKDG_IDENTITY_ORDER", 0)) {
        for (int p = 0; p < E; ++p) c->order[p] = p;
    } else {
        // Z-order of the barycentres: workgroup tiles become compact patches, so most
        // neighbours of a tile lie inside it (small halos) and the rest are nearby
        std::vector<double> bx(E), by(E);
        for (int e = 0; e < E; ++e) {
            double sx = 0, sy = 0;
            for (int k = 0; k < 3; ++k) {
                const int v = m->tri[3 * e + k];
                if (v < 0 || v >= m->n_vertices) { delete c; return fail(nullptr, RKDG_ERR_MESH, "vertex id out of range"); }
                sx += m->vx[v];
                sy += m->vy[v];
            }
            bx[e] = sx / 3;
            by[e] = sy / 3;
        }
        morton_order(E, bx.data(), by.data(), c->order.data());
    }
    std::vector<char> seen(E, 0);
    for (int p = 0; p < E; ++p) {
        int id = c->order[p];
        if (id < 0 || id >= E || seen[id]) { delete c; return fail(nullptr, RKDG_ERR_ARG, "order is not a permutation"); }
        seen[id] = 1;
        c->pos[id] = p;
    }
    // element coordinates (caller order)
    c->ex.resize((size_t)3 * E);
    c->ey.resize((size_t)3 * E);
    for (int e = 0; e < E; ++e)
        for (int k = 0; k < 3; ++k) {
            int v = m->tri[3 * e + k];
            if (v < 0 || v >= m->n_vertices) { delete c; return fail(nullptr, RKDG_ERR_MESH, "vertex id out of range"); }
            c->ex[3 * e + k] = m->vx[v];
            c->ey[3 * e + k] = m->vy[v];
        }
    // shared edges
    std::vector<int32_t> se((size_t)3 * E);
    if (m->shared_edge) {
        std::memcpy(se.data(), m->shared_edge, sizeof(int32_t) * 3 * E);
        for (size_t i = 0; i < se.size(); ++i)
            if (m->neighbour[i] >= 0 && (se[i] < 0 || se[i] > 2)) {
                delete c;
                return fail(nullptr, RKDG_ERR_MESH, "shared edge index outside 0..2 at element " + std::to_string(i / 3));
            }
    } else {
        std::string err;
        int rc = build_shared_edges(E, m->tri, m->neighbour, se.data(), err, m->vx, m->vy);
        if (rc) { delete c; return fail(nullptr, rc, err); }
    }
    // host images of the device arrays (device order)
    std::vector<double> hgeo((size_t)GEO_ROWS * ld, 0.0);
    std::vector<int> hnbr((size_t)3 * ld, -RKDG_BC_OUTFLOW);
    std::vector<unsigned char> hflag(ld, 0);
    std::string merr;
    for (int p = 0; p < E; ++p) {
        int e = c->order[p];
        double g[GEO_HOST];
        element_geometry(&c->ex[3 * e], &c->ey[3 * e], g);
        if (!element_ccw(g)) { merr = "degenerate or clockwise element " + std::to_string(e); break; }
        for (int r = 0; r < GEO_ROWS; ++r) hgeo[(size_t)r * ld + p] = g[r];
        for (int k = 0; k < 3; ++k) {
            int n = m->neighbour[3 * e + k];
            if (n >= 0) {
                if (n >= E || m->neighbour[3 * n + se[3 * e + k]] != e) { merr = "neighbour relation not symmetric at element " + std::to_string(e); break; }
                hnbr[(size_t)k * ld + p] = (c->pos[n] << 2) | se[3 * e + k];
            } else {
                int t = m->boundary ? m->boundary[3 * e + k] : RKDG_BC_OUTFLOW;
                if (t != RKDG_BC_FARFIELD && t != RKDG_BC_OUTFLOW && t != RKDG_BC_REFLECTION) { merr = "unknown boundary tag"; break; }
                hnbr[(size_t)k * ld + p] = -t;
            }
        }
        int id = m->dt_id ? m->dt_id[e] : e;
        bool inc = true;
        if (prm->dt_policy == RKDG_DT_REFERENCE) {
            // kernel_getTimeStep's tree stops at index > 64 and its unrolled tail reads
            // mx[0..63] only: threads tid % 128 >= 64 drop out when blockDim >= 128
            int R = prm->reduction_threads > 0 ? prm->reduction_threads : 512;
            if (R >= 128) inc = (id % 128) < 64;
        }
        hflag[p] = inc ? 1 : 0;
        if (!merr.empty()) break;
    }
    if (!merr.empty()) { delete c; return fail(nullptr, RKDG_ERR_MESH, merr); }
    // limiter geometry
    std::vector<double> hlim;
    std::vector<int> hpairs;
    if (prm->limiter == RKDG_LIMITER_TVB) {
        hlim.assign((size_t)LIM_ROWS * ld, 0.0);
        hpairs.assign(ld, 0);
        for (int p = 0; p < E; ++p) {
            int e = c->order[p];
            const double* x = &c->ex[3 * e];
            const double* y = &c->ey[3 * e];
            double bx = (x[0] + x[1] + x[2]) / 3, by = (y[0] + y[1] + y[2]) / 3;
            double nbx[3], nby[3];
            for (int k = 0; k < 3; ++k) {
                int a = (k + 1) % 3, b = (k + 2) % 3;
                double mx = 0.5 * (x[a] + x[b]), my = 0.5 * (y[a] + y[b]);
                int n = m->neighbour[3 * e + k];
                if (n >= 0) {
                    const double* xn = &c->ex[3 * n];
                    const double* yn = &c->ey[3 * n];
                    int sn = se[3 * e + k];
                    double nmx = 0.5 * (xn[(sn + 1) % 3] + xn[(sn + 2) % 3]);
                    double nmy = 0.5 * (yn[(sn + 1) % 3] + yn[(sn + 2) % 3]);
                    // neighbour barycentre carried into this element's frame (periodic images)
                    nbx[k] = (xn[0] + xn[1] + xn[2]) / 3 + (mx - nmx);
                    nby[k] = (yn[0] + yn[1] + yn[2]) / 3 + (my - nmy);
                } else {
                    nbx[k] = 2 * mx - bx;
                    nby[k] = 2 * my - by;
                }
            }
            double l[LIM_ROWS];
            int pk;
            limiter_geometry(x, y, nbx, nby, l, &pk);
            for (int r = 0; r < LIM_ROWS; ++r) hlim[(size_t)r * ld + p] = l[r];
            hpairs[p] = pk;
        }
    }

    // staged form: per-tile halos, used when the tile's LDS image fits
    std::vector<int> hoff, hids, hnloc;
    // (measured faster than the global-gather form at P1 and P2)
    if (px || (prm->precision == RKDG_PREC_FP64 && c->block == staged_block(c->P) &&
               env_int("RKDG_STAGED", c->P <= 2 ? 1 : 0))) {
        const int hmax = build_halo(hnbr.data(), ld, E, c->block, hoff, hids, hnloc);
        const int stride = px ? px_stride(c->P) : staged_stride(c->P);
        const long long lds = px ? px_lds_bytes(c->P) : (long long)4 * c->Np * stride * (long long)sizeof(double);
        if (c->block + hmax <= stride && lds <= max_lds_per_block(prm->device)) {
            c->staged = 1;
            c->lds_stride = stride;
            int cus = 0;
            // P1 FP64 with LF: the register-pipelined persistent tile (k_stage_rp), measured 8%
            // faster per step at 2.1M triangles; with Roe (arithmetic-bound) it is 1-2% slower
            c->pipe = env_int("RKDG_PIPE", (!px && c->P == 1 && prm->flux == RKDG_FLUX_LF) ? 2 : 0);
            const int per_cu = env_int("RKDG_PERSIST", c->pipe ? 2 : 0);
            if (per_cu > 0 && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, prm->device) == hipSuccess)
                c->persistent = per_cu * cus;
            if (c->pipe && c->persistent == 0) c->pipe = 0;
            c->stagger = env_int("RKDG_STAGGER", 0);
        }
    }
    if (px && !c->staged) {
        if (bf16) {  // the BF16 contractions run on the staged tiles only
            delete c;
            return fail(nullptr, RKDG_ERR_ARG, "RKDG_PREC_BF16_MFMA: a tile's halo does not fit the staged LDS image");
        }
        c->block = 64;  // a halo did not fit: four lanes per element
        c->ntiles = (c->E + c->block - 1) / c->block;
    }

    // device
    c->dev = prm->device;
    CREATECHK(hipSetDevice(c->dev));
    CREATECHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    // the allocation-free step's scratch on this device (resolved once per device)
    c->scratch_rc = noalloc_scratch(c->dev, &c->scratch);
    size_t ubytes = sizeof(double) * (size_t)4 * c->Np * ld;
    for (int b = 0; b < 3; ++b) CREATECHK(hipMalloc(&c->U[b], ubytes));
    CREATECHK(hipMalloc(&c->geo, sizeof(double) * hgeo.size()));
    CREATECHK(hipMalloc(&c->nbr, sizeof(int) * hnbr.size()));
    CREATECHK(hipMalloc(&c->dtflag, hflag.size()));
    CREATECHK(hipMalloc(&c->flags, sizeof(int) * ld));
    CREATECHK(hipMalloc(&c->st, sizeof(DevState)));
    CREATECHK(hipMalloc(&c->res_part, sizeof(double) * 2 * c->ntiles));
    c->bytes = 3 * ubytes + sizeof(double) * hgeo.size() + sizeof(int) * hnbr.size() + hflag.size() + sizeof(int) * ld +
               sizeof(double) * 2 * c->ntiles;
    for (int b = 0; b < 3; ++b) CREATECHK(hipMemsetAsync(c->U[b], 0, ubytes, c->stream));
    CREATECHK(hipMemsetAsync(c->flags, 0, sizeof(int) * ld, c->stream));
    CREATECHK(hipMemsetAsync(c->res_part, 0, sizeof(double) * 2 * c->ntiles, c->stream));
    // host images are [rows][ld]; the device layout is tile-major
    const std::vector<double> tgeo = to_tile_major(hgeo, GEO_ROWS, ld);
    const std::vector<int> tnbr = to_tile_major(hnbr, NBR_ROWS, ld);
    CREATECHK(hipMemcpyAsync(c->geo, tgeo.data(), sizeof(double) * tgeo.size(), hipMemcpyHostToDevice, c->stream));
    CREATECHK(hipMemcpyAsync(c->nbr, tnbr.data(), sizeof(int) * tnbr.size(), hipMemcpyHostToDevice, c->stream));
    CREATECHK(hipMemcpyAsync(c->dtflag, hflag.data(), hflag.size(), hipMemcpyHostToDevice, c->stream));
    if (c->staged) {
        CREATECHK(hipMalloc(&c->nloc, sizeof(int) * hnloc.size()));
        CREATECHK(hipMalloc(&c->hoff, sizeof(int) * hoff.size()));
        CREATECHK(hipMalloc(&c->hids, sizeof(int) * std::max<size_t>(1, hids.size())));
        c->bytes += sizeof(int) * (hnloc.size() + hoff.size() + hids.size());
        const std::vector<int> tnloc = to_tile_major(hnloc, NBR_ROWS, ld);
        CREATECHK(hipMemcpyAsync(c->nloc, tnloc.data(), sizeof(int) * tnloc.size(), hipMemcpyHostToDevice, c->stream));
        CREATECHK(hipMemcpyAsync(c->hoff, hoff.data(), sizeof(int) * hoff.size(), hipMemcpyHostToDevice, c->stream));
        if (!hids.empty())
            CREATECHK(hipMemcpyAsync(c->hids, hids.data(), sizeof(int) * hids.size(), hipMemcpyHostToDevice, c->stream));
    }
    if (prm->limiter == RKDG_LIMITER_TVB) {
        CREATECHK(hipMalloc(&c->lim, sizeof(double) * hlim.size()));
        CREATECHK(hipMalloc(&c->pairs, sizeof(int) * hpairs.size()));
        c->bytes += sizeof(double) * hlim.size() + sizeof(int) * hpairs.size();
        const std::vector<double> tlim = to_tile_major(hlim, LIM_ROWS, ld);
        CREATECHK(hipMemcpyAsync(c->lim, tlim.data(), sizeof(double) * tlim.size(), hipMemcpyHostToDevice, c->stream));
        CREATECHK(hipMemcpyAsync(c->pairs, hpairs.data(), sizeof(int) * hpairs.size(), hipMemcpyHostToDevice, c->stream));
    }
    int rc = upload_tables(c->P, ref_basis(c->P).tab);
    if (rc) { std::string m2 = hip_error_string(rc); rkdg_destroy(c); return fail(nullptr, RKDG_ERR_HIP, "table upload: " + m2); }
    rc = launch_state_init(c->st, 0, 0.0, 1, c->stream);
    if (rc) { std::string m2 = hip_error_string(rc); rkdg_destroy(c); return fail(nullptr, RKDG_ERR_HIP, "state init: " + m2); }
    CREATECHK(hipStreamSynchronize(c->stream));
    *out = c;
    return RKDG_OK;
}

int rkdg_destroy(rkdg_ctx* c) {
    if (!c) return RKDG_OK;
    hipSetDevice(c->dev);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->graph2) hipGraphExecDestroy(c->graph2);
    for (int b = 0; b < 3; ++b) hipFree(c->U[b]);
    hipFree(c->geo); hipFree(c->nbr); hipFree(c->dtflag); hipFree(c->lim); hipFree(c->pairs);
    hipFree(c->nloc); hipFree(c->hoff); hipFree(c->hids);
    hipFree(c->flags); hipFree(c->st); hipFree(c->res_part);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return RKDG_OK;
}

int rkdg_set_state(rkdg_ctx* c, int32_t basis, const double* U) {
    if (!c || !U || (basis != RKDG_BASIS_ELEMENT && basis != RKDG_BASIS_NATIVE)) return fail(c, RKDG_ERR_ARG, "set_state");
    HIPCHK(c, hipSetDevice(c->dev));
    const int E = c->E, Np = c->Np, P = c->P;
    const long long ld = c->ld;
    std::vector<double> h((size_t)4 * Np * ld, 0.0);
    parallel_for(E, [&](int a, int b) {
        double in[4 * MAXNP], outv[4 * MAXNP];
        for (int p = a; p < b; ++p) {
            int e = c->order[p];
            for (int v = 0; v < 4; ++v)
                for (int i = 0; i < Np; ++i) in[v * Np + i] = U[((size_t)v * Np + i) * E + e];
            if (basis == RKDG_BASIS_ELEMENT) element_to_native(P, &c->ex[3 * e], &c->ey[3 * e], in, outv);
            else std::memcpy(outv, in, sizeof(double) * 4 * Np);
            for (int r = 0; r < 4 * Np; ++r) h[tile_index(r, 4 * Np, p)] = outv[r];
        }
    });
    HIPCHK(c, hipMemcpyAsync(c->U[0], h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
    c->cur = 0;
    c->next_stage = 0;
    c->steps = 0;
    c->have_state = true;
    int rc = refresh_wavespeed(c, 0, 0.0);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return check_physical(c);
}

int rkdg_get_state(rkdg_ctx* c, int32_t basis, double* U) {
    if (!c || !U || (basis != RKDG_BASIS_ELEMENT && basis != RKDG_BASIS_NATIVE)) return fail(c, RKDG_ERR_ARG, "get_state");
    HIPCHK(c, hipSetDevice(c->dev));
    const int E = c->E, Np = c->Np, P = c->P;
    const long long ld = c->ld;
    std::vector<double> h((size_t)4 * Np * ld);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->U[c->cur], sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    parallel_for(E, [&](int a, int b) {
        double in[4 * MAXNP], outv[4 * MAXNP];
        for (int p = a; p < b; ++p) {
            int e = c->order[p];
            for (int r = 0; r < 4 * Np; ++r) in[r] = h[tile_index(r, 4 * Np, p)];
            if (basis == RKDG_BASIS_ELEMENT) native_to_element(P, &c->ex[3 * e], &c->ey[3 * e], in, outv);
            else std::memcpy(outv, in, sizeof(double) * 4 * Np);
            for (int v = 0; v < 4; ++v)
                for (int i = 0; i < Np; ++i) U[((size_t)v * Np + i) * E + e] = outv[v * Np + i];
        }
    });
    return RKDG_OK;
}

int rkdg_project_ic(rkdg_ctx* c, int32_t kind, const double* p, int32_t np) {
    if (!c || !p) return fail(c, RKDG_ERR_ARG, "project_ic");
    int need = kind == RKDG_IC_FREESTREAM ? 5 : kind == RKDG_IC_VORTEX ? 6 : kind == RKDG_IC_RIEMANN4 ? 19 : -1;
    if (need < 0 || np < need) return fail(c, RKDG_ERR_ARG, "initial condition kind or parameter count");
    const int E = c->E, Np = c->Np, P = c->P;
    std::vector<double> U((size_t)4 * Np * E, 0.0);
    parallel_for(E, [&](int a, int b) {
        double d[4 * MAXNP];
        for (int e = a; e < b; ++e) {
            if (kind == RKDG_IC_FREESTREAM) {
                // mode 0 only, higher modes zero (src/cudarkdgsolver.cu:348-367)
                double s[4];
                ic_state(kind, p, np, 0, 0, s);
                for (int r = 0; r < 4 * Np; ++r) d[r] = 0;
                for (int v = 0; v < 4; ++v) d[v * Np] = s[v];
            } else {
                project_native(P, &c->ex[3 * e], &c->ey[3 * e], kind, p, np, d);
            }
            for (int v = 0; v < 4; ++v)
                for (int i = 0; i < Np; ++i) U[((size_t)v * Np + i) * E + e] = d[v * Np + i];
        }
    });
    return rkdg_set_state(c, RKDG_BASIS_NATIVE, U.data());
}

int rkdg_compute_dt(rkdg_ctx* c, double* dt) {
    if (!c || !dt) return fail(c, RKDG_ERR_ARG, "compute_dt");
    if (!c->have_state) return fail(c, RKDG_ERR_STATE, "no state");
    if (c->prm.dt_policy == RKDG_DT_FIXED) { *dt = c->prm.dt_fixed; return RKDG_OK; }
    HIPCHK(c, hipSetDevice(c->dev));
    int slot = (int)(c->steps & 1);
    // a fresh reduction into this step's wave-speed slot only: the simulated time, the
    // step count and the dt history stay as they are
    int rc = launch_ws_clear(c->st, slot, 1, c->stream);
    if (rc) return fail(c, RKDG_ERR_HIP, hip_error_string(rc));
    StageArgs a = make_args(c, 0, slot, DTM_DEVICE, 0.0, 0.0);
    a.U = c->U[c->cur];
    rc = launch_wavespeed(c->P, a, c->block, c->stream);
    if (rc) return fail(c, RKDG_ERR_HIP, hip_error_string(rc));
    DevState h;
    HIPCHK(c, hipMemcpyAsync(&h, c->st, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double ws;
    std::memcpy(&ws, &h.ws[slot], sizeof ws);
    if (!(ws > 0)) return fail(c, RKDG_ERR_NONPHYSICAL, "no positive wave speed");
    *dt = c->prm.cfl / ws;
    return check_physical(c);
}

int rkdg_stage(rkdg_ctx* c, int32_t s, double dt) {
    if (!c || s < 0 || s > 2 || !(dt >= 0)) return fail(c, RKDG_ERR_ARG, "stage");
    if (!c->have_state) return fail(c, RKDG_ERR_STATE, "no state");
    if (s != c->next_stage) return fail(c, RKDG_ERR_STATE, "stages must be called in order 0, 1, 2");
    HIPCHK(c, hipSetDevice(c->dev));
    int slot = (int)(c->steps & 1);
    int rc = launch_one_stage(c, s, slot, DTM_HOST, dt, 0.0);
    if (rc) return rc;
    c->cur = s == 2 ? 0 : s + 1;
    c->next_stage = (s + 1) % 3;
    if (s == 2) c->steps++;
    return RKDG_OK;
}

int rkdg_step_noalloc(rkdg_ctx* c, double dt, double* res_dev, void* stream) {
    if (!c || !(dt > 0)) return fail(c, RKDG_ERR_ARG, "step_noalloc: dt must be positive");
    if (!c->have_state) return fail(c, RKDG_ERR_STATE, "no state");
    if (c->next_stage != 0) return fail(c, RKDG_ERR_STATE, "a step started with rkdg_stage is unfinished");
    if (c->ntiles > NOALLOC_MAX_TILES)
        return fail(c, RKDG_ERR_ARG, "step_noalloc: " + std::to_string(c->ntiles) + " workgroup tiles exceed the scratch's " +
                                         std::to_string(NOALLOC_MAX_TILES));
    if (!c->scratch)
        return fail(c, RKDG_ERR_HIP, std::string("step_noalloc scratch: ") + hip_error_string(c->scratch_rc));
    HIPCHK(c, hipSetDevice(c->dev));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    DevState* st = (DevState*)c->scratch;
    double* part = (double*)((char*)c->scratch + 256);
    // the counters start cleared (a fill kernel: nothing here may record a memset into a
    // caller's graph); every tile writes its residual pair in stage 2
    int rc = launch_state_init(st, 0, 0.0, 1, s);
    if (rc) return fail(c, RKDG_ERR_HIP, hip_error_string(rc));
    for (int stage = 0; stage < 3; ++stage) {
        StageArgs a = make_args(c, stage, 0, DTM_HOST, dt, 0.0);
        a.st = st;
        a.res_part = part;
        rc = launch_stage_and_limiter(c, a, s);
        if (rc) return rc;
    }
    if (res_dev) {
        rc = launch_res_reduce(part, c->ntiles, res_dev, s);
        if (rc) return fail(c, RKDG_ERR_HIP, hip_error_string(rc));
    }
    c->cur = 0;
    return RKDG_OK;
}

int rkdg_step(rkdg_ctx* c, double* dt_used) {
    if (!c) return fail(c, RKDG_ERR_ARG, "step");
    if (!c->have_state) return fail(c, RKDG_ERR_STATE, "no state");
    if (c->next_stage != 0) return fail(c, RKDG_ERR_STATE, "a step started with rkdg_stage is unfinished");
    HIPCHK(c, hipSetDevice(c->dev));
    int slot = (int)(c->steps & 1);
    int rc = launch_step(c, slot, 0.0);
    if (rc) return rc;
    c->steps++;
    c->cur = 0;
    if (dt_used) {
        DevState h;
        HIPCHK(c, hipMemcpyAsync(&h, c->st, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *dt_used = h.dt[slot];
        if (h.err_pos != INT_MAX) return check_physical(c);
    }
    return RKDG_OK;
}

static int build_graph(rkdg_ctx* c) {
    if (c->graph2) return RKDG_OK;
    hipGraph_t g;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_step(c, 0, 0.0);
    if (!rc) rc = launch_step(c, 1, 0.0);
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (e2 == hipSuccess) hipGraphDestroy(g);
        return rc;
    }
    if (e2 != hipSuccess) return fail(c, RKDG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e2));
    hipError_t e3 = hipGraphInstantiate(&c->graph2, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e3 != hipSuccess) { c->graph2 = nullptr; return fail(c, RKDG_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e3)); }
    return RKDG_OK;
}

static int advance_impl(rkdg_ctx* c, int nsteps, double t_end, int* done_out, double* t_out) {
    int use_graph = env_int("RKDG_GRAPH", 1) && t_end <= 0;
    int done = 0;
    long long nsteps0 = 0;  // the device's step counter before this call
    if (done_out) {
        DevState h0;
        HIPCHK(c, hipMemcpyAsync(&h0, c->st, sizeof h0, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        nsteps0 = h0.nsteps;
    }
    if (t_end > 0) {
        // chunks without host synchronisation inside; the end time turns the remaining
        // launches of a chunk into no-ops on the device
        const int chunk = 32;
        DevState h;
        while (done < nsteps) {
            int n = std::min(chunk, nsteps - done);
            for (int i = 0; i < n; ++i) {
                int rc = launch_step(c, (int)(c->steps & 1), t_end);
                if (rc) return rc;
                c->steps++;
            }
            done += n;
            HIPCHK(c, hipMemcpyAsync(&h, c->st, sizeof h, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (h.err_pos != INT_MAX) return check_physical(c);
            if (h.t[c->steps & 1] >= t_end) break;
        }
    } else {
        if (use_graph && (c->steps & 1)) {
            int rc = launch_step(c, 1, 0.0);
            if (rc) return rc;
            c->steps++;
            done++;
        }
        if (use_graph && nsteps - done >= 2) {
            int rc = build_graph(c);
            if (rc) return rc;
            while (nsteps - done >= 2) {
                HIPCHK(c, hipGraphLaunch(c->graph2, c->stream));
                c->steps += 2;
                done += 2;
            }
        }
        while (done < nsteps) {
            int rc = launch_step(c, (int)(c->steps & 1), 0.0);
            if (rc) return rc;
            c->steps++;
            done++;
        }
    }
    c->cur = 0;
    if (done_out || t_out) {
        DevState h;
        HIPCHK(c, hipMemcpyAsync(&h, c->st, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (done_out) *done_out = (int)(h.nsteps - nsteps0);  // steps taken (the end time stops it)
        if (t_out) *t_out = h.t[c->steps & 1];
        if (h.err_pos != INT_MAX) return check_physical(c);
    }
    (void)done;
    return RKDG_OK;
}

int rkdg_advance(rkdg_ctx* c, int32_t nsteps, double t_end, int32_t* steps_done, double* t) {
    if (!c || nsteps < 0) return fail(c, RKDG_ERR_ARG, "advance");
    if (!c->have_state) return fail(c, RKDG_ERR_STATE, "no state");
    if (c->next_stage != 0) return fail(c, RKDG_ERR_STATE, "a step started with rkdg_stage is unfinished");
    HIPCHK(c, hipSetDevice(c->dev));
    int d = 0;
    int rc = advance_impl(c, nsteps, t_end, &d, t);
    if (steps_done) *steps_done = d;
    return rc;
}

int rkdg_residual(rkdg_ctx* c, double res[2]) {
    if (!c || !res) return fail(c, RKDG_ERR_ARG, "residual");
    HIPCHK(c, hipSetDevice(c->dev));
    // the last stage 2 left one pair of maxima per tile
    std::vector<double> h((size_t)2 * c->ntiles);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->res_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    res[0] = res[1] = 0.0;
    for (int t = 0; t < c->ntiles; ++t) {
        res[0] = std::max(res[0], h[(size_t)2 * t]);
        res[1] = std::max(res[1], h[(size_t)2 * t + 1]);
    }
    return RKDG_OK;
}

int rkdg_time(rkdg_ctx* c, double* t, int64_t* steps) {
    if (!c) return fail(c, RKDG_ERR_ARG, "time");
    HIPCHK(c, hipSetDevice(c->dev));
    DevState h;
    HIPCHK(c, hipMemcpyAsync(&h, c->st, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (t) *t = h.t[c->steps & 1];
    if (steps) *steps = h.nsteps;
    return RKDG_OK;
}

int rkdg_get_flags(rkdg_ctx* c, int32_t* flags) {
    if (!c || !flags) return fail(c, RKDG_ERR_ARG, "get_flags");
    HIPCHK(c, hipSetDevice(c->dev));
    std::vector<int> h(c->ld);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->flags, sizeof(int) * c->ld, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int p = 0; p < c->E; ++p) flags[c->order[p]] = h[p];
    return RKDG_OK;
}

int rkdg_sync(rkdg_ctx* c) {
    if (!c) return fail(c, RKDG_ERR_ARG, "sync");
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RKDG_OK;
}

int rkdg_time_stage(rkdg_ctx* c, int32_t s, int32_t reps, double* avg_ms) {
    if (!c || s < 0 || s > 1 || reps < 1 || !avg_ms) return fail(c, RKDG_ERR_ARG, "time_stage (stages 0 and 1 re-run idempotently)");
    if (!c->have_state) return fail(c, RKDG_ERR_STATE, "no state");
    HIPCHK(c, hipSetDevice(c->dev));
    int slot = (int)(c->steps & 1);
    double dt;
    int rc = rkdg_compute_dt(c, &dt);
    if (rc) return rc;
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    // warm-up
    StageArgs args = make_args(c, s, slot, DTM_HOST, dt, 0.0);
    int lrc = launch_stage(c->P, c->prm.flux, c->prm.precision, c->block, args, c->stream);
    HIPCHK(c, hipEventRecord(a, c->stream));
    for (int r = 0; r < reps && !lrc; ++r) lrc = launch_stage(c->P, c->prm.flux, c->prm.precision, c->block, args, c->stream);
    if (lrc) {
        hipEventDestroy(a);
        hipEventDestroy(b);
        return fail(c, RKDG_ERR_HIP, hip_error_string(lrc));
    }
    HIPCHK(c, hipEventRecord(b, c->stream));
    HIPCHK(c, hipEventSynchronize(b));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *avg_ms = ms / reps;
    c->cur = 0;
    return RKDG_OK;
}

int rkdg_time_advance(rkdg_ctx* c, int32_t nsteps, double* out) {
    // out[0]: total ms over the step loop; out[1]: mean ms per stage kernel launch,
    // bracketed individually; out[2]: stage kernel launches; out[3]: mean ms of a
    // limiter launch (0 without limiter)
    if (!c || nsteps < 1 || !out) return fail(c, RKDG_ERR_ARG, "time_advance");
    if (!c->have_state) return fail(c, RKDG_ERR_STATE, "no state");
    HIPCHK(c, hipSetDevice(c->dev));
    const int n = 3 * nsteps;
    const bool lim = c->prm.limiter == RKDG_LIMITER_TVB;
    std::vector<hipEvent_t> ev((size_t)2 * n + 2 + (lim ? 2 * n : 0));
    for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
    int dt_mode = c->prm.dt_policy == RKDG_DT_FIXED ? DTM_FIXED : DTM_DEVICE;
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    for (int st = 0; st < nsteps; ++st) {
        int slot = (int)(c->steps & 1);
        for (int s = 0; s < 3; ++s) {
            int k = st * 3 + s;
            StageArgs a = make_args(c, s, slot, dt_mode, c->prm.dt_fixed, 0.0);
            HIPCHK(c, hipEventRecord(ev[2 + 2 * k], c->stream));
            int rc = launch_stage(c->P, c->prm.flux, c->prm.precision, c->block, a, c->stream);
            if (rc) return fail(c, RKDG_ERR_HIP, hip_error_string(rc));
            HIPCHK(c, hipEventRecord(ev[3 + 2 * k], c->stream));
            if (lim) {
                LimArgs la{};
                la.U = a.Uout; la.lim = c->lim; la.pairs = c->pairs; la.nbr = c->nbr; la.flags = c->flags;
                la.ld = c->ld; la.E = c->E; la.M = c->prm.tvb_M; la.gamma = c->prm.gamma;
                la.st = a.st; la.slot = a.slot; la.t_end = 0.0;
                HIPCHK(c, hipEventRecord(ev[2 + 2 * n + 2 * k], c->stream));
                rc = launch_limiter(c->P, la, c->block, c->stream);
                if (rc) return fail(c, RKDG_ERR_HIP, hip_error_string(rc));
                HIPCHK(c, hipEventRecord(ev[3 + 2 * n + 2 * k], c->stream));
            }
        }
        c->steps++;
    }
    HIPCHK(c, hipEventRecord(ev[1], c->stream));
    HIPCHK(c, hipEventSynchronize(ev[1]));
    float ms;
    HIPCHK(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
    out[0] = ms;
    double sum = 0, suml = 0;
    for (int k = 0; k < n; ++k) {
        HIPCHK(c, hipEventElapsedTime(&ms, ev[2 + 2 * k], ev[3 + 2 * k]));
        sum += ms;
        if (lim) {
            HIPCHK(c, hipEventElapsedTime(&ms, ev[2 + 2 * n + 2 * k], ev[3 + 2 * n + 2 * k]));
            suml += ms;
        }
    }
    out[1] = sum / n;
    out[2] = n;
    out[3] = lim ? suml / n : 0.0;
    for (auto& e : ev) hipEventDestroy(e);
    c->cur = 0;
    return check_physical(c);
}

int rkdg_time_steps(rkdg_ctx* c, int32_t nsteps, double* total_ms) {
    if (!c || nsteps < 1 || !total_ms) return fail(c, RKDG_ERR_ARG, "time_steps");
    if (!c->have_state) return fail(c, RKDG_ERR_STATE, "no state");
    if (c->next_stage != 0) return fail(c, RKDG_ERR_STATE, "a step started with rkdg_stage is unfinished");
    HIPCHK(c, hipSetDevice(c->dev));
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    hipError_t e = hipEventCreate(&b);
    if (e != hipSuccess) { hipEventDestroy(a); return fail(c, RKDG_ERR_HIP, hipGetErrorString(e)); }
    int rc = RKDG_OK;
    e = hipEventRecord(a, c->stream);
    if (e == hipSuccess) {
        rc = advance_impl(c, nsteps, 0.0, nullptr, nullptr);
        if (!rc) e = hipEventRecord(b, c->stream);
    }
    if (!rc && e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0;
    if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return fail(c, RKDG_ERR_HIP, hipGetErrorString(e));
    *total_ms = ms;
    return check_physical(c);
}

int rkdg_error_vs_exact(rkdg_ctx* c, int32_t kind, const double* p, int32_t np, double out[8]) {
    if (!c || !p || !out || (kind != RKDG_IC_VORTEX && kind != RKDG_IC_FREESTREAM)) return fail(c, RKDG_ERR_ARG, "error_vs_exact");
    const int E = c->E, Np = c->Np, P = c->P;
    std::vector<double> U((size_t)4 * Np * E);
    int rc = rkdg_get_state(c, RKDG_BASIS_NATIVE, U.data());
    if (rc) return rc;
    const TriRule& R = projection_rule();
    const std::vector<double>& psi = projection_psi(P);
    // per-chunk sums (any number of host threads), merged in chunk order
    std::vector<std::pair<int, std::array<double, 12>>> parts;
    std::mutex parts_lock;
    parallel_for(E, [&](int a, int b) {
        double s[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int e = a; e < b; ++e) {
            const double* x = &c->ex[3 * e];
            const double* y = &c->ey[3 * e];
            double area = 0.5 * std::fabs((x[2] - x[0]) * (y[1] - y[0]) - (y[2] - y[0]) * (x[1] - x[0]));
            for (int q = 0; q < R.n; ++q) {
                double px = x[0] + (x[1] - x[0]) * R.xi[q] + (x[2] - x[0]) * R.eta[q];
                double py = y[0] + (y[1] - y[0]) * R.xi[q] + (y[2] - y[0]) * R.eta[q];
                double ex_[4];
                ic_state(kind, p, np, px, py, ex_);
                for (int v = 0; v < 4; ++v) {
                    double u = 0;
                    for (int i = 0; i < Np; ++i) u += U[((size_t)v * Np + i) * E + e] * psi[(size_t)q * Np + i];
                    double d = u - ex_[v];
                    s[v] += R.w[q] * area * d * d;
                    s[4 + v] += R.w[q] * area * ex_[v] * ex_[v];
                    s[8 + v] = std::max(s[8 + v], std::fabs(d));
                }
            }
        }
        std::array<double, 12> part;
        std::copy(s, s + 12, part.begin());
        std::lock_guard<std::mutex> g(parts_lock);
        parts.emplace_back(a, part);
    });
    std::sort(parts.begin(), parts.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    double acc[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (const auto& pt : parts)
        for (int v = 0; v < 4; ++v) {
            acc[v] += pt.second[v];
            acc[4 + v] += pt.second[4 + v];
            acc[8 + v] = std::max(acc[8 + v], pt.second[8 + v]);
        }
    for (int v = 0; v < 4; ++v) {
        out[v] = acc[4 + v] > 0 ? std::sqrt(acc[v] / acc[4 + v]) : std::sqrt(acc[v]);
        out[4 + v] = acc[8 + v];
    }
    return RKDG_OK;
}

int rkdg_measure_copy(int32_t device, int64_t bytes, int32_t reps, double* gbs) {
    if (bytes < 16 || reps < 1 || !gbs) return fail(nullptr, RKDG_ERR_ARG, "measure_copy");
    bytes = bytes / 16 * 16;
    HIPCHK(nullptr, hipSetDevice(device));
    hipStream_t s;
    HIPCHK(nullptr, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    void *a = nullptr, *b = nullptr;
    hipError_t e1 = hipMalloc(&a, bytes), e2 = hipMalloc(&b, bytes);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        hipFree(a); hipFree(b); hipStreamDestroy(s);
        return fail(nullptr, RKDG_ERR_NOMEM, "measure_copy buffers");
    }
    // the first failing call decides the status; cleanup is best effort
    hipError_t st = hipMemsetAsync(a, 0, bytes, s);
    hipEvent_t e0 = nullptr, e9 = nullptr;
    if (st == hipSuccess) st = hipEventCreate(&e0);
    if (st == hipSuccess) st = hipEventCreate(&e9);
    if (st == hipSuccess) st = (hipError_t)launch_stream_copy(a, b, bytes, s);  // warm-up
    if (st == hipSuccess) st = hipEventRecord(e0, s);
    for (int r = 0; r < reps && st == hipSuccess; ++r) st = (hipError_t)launch_stream_copy(a, b, bytes, s);
    if (st == hipSuccess) st = hipEventRecord(e9, s);
    if (st == hipSuccess) st = hipEventSynchronize(e9);
    float ms = 0;
    if (st == hipSuccess) st = hipEventElapsedTime(&ms, e0, e9);
    if (e0) (void)hipEventDestroy(e0);
    if (e9) (void)hipEventDestroy(e9);
    (void)hipFree(a); (void)hipFree(b); (void)hipStreamDestroy(s);
    if (st != hipSuccess) return fail(nullptr, RKDG_ERR_HIP, hipGetErrorString(st));
    if (!(ms > 0)) return fail(nullptr, RKDG_ERR_HIP, "measure_copy: no elapsed time");
    *gbs = 2.0 * (double)bytes * reps / (ms / 1e3) / 1e9;
    return RKDG_OK;
}

int rkdg_info(rkdg_ctx* c, int64_t out[5]) {
    if (!c || !out) return RKDG_ERR_ARG;
    out[0] = c->Np; out[1] = c->Nqv; out[2] = c->Nqe; out[3] = c->E; out[4] = (int64_t)c->bytes;
    return RKDG_OK;
}

}  // extern "C"
