// gfx950 kernels of the plain FP64 form.
//
// One fused kernel per SSP-RK3 stage replaces the reference's nine launches per stage
// (calculateConVars, boundaryCondition, calculateVolumeRHS, calculateLFCoeff,
// calculateEdgeFG, calculateFlux x2, calculateEdgeRHS, rkdgStep*;
// src/kernelfunctions.cu:84-748).  One lane owns one element:
//   * its modal state is read once, coalesced, from the tile-major layout
//     U[ld / 256][var][mode][256] (a 256-element tile's modes are one contiguous block);
//   * the volume integral uses reference-element operators held in constant memory
//     (uniform across the wavefront, so they are scalar loads) and the element's
//     inverse Jacobian;
//   * each edge's flux is computed by the element itself from its own trace and the
//     neighbour's trace, evaluated from the neighbour's modes at the mirrored point
//     (element-owned gather: no atomics, no scatter, no trace buffers in HBM);
//   * the SSP-RK3 combination and the division by the diagonal mass are applied before
//     the single store;
//   * the last stage also reduces the CFL wave speed of the new cell means (the next
//     step's dt) and the reference's residual, so dt never leaves the device.
// Workgroups are mapped XCD-contiguously: with round-robin dispatch over the 8 XCDs,
// workgroup b runs on XCD b%8, and it processes tile (b%8)*(tiles/8) + b/8 so that each
// XCD sweeps one contiguous range of elements and neighbour gathers hit its own L2.
#include <hip/hip_runtime.h>
#include <climits>
#include <mutex>
#include <cstdlib>
#include <type_traits>
#include "internal.h"

namespace rkdg {

__constant__ DevTables c_tab[3];

template <int P>
struct Deg {
    static constexpr int Np = (P + 1) * (P + 2) / 2;
    static constexpr int Nqv = P <= 2 ? 7 : 12;
    static constexpr int Nqe = P <= 2 ? 3 : 4;
};

// Element e of row r of an array of R rows in the tile-major layout [ld / 256][R][256]
// (LAYOUT_TILE), addressed as the array base (scalar registers) plus one 32-bit byte
// offset (every array is below 4 GiB, checked when a context is created): the addressing
// form of global_load/store with an SGPR base, so no 64-bit address is formed per row.
template <typename T>
__device__ __forceinline__ T& tmx(T* base, int r, int R, int e) {
    const unsigned idx = ((unsigned)(e >> 8) * (unsigned)R + (unsigned)r) * 256u + (unsigned)(e & 255);
    auto b = reinterpret_cast<typename std::conditional<std::is_const<T>::value, const char*, char*>::type>(base);
    return *reinterpret_cast<T*>(b + idx * (unsigned)sizeof(T));
}

// The same arrays as buffer resources: one VGPR byte offset per element (tm_voff) and the
// row as a scalar byte offset (row r of a layout tile starts r * 2048 bytes after row 0), so
// a loop over rows costs no per-row address registers.  Raw buffer accesses are range-checked
// against the array size: reads outside return 0 and writes outside are dropped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tm_rsrc(const void* base, long long ld, int R, int size = 8) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)(unsigned)(ld * R * size), 0x00020000);
}
__device__ __forceinline__ unsigned tm_voff(int e, int R, int size = 8) {
    return (((unsigned)(e >> 8) * (unsigned)R) * 256u + (unsigned)(e & 255)) * (unsigned)size;
}
__device__ __forceinline__ int tm_load_i32(__amdgpu_buffer_rsrc_t rs, unsigned voff, int r) {
    return __builtin_bit_cast(int, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, r * 1024, 0));
}
__device__ __forceinline__ double tm_load(__amdgpu_buffer_rsrc_t rs, unsigned voff, int r) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, r * 2048, 0));
}
typedef unsigned int u32x2 __attribute__((__vector_size__(8)));
__device__ __forceinline__ void tm_store(__amdgpu_buffer_rsrc_t rs, unsigned voff, int r, double x) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), rs, (int)voff, r * 2048, 0);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Raises *p (the bits of a non-negative double, which order like the values) to x.  The
// atomic is skipped when the current value already covers x: one workgroup per tile would
// otherwise queue tens of thousands of atomics on one address per stage.  The final value is
// the maximum either way (values only increase), so the result stays order-independent.
__device__ __forceinline__ void atomic_raise(unsigned long long* p, double x) {
    if (!(x > 0)) return;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
    if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < bits) atomicMax(p, bits);
}

// Workgroup maxima of stage 2: the CFL wave speed ws is raised into *pws (one atomic at
// most); the two residual maxima r0, r1 are stored as this tile's pair rpart[0..1] with
// plain stores (NULL: not kept) and max-reduced on the host when asked for.  Device-wide
// atomics cost about 6 us per residual per stage at 0.5M triangles; the stores cost nothing.
__device__ __forceinline__ void block_stage_max(double ws, double r0, double r1, unsigned long long* pws,
                                                double* rpart) {
    __shared__ double red[3][16];
    ws = wave_max(ws);
    if (rpart) { r0 = wave_max(r0); r1 = wave_max(r1); }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) { red[0][wid] = ws; red[1][wid] = r0; red[2][wid] = r1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ma = 0, mb = 0, mc = 0;
        for (int w = 0; w < nw; ++w) { ma = fmax(ma, red[0][w]); mb = fmax(mb, red[1][w]); mc = fmax(mc, red[2][w]); }
        atomic_raise(pws, ma);
        if (rpart) { rpart[0] = mb; rpart[1] = mc; }
    }
}

// Physical fluxes F, G of the Euler equations (kernel_calculateEdgeFG's formulas,
// src/kernelfunctions.cu:413-428, with rho*u taken directly from the state).
__device__ __forceinline__ void euler_fg(double g, const double* U, double* F, double* G) {
    const double ir = 1.0 / U[0];
    const double u = U[1] * ir, v = U[2] * ir;
    const double p = (g - 1.0) * (U[3] - 0.5 * U[0] * (u * u + v * v));
    F[0] = U[1]; F[1] = U[1] * u + p; F[2] = U[1] * v; F[3] = u * (U[3] + p);
    G[0] = U[2]; G[1] = U[2] * u; G[2] = U[2] * v + p; G[3] = v * (U[3] + p);
}

// 1/x from the hardware reciprocal refined by two Newton steps (about 1 ulp), in place
// of the IEEE division sequence.
__device__ __forceinline__ double frcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double t = fma(-x, r, 1.0);
    r = fma(t, r, r);
    t = fma(-x, r, 1.0);
    return fma(t, r, r);
}

__device__ __forceinline__ void euler_fg_fast(double g, const double* U, double* F, double* G) {
    const double ir = frcp(U[0]);
    const double u = U[1] * ir, v = U[2] * ir;
    const double p = (g - 1.0) * (U[3] - 0.5 * (U[1] * u + U[2] * v));
    F[0] = U[1]; F[1] = U[1] * u + p; F[2] = U[1] * v; F[3] = u * (U[3] + p);
    G[0] = U[2]; G[1] = U[2] * u; G[2] = U[2] * v + p; G[3] = v * (U[3] + p);
}

// Flux of the Euler equations in direction (a, b): F a + G b, for any (a, b) (a unit
// normal on edges, a row of the inverse Jacobian in the volume).
__device__ __forceinline__ void euler_dir(double g, const double* U, double ir, double pr, double a, double b,
                                          double* Fd) {
    const double un = (U[1] * a + U[2] * b) * ir;
    Fd[0] = U[1] * a + U[2] * b;
    Fd[1] = fma(U[1], un, pr * a);
    Fd[2] = fma(U[2], un, pr * b);
    Fd[3] = un * (U[3] + pr);
}

__device__ __forceinline__ double pressure(double g, const double* U, double ir) {
    return (g - 1.0) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) * ir);
}

// Roe (1981) normal flux with the Harten-Hyman (1983) entropy fix on the acoustic waves.
// The central part 0.5 (F(UL) + F(UR)).n is formed from the two directional fluxes
// directly, so only four partial sums stay live instead of both F and G of both states.
// Square root of a positive normal double: two Goldschmidt steps from the hardware
// reciprocal-square-root estimate (no scaling for denormals, zero or infinity), for the
// Roe flux's five square roots per point (densities and sound speeds of physical states):
// about 6 instructions fewer than the general square root, 2-5% of a P1 Roe stage.
__device__ __forceinline__ double fsqrt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double gg = x * y, h = 0.5 * y;
    const double r = fma(-gg, h, 0.5);
    gg = fma(gg, r, gg);
    h = fma(h, r, h);
    double d = fma(-gg, gg, x);
    gg = fma(d, h, gg);
    d = fma(-gg, gg, x);
    return fma(d, h, gg);
}

__device__ __forceinline__ void roe_flux(double g, const double* UL, const double* UR, double nx, double ny,
                                         double* Fh) {
    const double rl = UL[0], irl = frcp(rl), ul = UL[1] * irl, vl = UL[2] * irl;
    const double rr = UR[0], irr = frcp(rr), ur = UR[1] * irr, vr = UR[2] * irr;
    const double pl = (g - 1) * (UL[3] - 0.5 * rl * (ul * ul + vl * vl));
    const double pr = (g - 1) * (UR[3] - 0.5 * rr * (ur * ur + vr * vr));
    double FL[4], FR[4];
    euler_dir(g, UL, irl, pl, nx, ny, FL);
    euler_dir(g, UR, irr, pr, nx, ny, FR);
#pragma unroll
    for (int v = 0; v < 4; ++v) Fh[v] = FL[v] + FR[v];
    const double hl = (UL[3] + pl) * irl, hr = (UR[3] + pr) * irr;
    // Roe averages with the weight ratio s = sqrt(rho_R / rho_L): (a_L + s a_R) / (1 + s),
    // and sqrt(rho_L rho_R) = rho_L s (one square root instead of two)
    const double sw = fsqrt(rr * irl), isum = frcp(1.0 + sw);
    const double rt = rl * sw;
    const double ut = (ul + sw * ur) * isum, vt = (vl + sw * vr) * isum, ht = (hl + sw * hr) * isum;
    const double q2 = ut * ut + vt * vt;
    const double c2 = (g - 1) * (ht - 0.5 * q2);
    // c and 1/c^2 from one reciprocal square root (two Newton steps)
    double ict = __builtin_amdgcn_rsq(c2);
    ict = ict * fma(-0.5 * c2 * ict, ict, 1.5);
    ict = ict * fma(-0.5 * c2 * ict, ict, 1.5);
    const double ct = c2 * ict, ic2 = ict * ict;
    const double unt = ut * nx + vt * ny, utt = -ut * ny + vt * nx;
    const double dp = pr - pl, dr = rr - rl;
    const double unl = ul * nx + vl * ny, unr = ur * nx + vr * ny;
    const double dun = unr - unl;
    const double dut = (-ur * ny + vr * nx) - (-ul * ny + vl * nx);
    const double a1 = 0.5 * (dp - rt * ct * dun) * ic2;
    const double a2 = dr - dp * ic2;
    const double a3 = rt * dut;
    const double a4 = 0.5 * (dp + rt * ct * dun) * ic2;
    // Harten-Hyman entropy fix on the acoustic waves (vanishes across shocks)
    const double cl = fsqrt(g * pl * irl), cr = fsqrt(g * pr * irr);
    double l1 = fabs(unt - ct), l4 = fabs(unt + ct);
    const double l2 = fabs(unt);
    const double d1 = fmax(0.0, fmax((unt - ct) - (unl - cl), (unr - cr) - (unt - ct)));
    const double d4 = fmax(0.0, fmax((unt + ct) - (unl + cl), (unr + cr) - (unt + ct)));
    if (l1 < d1) l1 = (l1 * l1 + d1 * d1) * (0.5 * frcp(d1));
    if (l4 < d4) l4 = (l4 * l4 + d4 * d4) * (0.5 * frcp(d4));
    const double b1 = l1 * a1, b2 = l2 * a2, b3 = l2 * a3, b4 = l4 * a4;
    Fh[0] = 0.5 * (Fh[0] - (b1 + b2 + b4));
    Fh[1] = 0.5 * (Fh[1] - (b1 * (ut - ct * nx) + b2 * ut + b3 * (-ny) + b4 * (ut + ct * nx)));
    Fh[2] = 0.5 * (Fh[2] - (b1 * (vt - ct * ny) + b2 * vt + b3 * nx + b4 * (vt + ct * ny)));
    Fh[3] = 0.5 * (Fh[3] - (b1 * (ht - unt * ct) + b2 * 0.5 * q2 + b3 * utt + b4 * (ht + unt * ct)));
}

struct TileInfo {
    int tile;
    int mesh;
    bool first;
    DevState* st;
    double cfl;
};

// Workgroup iteration `it` -> tile.  Grids are multiples of 8 and a persistent
// workgroup advances `it` by the grid size, so it % 8 (its XCD under round-robin
// dispatch) never changes: each XCD sweeps one contiguous block of tiles.
__device__ __forceinline__ bool tile_of_iter(const StageArgs& a, int it, int nit, TileInfo& ti) {
    int tile = it;
    if (a.xcd_remap) {
        const int per = nit >> 3;  // nit is a multiple of 8
        tile = (it & 7) * per + (it >> 3);
    }
    if (tile >= a.ntiles) return false;
    ti.tile = tile;
    if (a.tile_mesh) {
        const int tm = a.tile_mesh[tile];
        ti.mesh = tm >> 1;
        ti.first = tm & 1;
        ti.st = a.st_batch + ti.mesh;
        ti.cfl = a.cfl_batch[ti.mesh];
    } else {
        ti.mesh = 0;
        ti.first = tile == 0;
        ti.st = a.st;
        ti.cfl = a.cfl;
    }
    return true;
}

__device__ __forceinline__ int iter_count(const StageArgs& a) {
    return a.xcd_remap ? (a.ntiles + 7) / 8 * 8 : a.ntiles;
}

// dt of the current step, identical in every workgroup of the mesh; done = end time reached
__device__ __forceinline__ double stage_dt(const StageArgs& a, const TileInfo& ti, bool& done) {
    double dt;
    if (a.dt_mode == DTM_DEVICE) dt = ti.cfl / __longlong_as_double((long long)ti.st->ws[a.slot]);
    else dt = a.dt_host;
    done = false;
    if (a.t_end > 0) {
        const double rem = a.t_end - ti.st->t[a.slot];
        if (!(rem > 0)) done = true;
        else if (rem < dt) dt = rem;
    }
    return dt;
}

// This tile's pair of residual maxima (stage 2), or NULL when they are not kept.
__device__ __forceinline__ double* tile_residual(const StageArgs& a, const TileInfo& ti) {
    return a.res_part ? a.res_part + 2 * ti.tile : nullptr;
}

// Neighbour prefetch: the neighbour's modes are loaded one edge ahead (during the volume
// integral for the first edge), so their latency hides behind the flux arithmetic.  It
// costs 8*4*Np bytes of registers per lane, affordable up to P2.
template <int P>
struct Prefetch {
    static constexpr bool on = P <= 2;
};

template <int Np>
__device__ __forceinline__ void load_modes(const double* __restrict__ U, int n, double (&c)[4][Np]) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int i = 0; i < Np; ++i) c[v][i] = tmx(U, v * Np + i, 4 * Np, n);
}

// Element geometry: the rows of the inverse Jacobian and the area (edge normals and lengths
// are derived from them, edge_normal).
struct ElemGeo {
    double xix, xiy, etx, ety, area;
};

// Edge k's outward unit normal and length.  For the affine map of a CCW triangle (edge k
// opposite vertex k) the outward normal times the length is n_k l_k = 2A s_k with
// s_0 = xi + eta, s_1 = -xi, s_2 = -eta (xi = (xi_x, xi_y), eta = (eta_x, eta_y): the rows
// of the inverse Jacobian, 2A its determinant); so n_k = s_k / |s_k|, l_k = 2A |s_k|.
__device__ __forceinline__ void edge_frame(double sx, double sy, double area, double& nx, double& ny, double& lk) {
    const double q = sqrt(sx * sx + sy * sy);
    const double iq = frcp(q);
    nx = sx * iq;
    ny = sy * iq;
    lk = 2.0 * area * q;
}

__device__ __forceinline__ void edge_normal(const ElemGeo& eg, int k, double& nx, double& ny, double& lk) {
    const double sx = k == 0 ? eg.xix + eg.etx : (k == 1 ? -eg.xix : -eg.etx);
    const double sy = k == 0 ? eg.xiy + eg.ety : (k == 1 ? -eg.xiy : -eg.ety);
    edge_frame(sx, sy, eg.area, nx, ny, lk);
}

// Neighbour sources of the element core.
// NB_LDS_LEAN reads the LDS image like NB_LDS but evaluates one edge point at a time
// (fewer live registers, less instruction-level parallelism).
enum { NB_GLOBAL = 0, NB_LDS = 1, NB_LDS_LEAN = 2 };

// Residual, SSP-RK3 update and store of one element, from its modes Uo (already in
// registers), its geometry and its three neighbour codes.  Neighbour traces come from
// global memory (NB_GLOBAL, prefetched one edge ahead up to P2; the step-start state of
// stages 1 and 2 is then prefetched into the same registers during the last edge) or
// from the tile's LDS image (NB_LDS, slots of a.nloc).  Stage 2 folds the element's CFL
// wave speed and residual into ws_l, r0_l, r1_l (running maxima).
// Work registers of one element between its residual and its update.
template <int P>
struct ElemWork {
    double R[4][Deg<P>::Np];   // residual (mass not yet applied)
    double Nb[4][Deg<P>::Np];  // neighbour modes; with prefetch, the step-start modes afterwards
    double per;                // perimeter
};

// Volume integral sum_q w_q A (F . grad psi_i), added into R.  The point loop stays rolled:
// each iteration reads one row of the reference operators into scalar registers instead of
// keeping every table live.
template <int P>
__device__ __forceinline__ void element_volume(const StageArgs& a, const double (&Uo)[4][Deg<P>::Np],
                                               const ElemGeo& eg, double (&R)[4][Deg<P>::Np]) {
    constexpr int Np = Deg<P>::Np, Nqv = Deg<P>::Nqv;
    const DevTables& T = c_tab[P - 1];
    const double g = a.gamma;
#pragma unroll 1
    for (int q = 0; q < Nqv; ++q) {
        double uq[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double s = Uo[v][0];  // psi_0 = 1 (the mean)
#pragma unroll
            for (int i = 1; i < Np; ++i) s = fma(Uo[v][i], T.Vv[q][i], s);
            uq[v] = s;
        }
        // contravariant fluxes w A F.(xi_x, xi_y) and w A F.(eta_x, eta_y)
        const double ir = frcp(uq[0]);
        const double pr = pressure(g, uq, ir);
        const double wa = T.wv[q] * eg.area;
        double Fx[4], Fe[4];
        euler_dir(g, uq, ir, pr, wa * eg.xix, wa * eg.xiy, Fx);
        euler_dir(g, uq, ir, pr, wa * eg.etx, wa * eg.ety, Fe);
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int i = 1; i < Np; ++i) R[v][i] = fma(Fx[v], T.Dxi[q][i], fma(Fe[v], T.Deta[q][i], R[v][i]));
    }
}

template <int P, int FLUX, int NBSRC>
__device__ __forceinline__ void element_residual(const StageArgs& a, int e, const double (&Uo)[4][Deg<P>::Np],
                                                 const ElemGeo& eg, int w0, int w1, int w2, const double* sVe,
                                                 const double* sU, ElemWork<P>& W, const double* sGe = nullptr) {
    using D = Deg<P>;
    constexpr int Np = D::Np, Nqe = D::Nqe;
    constexpr bool PF = NBSRC == NB_GLOBAL && Prefetch<P>::on;
    const DevTables& T = c_tab[P - 1];
    const double g = a.gamma;
    auto& R = W.R;
    auto& Nb = W.Nb;
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int i = 0; i < Np; ++i) R[v][i] = 0.0;
    if (PF && w0 >= 0) load_modes<Np>(a.U, w0 >> 2, Nb);

    element_volume<P>(a, Uo, eg, R);

    // ---- surface integral, element-owned -----------------------------------------
    // Both loops stay rolled: one quadrature point's traces and flux are live at a time.
    // Edge k's direction s_k (edge_normal) for k = 1, 2 is re-read one edge ahead (L2 hits,
    // or the LDS copy sGe: rows of 256 doubles) rather than keeping the inverse Jacobian
    // live across the loop.
    double sx_n = eg.xix + eg.etx, sy_n = eg.xiy + eg.ety, per = 0.0;
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {
        const int w = k == 0 ? w0 : (k == 1 ? w1 : w2);
        const double sx = sx_n, sy = sy_n;
        if (k < 2) {
            sx_n = -(sGe ? sGe[(2 * k) * 256] : tmx(a.geo, 2 * k, GEO_ROWS, e));
            sy_n = -(sGe ? sGe[(2 * k + 1) * 256] : tmx(a.geo, 2 * k + 1, GEO_ROWS, e));
        }
        double nx, ny, lk;
        edge_frame(sx, sy, eg.area, nx, ny, lk);
        per += lk;
        // neighbour modes: LDS slot, registers (prefetched), or a load now
        const double* src = nullptr;
        if (w >= 0) {
            if (NBSRC != NB_GLOBAL) src = sU + (w >> 2);
            else if (!PF) load_modes<Np>(a.U, w >> 2, Nb);
        }
        constexpr int S = staged_stride(P);
        const bool refl = w == -RKDG_BC_REFLECTION;
        const double cxx = refl ? ny * ny - nx * nx : 1.0, cyy = refl ? nx * nx - ny * ny : 1.0;
        const double cxy = refl ? 2.0 * nx * ny : 0.0;
        // neighbour (or ghost) cell mean
        double mR[4];
        if (w >= 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) mR[v] = NBSRC != NB_GLOBAL ? src[(v * Np) * S] : Nb[v][0];
        } else if (w == -RKDG_BC_FARFIELD) {
            // kernel_boundaryCondition CELL_FARFIELD (src/kernelfunctions.cu:184-201)
#pragma unroll
            for (int v = 0; v < 4; ++v) mR[v] = a.fs[v];
        } else {
            // CELL_REFLECTION (232-268) mirrors the velocity about the edge normal;
            // CELL_OUTFLOW (203-230) copies the interior trace
            mR[0] = Uo[0][0];
            mR[3] = Uo[3][0];
            mR[1] = Uo[1][0] * cxx - cxy * Uo[2][0];
            mR[2] = Uo[2][0] * cyy - cxy * Uo[1][0];
        }
        double alpha = 0.0;
        if (FLUX == RKDG_FLUX_LF) {
            // kernel_calculateLFCoeff (src/kernelfunctions.cu:370-387)
            const double rho = 0.5 * (Uo[0][0] + mR[0]);
            const double ir = frcp(rho);
            const double u = 0.5 * (Uo[1][0] + mR[1]) * ir;
            const double v = 0.5 * (Uo[2][0] + mR[2]) * ir;
            const double rE = 0.5 * (Uo[3][0] + mR[3]);
            const double p = (g - 1.0) * (rE - 0.5 * rho * (u * u + v * v));
            alpha = fabs(u * nx + v * ny) + sqrt(g * p * ir);
        }
        const double hl = 0.5 * lk;
        const double* tb = sVe + (w & 3) * (Nqe * Np);
        if constexpr (NBSRC == NB_LDS && FLUX == RKDG_FLUX_LF && P <= 2) {
            // Registers are not the occupancy limit here (the LDS image is), so all points
            // of the edge are evaluated together for instruction-level parallelism: the
            // neighbour trace is streamed from LDS over its modes into Nqe x 4 chains.
            double uRa[Nqe][4];
            if (w >= 0) {
#pragma unroll
                for (int i = 0; i < Np; ++i) {
                    double c[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) c[v] = src[(v * Np + i) * S];
#pragma unroll
                    for (int j = 0; j < Nqe; ++j) {
                        const double t = tb[(Nqe - 1 - j) * Np + i];
#pragma unroll
                        for (int v = 0; v < 4; ++v) uRa[j][v] = i == 0 ? c[v] : fma(c[v], t, uRa[j][v]);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < Nqe; ++j) {
                double uL[4], uR[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double s = Uo[v][0];  // psi_0 = 1
#pragma unroll
                    for (int i = 1; i < Np; ++i) s = fma(Uo[v][i], T.Ve[k][j][i], s);
                    uL[v] = s;
                }
                if (w >= 0) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) uR[v] = uRa[j][v];
                } else if (w == -RKDG_BC_FARFIELD) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) uR[v] = a.fs[v];
                } else {
                    uR[0] = uL[0];
                    uR[3] = uL[3];
                    uR[1] = uL[1] * cxx - cxy * uL[2];
                    uR[2] = uL[2] * cyy - cxy * uL[1];
                }
                double Fh[4];
                if (FLUX == RKDG_FLUX_LF) {
                    double FL[4], FR[4];
                    const double irl = frcp(uL[0]), irr = frcp(uR[0]);
                    euler_dir(g, uL, irl, pressure(g, uL, irl), nx, ny, FL);
                    euler_dir(g, uR, irr, pressure(g, uR, irr), nx, ny, FR);
#pragma unroll
                    for (int v = 0; v < 4; ++v) Fh[v] = 0.5 * (FL[v] + FR[v] - alpha * (uR[v] - uL[v]));
                } else {
                    roe_flux(g, uL, uR, nx, ny, Fh);
                }
                const double wj = T.we[j] * hl;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const double f = wj * Fh[v];
#pragma unroll
                    for (int i = 0; i < Np; ++i) R[v][i] = fma(-f, T.Ve[k][j][i], R[v][i]);
                }
            }
        } else {
    #pragma unroll 1
            for (int j = 0; j < Nqe; ++j) {
                double uL[4], uR[4];
    #pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double s = Uo[v][0];  // psi_0 = 1
    #pragma unroll
                    for (int i = 1; i < Np; ++i) s = fma(Uo[v][i], T.Ve[k][j][i], s);
                    uL[v] = s;
                }
                if (w >= 0) {
                    // neighbour trace at the mirrored point
                    const double* t = tb + (Nqe - 1 - j) * Np;
    #pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        double s = NBSRC != NB_GLOBAL ? src[(v * Np) * S] : Nb[v][0];  // psi_0 = 1
    #pragma unroll
                        for (int i = 1; i < Np; ++i) s = fma(NBSRC != NB_GLOBAL ? src[(v * Np + i) * S] : Nb[v][i], t[i], s);
                        uR[v] = s;
                    }
                } else if (w == -RKDG_BC_FARFIELD) {
    #pragma unroll
                    for (int v = 0; v < 4; ++v) uR[v] = a.fs[v];
                } else {
                    uR[0] = uL[0];
                    uR[3] = uL[3];
                    uR[1] = uL[1] * cxx - cxy * uL[2];
                    uR[2] = uL[2] * cyy - cxy * uL[1];
                }
                double Fh[4];
                if (FLUX == RKDG_FLUX_LF) {
                    // 0.5 (F(uL).n + F(uR).n - alpha (uR - uL)), kernel_calculateFlux's formula
                    double FL[4], FR[4];
                    const double irl = frcp(uL[0]), irr = frcp(uR[0]);
                    euler_dir(g, uL, irl, pressure(g, uL, irl), nx, ny, FL);
                    euler_dir(g, uR, irr, pressure(g, uR, irr), nx, ny, FR);
    #pragma unroll
                    for (int v = 0; v < 4; ++v) Fh[v] = 0.5 * (FL[v] + FR[v] - alpha * (uR[v] - uL[v]));
                } else {
                    roe_flux(g, uL, uR, nx, ny, Fh);
                }
                const double wj = T.we[j] * hl;
    #pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const double f = wj * Fh[v];
    #pragma unroll
                    for (int i = 0; i < Np; ++i) R[v][i] = fma(-f, T.Ve[k][j][i], R[v][i]);
                }
            }
        }
        // next edge's neighbour modes (after the last edge: the step-start state), issued
        // once this edge no longer needs the registers
        if (k < 2) {
            const int wn = k == 0 ? w1 : w2;
            if (PF && wn >= 0) load_modes<Np>(a.U, wn >> 2, Nb);
        } else if (PF && a.stage > 0) {
            load_modes<Np>(a.U0, e, Nb);
        }
    }
    W.per = per;
}

// SSP-RK3 combination, mass division and store of one element (stage 2 also folds its
// CFL wave speed and residual into ws_l, r0_l, r1_l, running maxima).
template <int P, int NBSRC, bool U0_IN_NB = false>
__device__ __forceinline__ void element_update(const StageArgs& a, int e, const double (&Uo)[4][Deg<P>::Np],
                                               const ElemGeo& eg, ElemWork<P>& W, double dt, bool dtf,
                                               DevState* st, double& ws_l, double& r0_l, double& r1_l) {
    constexpr int Np = Deg<P>::Np;
    constexpr bool PF = NBSRC == NB_GLOBAL && Prefetch<P>::on;
    const DevTables& T = c_tab[P - 1];
    const double g = a.gamma;
    auto& R = W.R;
    auto& Nb = W.Nb;
    const double per = W.per;
    // ---- SSP-RK3 stage (src/kernelfunctions.cu:640-748) ---------------------------
    // The step-start modes are all loaded before the first store: U0 and Uout are the same
    // buffer in stage 2, so a load placed after a store could not be moved ahead of it and
    // every mode would cost a full memory round trip.
    // (buffer-resource addressing: the 4 Np rows cost no per-row address registers)
    const __amdgpu_buffer_rsrc_t rO = tm_rsrc(a.Uout, a.ld, 4 * Np);
    const unsigned vo = tm_voff(e, 4 * Np);
    if (!PF && !U0_IN_NB && a.stage > 0) {
        const __amdgpu_buffer_rsrc_t r0 = tm_rsrc(a.U0, a.ld, 4 * Np);
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int i = 0; i < Np; ++i) Nb[v][i] = tm_load(r0, vo, v * Np + i);
    }
    const double ia = frcp(eg.area);
    double mean[4];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int i = 0; i < Np; ++i) {
            const double r = R[v][i] * (T.minv[i] * ia);
            double out;
            if (a.stage == 0) {
                out = Uo[v][i] + dt * r;
            } else {
                const double o0 = Nb[v][i];
                if (a.stage == 1) {
                    out = 0.75 * o0 + 0.25 * Uo[v][i] + 0.25 * dt * r;
                } else {
                    out = (1.0 / 3.0) * o0 + (2.0 / 3.0) * Uo[v][i] + (2.0 / 3.0) * dt * r;
                    if (i == 0 && dtf) {
                        if (v == 0) r0_l = fmax(r0_l, fabs(out - o0));
                        if (v == 3) r1_l = fmax(r1_l, fabs(out - o0));
                    }
                }
            }
            if (i == 0) mean[v] = out;
            tm_store(rO, vo, v * Np + i, out);
        }

    if (a.stage == 2) {
        // next step's CFL wave speed on the new means (kernel_getTimeStep's formula,
        // src/kernelfunctions.cu:19-25)
        const double rho = mean[0], u = mean[1] / rho, v = mean[2] / rho;
        const double p = (g - 1.0) * (mean[3] - 0.5 * rho * (u * u + v * v));
        if (!(rho > 0.0) || !(p > 0.0)) atomicMin(&st->err_pos, e);
        const double c = sqrt(g * p / rho);
        const double s = (sqrt(u * u + v * v) + c) * per / eg.area;
        if (dtf && s > 0.0) ws_l = fmax(ws_l, s);
    }
}


template <int P, int FLUX, int NBSRC>
__device__ __forceinline__ void element_core(const StageArgs& a, int e, const double (&Uo)[4][Deg<P>::Np],
                                             const ElemGeo& eg, int w0, int w1, int w2, double dt, bool dtf,
                                             const double* sVe, const double* sU, DevState* st, double& ws_l,
                                             double& r0_l, double& r1_l, const double* sGe = nullptr) {
    ElemWork<P> W;
    element_residual<P, FLUX, NBSRC>(a, e, Uo, eg, w0, w1, w2, sVe, sU, W, sGe);
    element_update<P, NBSRC>(a, e, Uo, eg, W, dt, dtf, st, ws_l, r0_l, r1_l);
}

// Per-mesh bookkeeping done once per stage by the workgroup owning the mesh's first tile:
// stage 0 fixes the step's dt and clears the reductions, stage 2 advances the time.
__device__ __forceinline__ void stage_bookkeeping(const StageArgs& a, const TileInfo& ti, double dt, bool done) {
    if (ti.first && threadIdx.x == 0) {
        if (a.stage == 0 && !done) {
            ti.st->ws[1 - a.slot] = 0ull;
            ti.st->dt[a.slot] = dt;
        }
        if (a.stage == 2) {
            ti.st->t[1 - a.slot] = done ? ti.st->t[a.slot] : ti.st->t[a.slot] + dt;
            if (done) ti.st->ws[1 - a.slot] = ti.st->ws[a.slot];
            else ti.st->nsteps += 1;
        }
    }
}

// One workgroup tile, operands loaded from global memory.  STAGED: the tile's own modes
// and its halo (the out-of-tile neighbours, listed per tile on the host) are first copied
// into LDS, and every neighbour trace is evaluated from LDS.
// Operands an element needs before its first edge: neighbour codes, geometry, dt flag.
struct ElemOps {
    int w0, w1, w2;
    ElemGeo eg;
    bool dtf;
};

__device__ __forceinline__ void load_element_ops(const StageArgs& a, const int* nb_ids, int e, ElemOps& o) {
    o.w0 = tmx(nb_ids, 0, NBR_ROWS, e);
    o.w1 = tmx(nb_ids, 1, NBR_ROWS, e);
    o.w2 = tmx(nb_ids, 2, NBR_ROWS, e);
    o.eg.area = tmx(a.geo, 4, GEO_ROWS, e);
    o.eg.xix = tmx(a.geo, 0, GEO_ROWS, e);
    o.eg.xiy = tmx(a.geo, 1, GEO_ROWS, e);
    o.eg.etx = tmx(a.geo, 2, GEO_ROWS, e);
    o.eg.ety = tmx(a.geo, 3, GEO_ROWS, e);
    o.dtf = a.stage == 2 && a.dtflag[e] != 0;
}

// Staged form.  The LDS image of a tile (its own modes, then its halo: the out-of-tile
// neighbours listed per tile on the host) is filled by the direct-to-LDS load engine: the
// own rows as 16-byte lanes (1 KiB per wave instruction), the halo as 4-byte lanes (two
// lanes per double, 32 halo elements per instruction).  No data passes through registers.
template <int P>
struct Staged {
    static constexpr int Np = Deg<P>::Np, S = staged_stride(P), B = staged_block(P), NR = 4 * Np;
    static constexpr int NW = B / 64, PARTS = B / 128, HCH = (S - B + 31) / 32;
};

// This lane's halo ids of a tile (element hb * 32 + lane / 2 of the halo, chunk hb).
template <int P>
__device__ __forceinline__ int staged_halo_ids(const StageArgs& a, int tile, int (&hid)[Staged<P>::HCH],
                                               int tid = threadIdx.x) {
    const int lane = tid & 63;
    const int h0 = a.hoff[tile], nh = a.hoff[tile + 1] - h0;
#pragma unroll
    for (int hb = 0; hb < Staged<P>::HCH; ++hb) {
        const int h = hb * 32 + (lane >> 1);
        hid[hb] = h < nh ? a.hids[h0 + h] : 0;
    }
    return nh;
}

// Issues this wave's share of the direct loads of a tile's image.
template <int P>
__device__ __forceinline__ void staged_fill(const StageArgs& a, int tile, int nh, const int (&hid)[Staged<P>::HCH],
                                            double* sU, int tid = threadIdx.x) {
    using L = Staged<P>;
    const int lane = tid & 63, wave = tid >> 6;
    char* lds = (char*)sU;
    // the tile's rows inside its layout tile: row r starts 2 KiB after row r - 1
    const long long e0 = (long long)tile * L::B;
    const double* own = a.U + ((e0 >> 8) * L::NR) * 256 + (e0 & 255);
    for (int c = wave; c < L::NR * L::PARTS; c += L::NW) {
        const int r = c / L::PARTS, part = c - r * L::PARTS;
        const char* src = (const char*)(own + r * 256) + part * 1024 + lane * 16;
        __builtin_amdgcn_global_load_lds((const void*)src,
                                         (__attribute__((address_space(3))) void*)(lds + r * L::S * 8 + part * 1024), 16, 0, 0);
    }
#pragma unroll
    for (int hb = 0; hb < L::HCH; ++hb) {
        if (hb * 32 >= nh) break;  // uniform across the workgroup
        const bool on = hb * 32 + (lane >> 1) < nh;
        const int h = hid[hb];
        const char* src0 = (const char*)(a.U + ((size_t)(h >> 8) * L::NR) * 256 + (h & 255)) + (lane & 1) * 4;
        for (int r = wave; r < L::NR; r += L::NW) {
            if (on)
                __builtin_amdgcn_global_load_lds((const void*)(src0 + (size_t)r * 2048),
                                                 (__attribute__((address_space(3))) void*)(lds + (r * L::S + L::B + hb * 32) * 8),
                                                 4, 0, 0);
        }
    }
}

// One workgroup tile, operands loaded from global memory (STAGED: neighbour traces from
// the tile's LDS image).
template <int P, int FLUX, bool STAGED>
__device__ __forceinline__ void stage_tile(const StageArgs& a, const TileInfo& ti, const double* sVe, double* sU,
                                           int tid) {
    constexpr int Np = Deg<P>::Np;
    bool done;
    const double dt = stage_dt(a, ti, done);
    stage_bookkeeping(a, ti, dt, done);
    if (done) return;

    const int e = ti.tile * blockDim.x + tid;
    double ws_l = 0.0, r0_l = 0.0, r1_l = 0.0;
    double Uo[4][Np];
    ElemOps o{};
    if (STAGED) {
        int hid[Staged<P>::HCH];
        const int nh = staged_halo_ids<P>(a, ti.tile, hid, tid);
        staged_fill<P>(a, ti.tile, nh, hid, sU, tid);
        if (e < a.E) load_element_ops(a, a.nloc, e, o);  // same round trip as the image
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int i = 0; i < Np; ++i) Uo[v][i] = sU[(v * Np + i) * Staged<P>::S + tid];
    } else if (e < a.E) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int i = 0; i < Np; ++i) Uo[v][i] = tmx(a.U, v * Np + i, 4 * Np, e);
        load_element_ops(a, a.nbr, e, o);
    }

    if (e < a.E)
        element_core<P, FLUX, STAGED ? (P == 1 && FLUX == RKDG_FLUX_LF ? NB_LDS_LEAN : NB_LDS) : NB_GLOBAL>(a, e, Uo, o.eg, o.w0, o.w1, o.w2, dt, o.dtf, sVe, sU,
                                                          ti.st, ws_l, r0_l, r1_l);
    if (a.stage == 2)
        block_stage_max(ws_l, r0_l, r1_l, &ti.st->ws[1 - a.slot], tile_residual(a, ti));
}

// Persistent staged form with the next tile's image in flight during the current tile's
// SSP-RK3 update.  Loads complete in order, so the order of issue decides what a wait waits
// for: once every wave has finished its residual (a barrier), the update's step-start modes
// are requested first, then the next tile's image and operands; the update then waits only
// for its own loads — the image loads behind them have a fixed count per wave (every own-row
// part and every halo chunk is issued, lanes past the halo's end reading element 0 into
// unused slots; the operands are read at an index clamped into the mesh), so the compiler's
// wait counts them instead of draining the queue.  The next tile's halo ids are requested
// before the residual, so their values have arrived when the image addresses are formed.
// The last tile of a workgroup runs without the prefetch (a separate instantiation of the
// body, so that no path through the prefetching body lacks its loads).
// Byte offset of halo element h's first row in the tile-major state, plus the lane's half of
// the double (two lanes per double).
template <int P>
__device__ __forceinline__ unsigned halo_offset(int h) {
    using L = Staged<P>;
    return (unsigned)((((h >> 8) * L::NR) * 256 + (h & 255)) * 8) + (unsigned)(threadIdx.x & 1) * 4u;
}

template <int P>
__device__ __forceinline__ void staged_fill_fixed(const StageArgs& a, int tile, const unsigned (&hoff)[Staged<P>::HCH],
                                                  double* sU) {
    using L = Staged<P>;
    static_assert((L::NR * L::PARTS) % L::NW == 0 && L::NR % L::NW == 0, "rows divide among the waves");
    static_assert(L::S - L::B == 32 * L::HCH, "every halo chunk lands inside the halo slots");
    // Addresses as a uniform base (scalar registers: the wave index, rows and LDS slots) plus
    // one 32-bit lane offset, so that the loads cost almost no vector registers at the point
    // of the update where they are issued.
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    char* lds = (char*)sU;
    const char* U = (const char*)a.U;
    const long long e0 = (long long)tile * L::B;
    const char* own = U + (((e0 >> 8) * L::NR) * 256 + (e0 & 255)) * 8;
    const unsigned lo = (unsigned)lane * 16u;
#pragma unroll
    for (int j = 0; j < L::NR * L::PARTS / L::NW; ++j) {
        const int c = wave + j * L::NW;
        const int r = c / L::PARTS, part = c - r * L::PARTS;
        __builtin_amdgcn_global_load_lds((const void*)(own + r * 2048 + part * 1024 + lo),
                                         (__attribute__((address_space(3))) void*)(lds + r * L::S * 8 + part * 1024), 16, 0, 0);
    }
#pragma unroll
    for (int hb = 0; hb < L::HCH; ++hb) {
        const unsigned ho = hoff[hb];
#pragma unroll
        for (int j = 0; j < L::NR / L::NW; ++j) {
            const int r = wave + j * L::NW;
            __builtin_amdgcn_global_load_lds((const void*)(U + (size_t)r * 2048 + ho),
                                             (__attribute__((address_space(3))) void*)(lds + (r * L::S + L::B + hb * 32) * 8),
                                             4, 0, 0);
        }
    }
}

// The next tile's operands as loaded (the dt flag raw: a test of it here would wait for
// every load in flight), read at an index clamped into the mesh.
struct ElemOpsRaw {
    int w0, w1, w2;
    ElemGeo eg;
    int flag;
};

__device__ __forceinline__ void load_element_ops_clamped(const StageArgs& a, int e, ElemOpsRaw& o) {
    e = min(e, a.E - 1);
    o.w0 = tmx(a.nloc, 0, NBR_ROWS, e);
    o.w1 = tmx(a.nloc, 1, NBR_ROWS, e);
    o.w2 = tmx(a.nloc, 2, NBR_ROWS, e);
    o.eg.area = tmx(a.geo, 4, GEO_ROWS, e);
    o.eg.xix = tmx(a.geo, 0, GEO_ROWS, e);
    o.eg.xiy = tmx(a.geo, 1, GEO_ROWS, e);
    o.eg.etx = tmx(a.geo, 2, GEO_ROWS, e);
    o.eg.ety = tmx(a.geo, 3, GEO_ROWS, e);
    o.flag = a.dtflag[e];
}

// The next valid tile of this workgroup after iteration `it` (it advances).
__device__ __forceinline__ bool next_tile(const StageArgs& a, int& it, int nit, TileInfo& t) {
    for (it += gridDim.x; it < nit; it += gridDim.x)
        if (tile_of_iter(a, it, nit, t)) return true;
    return false;
}

template <int P, int FLUX, bool PF>
__device__ __forceinline__ void pipe_tile(const StageArgs& a, const TileInfo& ti, const ElemOpsRaw& o, const double* sVe,
                                          double* sU, const TileInfo& tn, ElemOpsRaw& on) {
    using L = Staged<P>;
    constexpr int Np = Deg<P>::Np;
    const int tid = threadIdx.x;
    bool done;
    const double dt = stage_dt(a, ti, done);
    stage_bookkeeping(a, ti, dt, done);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();  // the tile's image and operands have landed (and sVe on entry)
    const int e = ti.tile * L::B + tid;
    const bool live = e < a.E && !done;
    const bool dtf = a.stage == 2 && o.flag != 0;
    double Uo[4][Np];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int i = 0; i < Np; ++i) Uo[v][i] = sU[(v * Np + i) * L::S + tid];
    int hid[L::HCH];
    if constexpr (PF) staged_halo_ids<P>(a, tn.tile, hid);
    ElemWork<P> W;
    if (live) element_residual<P, FLUX, NB_LDS>(a, e, Uo, o.eg, o.w0, o.w1, o.w2, sVe, sU, W);
    // the next image's halo offsets are formed here, before the step-start loads are issued,
    // so that the wait for the halo ids cannot stop at those loads
    unsigned hoff[L::HCH];
    if constexpr (PF) {
#pragma unroll
        for (int hb = 0; hb < L::HCH; ++hb) {
            hoff[hb] = halo_offset<P>(hid[hb]);
            asm volatile("" ::"v"(hoff[hb]));  // formed here, not after the loads below
        }
    }
    if constexpr (PF) load_element_ops_clamped(a, tn.tile * L::B + tid, on);
    __syncthreads();  // every wave is done with the image
    if (a.stage > 0) {
        const __amdgpu_buffer_rsrc_t r0 = tm_rsrc(a.U0, a.ld, 4 * Np);
        const unsigned vo = tm_voff(e, 4 * Np);
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int i = 0; i < Np; ++i) W.Nb[v][i] = tm_load(r0, vo, v * Np + i);
    }
    if constexpr (PF) staged_fill_fixed<P>(a, tn.tile, hoff, sU);
    double ws_l = 0.0, r0_l = 0.0, r1_l = 0.0;
    if (live) element_update<P, NB_LDS, true>(a, e, Uo, o.eg, W, dt, dtf, ti.st, ws_l, r0_l, r1_l);
    // (past the end time the tile keeps its last residual pair, as the staged kernel does;
    // `done` is the same for every workgroup of the mesh)
    if (a.stage == 2 && !done)
        block_stage_max(ws_l, r0_l, r1_l, &ti.st->ws[1 - a.slot], tile_residual(a, ti));
}

template <int P, int FLUX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_stage_pipe(StageArgs a) {
    using L = Staged<P>;
    constexpr int Np = Deg<P>::Np, Nqe = Deg<P>::Nqe;
    const DevTables& T = c_tab[P - 1];
    __shared__ double sVe[3 * Nqe * Np];
    extern __shared__ __align__(16) double sU[];
    for (int t = threadIdx.x; t < 3 * Nqe * Np; t += blockDim.x) {
        const int k = t / (Nqe * Np), r = t - k * (Nqe * Np), j = r / Np, i = r - j * Np;
        sVe[t] = T.Ve[k][j][i];
    }
    const int nit = iter_count(a);
    int it = (int)blockIdx.x - (int)gridDim.x;
    TileInfo ti;
    if (!next_tile(a, it, nit, ti)) return;
    ElemOpsRaw o;
    {
        int hid[L::HCH];
        staged_halo_ids<P>(a, ti.tile, hid);
        unsigned hoff[L::HCH];
#pragma unroll
        for (int hb = 0; hb < L::HCH; ++hb) hoff[hb] = halo_offset<P>(hid[hb]);
        staged_fill_fixed<P>(a, ti.tile, hoff, sU);
        load_element_ops_clamped(a, ti.tile * L::B + (int)threadIdx.x, o);
    }
    TileInfo tn;
    while (next_tile(a, it, nit, tn)) {  // uniform across the workgroup
        ElemOpsRaw on;
        pipe_tile<P, FLUX, true>(a, ti, o, sVe, sU, tn, on);
        ti = tn;
        o = on;
    }
    ElemOpsRaw unused;
    pipe_tile<P, FLUX, false>(a, ti, o, sVe, sU, ti, unused);
}

// Register-pipelined persistent staged form (P1): the next tile's own modes, halo modes and
// operands are loaded into registers while the current tile computes its residual and
// update, and written into the single LDS image at the top of the next iteration.  Loads
// complete in order, so a wait for any load also waits for every load issued before it:
// everything the current tile reads from memory (its dt, then the update's step-start
// modes, which come last) is issued before the prefetch or long after it, and the halo's
// indices run ahead of its data (the per-tile halo offsets three tiles ahead, the lane's
// halo element id two tiles ahead), so no wait inside an iteration stops at the prefetch.
// Halo lane layout: lane t holds halo element t % 128 (at most S - B = 96), rows t / 128 + 2j.
template <int P>
struct RpImage {
    static constexpr int NR = 4 * Deg<P>::Np, NH = NR / 2;
    double u[NR];    // own modes of the lane's element
    double hx[NH];   // rows of the lane's halo element
    double g[5];     // inverse Jacobian rows and area
    int w[3];        // neighbour codes (LDS slots)
    int dtf;         // dt-subset byte, raw (compared when used)
    int nh;          // halo elements of the tile
};

template <int P>
__device__ __forceinline__ void rp_data(const StageArgs& a, int tile, int nh, int hid, RpImage<P>& im) {
    using L = Staged<P>;
    constexpr int NR = RpImage<P>::NR, NH = RpImage<P>::NH;
    const int tid = threadIdx.x, hh = tid & 127, hr = tid >> 7;
    const int e = tile * L::B + tid;  // < ld: the last tile's padded columns lie inside the arrays
    const __amdgpu_buffer_rsrc_t rU = tm_rsrc(a.U, a.ld, NR);
    const unsigned vo = tm_voff(e, NR);
#pragma unroll
    for (int r = 0; r < NR; ++r) im.u[r] = tm_load(rU, vo, r);
    im.nh = nh;
    if (hh < nh) {
        const unsigned vh = tm_voff(hid, NR);
#pragma unroll
        for (int j = 0; j < NH; ++j) im.hx[j] = tm_load(rU, vh, hr + 2 * j);
    }
    // operands (also read for the padded lanes: their columns hold the padding's values)
    const __amdgpu_buffer_rsrc_t rG = tm_rsrc(a.geo, a.ld, GEO_ROWS);
    const unsigned vg = tm_voff(e, GEO_ROWS);
#pragma unroll
    for (int r = 0; r < GEO_ROWS; ++r) im.g[r] = tm_load(rG, vg, r);
#pragma unroll
    for (int k = 0; k < 3; ++k) im.w[k] = tmx(a.nloc, k, NBR_ROWS, e);
    im.dtf = a.dtflag[e];
}

// Lane's halo element id of a tile (0 when the lane holds none) from its halo offsets.
__device__ __forceinline__ int rp_halo_id(const StageArgs& a, int2 ho) {
    const int hh = threadIdx.x & 127;
    return hh < ho.y - ho.x ? a.hids[ho.x + hh] : 0;
}

__device__ __forceinline__ int2 rp_halo_off(const StageArgs& a, int tile) {
    return make_int2(a.hoff[tile], a.hoff[tile + 1]);
}

template <int P, int FLUX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_stage_rp(StageArgs a) {
    using L = Staged<P>;
    constexpr int Np = Deg<P>::Np, Nqe = Deg<P>::Nqe, NR = RpImage<P>::NR, NH = RpImage<P>::NH;
    const DevTables& T = c_tab[P - 1];
    __shared__ double sVe[3 * Nqe * Np];
    extern __shared__ __align__(16) double sU[];
    for (int t = threadIdx.x; t < 3 * Nqe * Np; t += blockDim.x) {
        const int k = t / (Nqe * Np), r = t - k * (Nqe * Np), j = r / Np, i = r - j * Np;
        sVe[t] = T.Ve[k][j][i];
    }
    const int nit = iter_count(a);
    const int G = gridDim.x, tid = threadIdx.x, hh = tid & 127, hr = tid >> 7;
    // tiles of this workgroup's iterations it, it + G, it + 2G, it + 3G (uniform)
    int it = blockIdx.x;
    TileInfo t0, t1, t2;
    bool h0 = it < nit && tile_of_iter(a, it, nit, t0);
    bool h1 = h0 && it + G < nit && tile_of_iter(a, it + G, nit, t1);
    bool h2 = h1 && it + 2 * G < nit && tile_of_iter(a, it + 2 * G, nit, t2);
    RpImage<P> im;
    int2 ho1 = make_int2(0, 0), ho2 = make_int2(0, 0);
    int hid1 = 0;
    if (h0) {
        const int2 ho0 = rp_halo_off(a, t0.tile);
        rp_data<P>(a, t0.tile, ho0.y - ho0.x, rp_halo_id(a, ho0), im);
    }
    if (h1) {
        ho1 = rp_halo_off(a, t1.tile);
        hid1 = rp_halo_id(a, ho1);
    }
    if (h2) ho2 = rp_halo_off(a, t2.tile);
    // one mesh: the step's dt and end-time state are the same for every tile of the launch
    bool done = false;
    const double dt = h0 ? stage_dt(a, t0, done) : 0.0;
    while (h0) {  // uniform across the workgroup
        // this tile's image from the registers (the previous tile is done with the LDS)
#pragma unroll
        for (int r = 0; r < NR; ++r) sU[r * L::S + tid] = im.u[r];
        if (hh < im.nh) {
#pragma unroll
            for (int j = 0; j < NH; ++j) sU[(hr + 2 * j) * L::S + L::B + hh] = im.hx[j];
        }
        // the inverse Jacobian rows, re-read by the edge loop (a load from memory there would
        // wait for the whole prefetch)
        double* sG = sU + NR * L::S;
#pragma unroll
        for (int r = 0; r < 4; ++r) sG[r * L::B + tid] = im.g[r];
        double Uo[4][Np];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int i = 0; i < Np; ++i) Uo[v][i] = im.u[v * Np + i];
        ElemOps o;
        o.eg = ElemGeo{im.g[0], im.g[1], im.g[2], im.g[3], im.g[4]};
        o.w0 = im.w[0]; o.w1 = im.w[1]; o.w2 = im.w[2];
        o.dtf = a.stage == 2 && im.dtf != 0;
        __syncthreads();  // the image (and sVe on entry) is complete
        // prefetch: the next tile's data, the halo ids of the one after, the offsets of the third
        TileInfo t3;
        const bool h3 = h2 && it + 3 * G < nit && tile_of_iter(a, it + 3 * G, nit, t3);
        if (h1) rp_data<P>(a, t1.tile, ho1.y - ho1.x, hid1, im);
        int hid2 = 0;
        if (h2) hid2 = rp_halo_id(a, ho2);
        int2 ho3 = make_int2(0, 0);
        if (h3) ho3 = rp_halo_off(a, t3.tile);
        const int e = t0.tile * L::B + tid;
        double ws_l = 0.0, r0_l = 0.0, r1_l = 0.0;
        if (e < a.E && !done)
            element_core<P, FLUX, NB_LDS>(a, e, Uo, o.eg, o.w0, o.w1, o.w2, dt, o.dtf, sVe, sU, t0.st, ws_l, r0_l, r1_l,
                                          sU + NR * L::S + tid);
        stage_bookkeeping(a, t0, dt, done);
        if (a.stage == 2 && !done)  // (done is uniform: one end time per mesh)
            block_stage_max(ws_l, r0_l, r1_l, &t0.st->ws[1 - a.slot], tile_residual(a, t0));
        __syncthreads();  // every wave is done with the image before the next one is written
        it += G;
        t0 = t1; t1 = t2; t2 = t3;
        h0 = h1; h1 = h2; h2 = h3;
        ho1 = ho2; ho2 = ho3;
        hid1 = hid2;
    }
}

// One workgroup per tile, operands from global memory (STAGED: neighbours from the
// tile's LDS image, passed as dynamic shared memory of 4 Np rows of staged_stride(P)
// doubles; the workgroup then has staged_block(P) lanes).
template <int P, int FLUX, bool STAGED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STAGED && P == 1 ? (FLUX == RKDG_FLUX_LF ? 4 : 3) : (STAGED && P <= 2 ? 2 : 1)))) k_stage(StageArgs a) {
    using D = Deg<P>;
    constexpr int Np = D::Np, Nqe = D::Nqe;
    const DevTables& T = c_tab[P - 1];
    __shared__ double sVe[3 * Nqe * Np];
    extern __shared__ __align__(16) double sU[];
    for (int t = threadIdx.x; t < 3 * Nqe * Np; t += blockDim.x) {
        const int k = t / (Nqe * Np), r = t - k * (Nqe * Np), j = r / Np, i = r - j * Np;
        sVe[t] = T.Ve[k][j][i];
    }
    __syncthreads();
    if (a.stagger > 0 && blockIdx.x >= (gridDim.x >> 1))
        for (int i = 0; i < a.stagger; ++i) __builtin_amdgcn_s_sleep(127);
    const int nit = iter_count(a);
    for (int it = blockIdx.x; it < nit; it += gridDim.x) {
        TileInfo ti;
        if (!tile_of_iter(a, it, nit, ti)) continue;  // uniform across the workgroup
        int tid = (int)threadIdx.x;
        // P1 staged tiles run four (LF, one edge point at a time) or three (Roe) waves per
        // SIMD: the lane index through an opaque copy per tile keeps the addresses derived
        // from it inside the tile instead of live (spilled)
        if constexpr (STAGED && P == 1) asm volatile("" : "+v"(tid));
        stage_tile<P, FLUX, STAGED>(a, ti, sVe, sU, tid);
        __syncthreads();  // LDS and the reduction scratch are reused by the next tile
    }
}

#include "stage_mx.inc"
#include "stage_px.inc"
#include "stage_pz.inc"

// CFL wave speed of the current state, max-reduced into ws[slot]
template <int P>
__global__ void __launch_bounds__(256) k_wavespeed(StageArgs a) {
    constexpr int Np = Deg<P>::Np;
    const int nit = iter_count(a);
    for (int it = blockIdx.x; it < nit; it += gridDim.x) {
        TileInfo ti;
        if (!tile_of_iter(a, it, nit, ti)) continue;
        const int e = ti.tile * blockDim.x + threadIdx.x;
        double s_l = 0.0;
        if (e < a.E) {
            const double g = a.gamma;
            const double rho = tmx(a.U, 0 * Np, 4 * Np, e);
            const double u = tmx(a.U, 1 * Np, 4 * Np, e) / rho, v = tmx(a.U, 2 * Np, 4 * Np, e) / rho;
            const double p = (g - 1.0) * (tmx(a.U, 3 * Np, 4 * Np, e) - 0.5 * rho * (u * u + v * v));
            if (!(rho > 0.0) || !(p > 0.0)) atomicMin(&ti.st->err_pos, e);
            const double c = sqrt(g * p / rho);
            const ElemGeo eg{tmx(a.geo, 0, GEO_ROWS, e), tmx(a.geo, 1, GEO_ROWS, e), tmx(a.geo, 2, GEO_ROWS, e),
                             tmx(a.geo, 3, GEO_ROWS, e), tmx(a.geo, 4, GEO_ROWS, e)};
            double per = 0.0;
            for (int k = 0; k < 3; ++k) {
                double nx, ny, lk;
                edge_normal(eg, k, nx, ny, lk);
                per += lk;
            }
            const double s = (sqrt(u * u + v * v) + c) * per / eg.area;
            if (a.dtflag[e] && s > 0.0) s_l = s;
        }
        block_stage_max(s_l, 0.0, 0.0, &ti.st->ws[a.slot], nullptr);
        __syncthreads();
    }
}

// Clears only the wave-speed maximum of one step parity (a fresh CFL reduction of the
// current state), leaving the times, dts, step count and error position as they are.
__global__ void k_ws_clear(DevState* st, int n, int slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st[i].ws[slot] = 0ull;
}

__global__ void k_state_init(DevState* st, int n, int slot, double t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        st[i].ws[0] = 0ull; st[i].ws[1] = 0ull;
        st[i].t[slot] = t; st[i].t[1 - slot] = t;
        st[i].dt[0] = 0.0; st[i].dt[1] = 0.0;
        st[i].err_pos = INT_MAX;
        st[i].nsteps = 0;
    }
}

// TVB minmod limiter (Cockburn & Shu 1998) on the P1 part, component-wise, followed by
// the positivity-preserving scaling of Zhang & Shu (2010), in place.  The limiter reads
// the neighbours' means only and never modifies a mean, so the in-place update is
// race-free.  Flags: bit 0 TVB-limited, bit 1 positivity-scaled.
constexpr double POS_EPS = 1e-13;

template <int P>
__global__ void __launch_bounds__(256) k_limit(LimArgs a) {
    using D = Deg<P>;
    constexpr int Np = D::Np, Nqv = D::Nqv, Nqe = D::Nqe;
    const DevTables& T = c_tab[P - 1];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    if (a.t_end > 0 && !(a.t_end - a.st->t[a.slot] > 0)) return;  // past the end time (as its stage)
    double c[4][Np], nm[3][4];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int i = 0; i < Np; ++i) c[v][i] = tmx(a.U, v * Np + i, 4 * Np, e);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int w = tmx(a.nbr, k, NBR_ROWS, e);
        if (w >= 0) {
            const int n = w >> 2;
#pragma unroll
            for (int v = 0; v < 4; ++v) nm[k][v] = tmx(a.U, v * Np, 4 * Np, n);
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) nm[k][v] = c[v][0];
        }
    }
    double al[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) al[r] = tmx(a.lim, r, LIM_ROWS, e);
    const double tol = a.M * tmx(a.lim, 6, LIM_ROWS, e);
    const int pk = a.pairs[e];
    const double nu = 1.5;
    int flag = 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double m0 = c[v][0];
        double d[3];
        bool changed = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double ut = T.Vmid[k][1] * c[v][1] + T.Vmid[k][2] * c[v][2];
            const int k1 = (pk >> (4 * k)) & 3, k2 = (pk >> (4 * k + 2)) & 3;
            double n1 = nm[0][v], n2 = nm[0][v];
            if (k1 == 1) n1 = nm[1][v]; else if (k1 == 2) n1 = nm[2][v];
            if (k2 == 1) n2 = nm[1][v]; else if (k2 == 2) n2 = nm[2][v];
            const double b = nu * (al[2 * k] * (n1 - m0) + al[2 * k + 1] * (n2 - m0));
            double r;
            if (fabs(ut) <= tol) r = ut;
            else if (ut > 0 && b > 0) r = fmin(ut, b);
            else if (ut < 0 && b < 0) r = fmax(ut, b);
            else r = 0.0;
            if (r != ut) changed = true;
            d[k] = r;
        }
        if (!changed) continue;
        flag = 1;
        const double sum = d[0] + d[1] + d[2];
        if (sum != 0.0) {
            double pos = 0, neg = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) { pos += fmax(0.0, d[k]); neg += fmax(0.0, -d[k]); }
            const double tp = pos > 0 ? fmin(1.0, neg / pos) : 0.0;
            const double tn = neg > 0 ? fmin(1.0, pos / neg) : 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) d[k] = tp * fmax(0.0, d[k]) - tn * fmax(0.0, -d[k]);
        }
        c[v][1] = T.recon[0][0] * d[0] + T.recon[0][1] * d[1] + T.recon[0][2] * d[2];
        c[v][2] = T.recon[1][0] * d[0] + T.recon[1][1] * d[1] + T.recon[1][2] * d[2];
#pragma unroll
        for (int i = 3; i < Np; ++i) c[v][i] = 0.0;
    }
    // positivity: density first, then pressure, at every volume and edge point
    const double g = a.gamma;
    const double rm = c[0][0];
    double rmin = rm;
#pragma unroll
    for (int q = 0; q < Nqv + 3 * Nqe; ++q) {
        double s = c[0][0];
#pragma unroll
        for (int i = 1; i < Np; ++i) s += (q < Nqv ? T.Vv[q][i] : T.Ve[(q - Nqv) / Nqe][(q - Nqv) % Nqe][i]) * c[0][i];
        rmin = fmin(rmin, s);
    }
    double th1 = 1.0;
    if (rmin < POS_EPS) th1 = fmin(1.0, (rm - POS_EPS) / (rm - rmin));
    if (th1 < 1.0) {
#pragma unroll
        for (int i = 1; i < Np; ++i) c[0][i] *= th1;
    }
    const double pm = (g - 1.0) * (c[3][0] - 0.5 * (c[1][0] * c[1][0] + c[2][0] * c[2][0]) / rm);
    double th2 = 1.0;
#pragma unroll
    for (int q = 0; q < Nqv + 3 * Nqe; ++q) {
        double U[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double s = c[v][0];
#pragma unroll
            for (int i = 1; i < Np; ++i) s += (q < Nqv ? T.Vv[q][i] : T.Ve[(q - Nqv) / Nqe][(q - Nqv) % Nqe][i]) * c[v][i];
            U[v] = s;
        }
        const double pq = (g - 1.0) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0]);
        if (pq < POS_EPS) th2 = fmin(th2, (pm - POS_EPS) / (pm - pq));
    }
    if (th1 < 1.0 || th2 < 1.0) {
        flag |= 2;
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int i = 1; i < Np; ++i) c[v][i] *= th2;
    }
    if (flag) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int i = 1; i < Np; ++i) tmx(a.U, v * Np + i, 4 * Np, e) = c[v][i];
    }
    a.flags[e] = flag;
}

// Plain streaming copy (16 bytes per lane per iteration), the calibration kernel for
// bandwidth and memory counters.
typedef double dvec2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(256) k_stream_copy(const dvec2* __restrict__ src, dvec2* __restrict__ dst,
                                                     long long n) {
    // four independent 16-byte loads in flight per lane, streamed past the caches
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        dvec2 v0 = __builtin_nontemporal_load(src + i), v1 = __builtin_nontemporal_load(src + i + stride);
        dvec2 v2 = __builtin_nontemporal_load(src + i + 2 * stride), v3 = __builtin_nontemporal_load(src + i + 3 * stride);
        __builtin_nontemporal_store(v0, dst + i);
        __builtin_nontemporal_store(v1, dst + i + stride);
        __builtin_nontemporal_store(v2, dst + i + 2 * stride);
        __builtin_nontemporal_store(v3, dst + i + 3 * stride);
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

int launch_stream_copy(const void* src, void* dst, long long bytes, void* stream) {
    const long long n = bytes / 16;
    hipLaunchKernelGGL(k_stream_copy, dim3(8192), dim3(256), 0, (hipStream_t)stream, (const dvec2*)src,
                       (dvec2*)dst, n);
    return (int)hipGetLastError();
}

int upload_tables(int P, const DevTables& t) {
    hipError_t rc = hipMemcpyToSymbol(HIP_SYMBOL(c_tab), &t, sizeof(DevTables), (size_t)(P - 1) * sizeof(DevTables),
                                      hipMemcpyHostToDevice);
    return (int)rc;
}

static int grid_blocks(const StageArgs& a) {
    int nb = a.xcd_remap ? (a.ntiles + 7) / 8 * 8 : a.ntiles;
    if (a.persistent > 0 && a.persistent < nb) nb = a.xcd_remap ? (a.persistent + 7) / 8 * 8 : a.persistent;
    return nb;
}

template <int P, int FLUX>
static void launch_stage_pf(int block, const StageArgs& a, hipStream_t s) {
    dim3 grid(grid_blocks(a)), blk(block);
    if (a.staged) {
        constexpr size_t lds = (size_t)4 * Deg<P>::Np * staged_stride(P) * sizeof(double);
        static bool opted = false;  // above 64 KiB the kernels must opt in
        if (lds > 65536 && !opted) {
            // (a failure leaves the attribute unset: the launch then fails and reports it)
            opted = hipFuncSetAttribute((const void*)k_stage<P, FLUX, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess &&
                    hipFuncSetAttribute((const void*)k_stage_pipe<P <= 2 ? P : 2, FLUX>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess;
        }
        if (P == 1 && a.pipe == 2 && !a.tile_mesh)  // + the LDS copy of 4 geometry rows
            hipLaunchKernelGGL((k_stage_rp<1, FLUX>), grid, blk, lds + 4 * staged_block(1) * sizeof(double), s, a);
        else if (P <= 2 && a.pipe) hipLaunchKernelGGL((k_stage_pipe<P <= 2 ? P : 2, FLUX>), grid, blk, lds, s, a);
        else hipLaunchKernelGGL((k_stage<P, FLUX, true>), grid, blk, lds, s, a);
        return;
    }
    hipLaunchKernelGGL((k_stage<P, FLUX, false>), grid, blk, 0, s, a);
}

template <int P>
static void launch_stage_p(int flux, int block, const StageArgs& a, hipStream_t s) {
    if (flux == RKDG_FLUX_ROE) launch_stage_pf<P, RKDG_FLUX_ROE>(block, a, s);
    else launch_stage_pf<P, RKDG_FLUX_LF>(block, a, s);
}

template <int P>
static void launch_stage_mx_p(int flux, const StageArgs& a, hipStream_t s) {
    dim3 grid(grid_blocks(a)), blk(256);
    if (flux == RKDG_FLUX_ROE) hipLaunchKernelGGL((k_stage_mx<P, RKDG_FLUX_ROE>), grid, blk, 0, s, a);
    else hipLaunchKernelGGL((k_stage_mx<P, RKDG_FLUX_LF>), grid, blk, 0, s, a);
}

template <int P, int FLUX, int CT>
static void launch_stage_px_pf(const StageArgs& a, hipStream_t s) {
    constexpr size_t lds = (size_t)4 * Deg<P>::Np * staged_stride(P) * sizeof(double);
    static bool opted = false;  // above 64 KiB the kernel must opt in
    if (lds > 65536 && !opted) {
        opted = hipFuncSetAttribute((const void*)k_stage_px<P, FLUX, CT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess;
    }
    hipLaunchKernelGGL((k_stage_px<P, FLUX, CT>), dim3(grid_blocks(a)), dim3(staged_block(P)), lds, s, a);
}

template <int P, int CT>
static void launch_stage_px_p(int flux, const StageArgs& a, hipStream_t s) {
    if (flux == RKDG_FLUX_ROE) launch_stage_px_pf<P, RKDG_FLUX_ROE, CT>(a, s);
    else launch_stage_px_pf<P, RKDG_FLUX_LF, CT>(a, s);
}

template <int FLUX, int CT>
static void launch_stage_pz_f(const StageArgs& a, hipStream_t s) {
    static bool opted = false;  // above 64 KiB the kernel must opt in
    if (!opted) {
        opted = hipFuncSetAttribute((const void*)k_stage_pz<3, FLUX, CT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)Pz<3>::lds(CT)) == hipSuccess;
    }
    hipLaunchKernelGGL((k_stage_pz<3, FLUX, CT>), dim3(grid_blocks(a)), dim3(Pz<3>::B), Pz<3>::lds(CT), s, a);
}

int launch_stage(int P, int flux, int prec, int block, const StageArgs& a, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (prec == RKDG_PREC_BF16_MFMA) {
        // BF16 contractions: staged tiles of 256 elements, one lane per element (P1, P2)
        if (!a.staged || block != px_block(P)) return (int)hipErrorInvalidValue;
        switch (P) {
            case 1: launch_stage_px_p<1, CT_BF16>(flux, a, s); break;
            case 2: launch_stage_px_p<2, CT_BF16>(flux, a, s); break;
            case 3:
                if (flux == RKDG_FLUX_ROE) launch_stage_pz_f<RKDG_FLUX_ROE, CT_BF16>(a, s);
                else launch_stage_pz_f<RKDG_FLUX_LF, CT_BF16>(a, s);
                break;
            default: return (int)hipErrorInvalidValue;
        }
        return (int)hipGetLastError();
    }
    if (prec == RKDG_PREC_FP32_MFMA) {
        // staged tiles of 256 elements, one lane per element (P1, P2)
        if (a.staged) {
            if (block != px_block(P)) return (int)hipErrorInvalidValue;
            switch (P) {
                case 1: launch_stage_px_p<1, CT_F32>(flux, a, s); break;
                case 2: launch_stage_px_p<2, CT_F32>(flux, a, s); break;
                case 3:
                    if (flux == RKDG_FLUX_ROE) launch_stage_pz_f<RKDG_FLUX_ROE, CT_F32>(a, s);
                    else launch_stage_pz_f<RKDG_FLUX_LF, CT_F32>(a, s);
                    break;
                default: return (int)hipErrorInvalidValue;
            }
            return (int)hipGetLastError();
        }
        // four lanes per element: tiles of 64 elements on 256 threads
        if (block != 64) return (int)hipErrorInvalidValue;
        switch (P) {
            case 1: launch_stage_mx_p<1>(flux, a, s); break;
            case 2: launch_stage_mx_p<2>(flux, a, s); break;
            case 3: launch_stage_mx_p<3>(flux, a, s); break;
            default: return (int)hipErrorInvalidValue;
        }
        return (int)hipGetLastError();
    }
    switch (P) {
        case 1: launch_stage_p<1>(flux, block, a, s); break;
        case 2: launch_stage_p<2>(flux, block, a, s); break;
        case 3: launch_stage_p<3>(flux, block, a, s); break;
        default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

int launch_wavespeed(int P, const StageArgs& a, int block, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    dim3 grid(grid_blocks(a)), blk(block);
    switch (P) {
        case 1: hipLaunchKernelGGL(k_wavespeed<1>, grid, blk, 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_wavespeed<2>, grid, blk, 0, s, a); break;
        case 3: hipLaunchKernelGGL(k_wavespeed<3>, grid, blk, 0, s, a); break;
        default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

// ---- allocation-free step: its scratch lives in one static device array -------------
// Layout: a DevState (the step's counters and wave-speed maxima) in the first 256 bytes,
// then one pair of residual maxima per workgroup tile.  Each GPU holds its own copy of the
// array until the process ends; host code never names it in a kernel argument or copy
// (on the host the name is a stand-in, not a device address): it uses the address
// hipGetSymbolAddress returned for the device.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};
}  // namespace

int noalloc_scratch(int dev, void** out) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        // with dev current: this device's copy of the array
        void* ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (rc != hipSuccess) return (int)rc;
        g_scratch_dev[dev] = ptr;
    }
    *out = g_scratch_dev[dev];
    return (int)hipSuccess;
}

// Maxima of the per-tile residual pairs into out[0..1] (one workgroup).
__global__ void __launch_bounds__(256) k_res_reduce(const double* __restrict__ part, int ntiles,
                                                    double* __restrict__ out) {
    __shared__ double red[2][256];
    double a = 0.0, b = 0.0;
    for (int t = threadIdx.x; t < ntiles; t += blockDim.x) {
        a = fmax(a, part[2 * t]);
        b = fmax(b, part[2 * t + 1]);
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + o]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0][0];
        out[1] = red[1][0];
    }
}

int launch_res_reduce(const double* part, int ntiles, double* out, void* stream) {
    hipLaunchKernelGGL(k_res_reduce, dim3(1), dim3(256), 0, (hipStream_t)stream, part, ntiles, out);
    return (int)hipGetLastError();
}

int launch_ws_clear(DevState* st, int slot, int n, void* stream) {
    hipLaunchKernelGGL(k_ws_clear, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, st, n, slot);
    return (int)hipGetLastError();
}

int launch_state_init(DevState* st, int slot, double t, int n, void* stream) {
    hipLaunchKernelGGL(k_state_init, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, st, n, slot, t);
    return (int)hipGetLastError();
}

int launch_limiter(int P, const LimArgs& a, int block, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    dim3 grid((a.E + block - 1) / block), blk(block);
    switch (P) {
        case 1: hipLaunchKernelGGL(k_limit<1>, grid, blk, 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_limit<2>, grid, blk, 0, s, a); break;
        case 3: hipLaunchKernelGGL(k_limit<3>, grid, blk, 0, s, a); break;
        default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

const char* hip_error_string(int code) { return hipGetErrorString((hipError_t)code); }

}  // namespace rkdg
