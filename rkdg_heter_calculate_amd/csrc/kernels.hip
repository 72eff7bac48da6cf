// gfx950 kernels of the plain FP64 form.
//
// One fused kernel per SSP-RK3 stage replaces the reference's nine launches per stage
// (calculateConVars, boundaryCondition, calculateVolumeRHS, calculateLFCoeff,
// calculateEdgeFG, calculateFlux x2, calculateEdgeRHS, rkdgStep*;
// src/kernelfunctions.cu:84-748).  One lane owns one element:
//   * its modal state is read once, coalesced, from the SoA layout U[var][mode][ld];
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
#include <cstdlib>
#include "internal.h"

namespace rkdg {

__constant__ DevTables c_tab[3];

template <int P>
struct Deg {
    static constexpr int Np = (P + 1) * (P + 2) / 2;
    static constexpr int Nqv = P <= 2 ? 7 : 12;
    static constexpr int Nqe = P <= 2 ? 3 : 4;
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block-wide max of three non-negative values, then one atomic per value.
__device__ __forceinline__ void block_max3_atomic(double a, double b, double c, unsigned long long* pa,
                                                  unsigned long long* pb, unsigned long long* pc, bool do_bc) {
    __shared__ double red[3][16];
    a = wave_max(a);
    if (do_bc) { b = wave_max(b); c = wave_max(c); }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) { red[0][wid] = a; red[1][wid] = b; red[2][wid] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ma = 0, mb = 0, mc = 0;
        for (int w = 0; w < nw; ++w) { ma = fmax(ma, red[0][w]); mb = fmax(mb, red[1][w]); mc = fmax(mc, red[2][w]); }
        if (ma > 0) atomicMax(pa, (unsigned long long)__double_as_longlong(ma));
        if (do_bc) {
            if (mb > 0) atomicMax(pb, (unsigned long long)__double_as_longlong(mb));
            if (mc > 0) atomicMax(pc, (unsigned long long)__double_as_longlong(mc));
        }
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
    const double sl = sqrt(rl), sr = sqrt(rr), isum = frcp(sl + sr);
    const double rt = sl * sr;
    const double ut = (sl * ul + sr * ur) * isum, vt = (sl * vl + sr * vr) * isum, ht = (sl * hl + sr * hr) * isum;
    const double q2 = ut * ut + vt * vt;
    const double c2 = (g - 1) * (ht - 0.5 * q2);
    const double ct = sqrt(c2), ic2 = frcp(c2);
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
    const double cl = sqrt(g * pl * irl), cr = sqrt(g * pr * irr);
    double l1 = fabs(unt - ct), l4 = fabs(unt + ct);
    const double l2 = fabs(unt);
    const double d1 = fmax(0.0, fmax((unt - ct) - (unl - cl), (unr - cr) - (unt - ct)));
    const double d4 = fmax(0.0, fmax((unt + ct) - (unl + cl), (unr + cr) - (unt + ct)));
    if (l1 < d1) l1 = (l1 * l1 + d1 * d1) / (2 * d1);
    if (l4 < d4) l4 = (l4 * l4 + d4 * d4) / (2 * d4);
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

// Neighbour prefetch: the neighbour's modes are loaded one edge ahead (during the volume
// integral for the first edge), so their latency hides behind the flux arithmetic.  It
// costs 8*4*Np bytes of registers per lane, affordable up to P2.
template <int P>
struct Prefetch {
    static constexpr bool on = P <= 2;
};

template <int Np>
__device__ __forceinline__ void load_modes(const double* __restrict__ U, long long ld, int n, double (&c)[4][Np]) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int i = 0; i < Np; ++i) c[v][i] = U[(v * Np + i) * ld + n];
}

// STAGED: the tile's own modes and its halo (the out-of-tile neighbours, listed per tile
// on the host) are first copied into LDS with independent coalesced loads; every
// neighbour trace is then evaluated from LDS, so no load in the element's arithmetic
// depends on another load.  Otherwise neighbours are gathered from global memory
// (prefetched one edge ahead up to P2).
template <int P, int FLUX, bool STAGED>
__device__ __forceinline__ void stage_tile(const StageArgs& a, const TileInfo& ti, const double* sVe, double* sU) {
    using D = Deg<P>;
    constexpr int Np = D::Np, Nqv = D::Nqv, Nqe = D::Nqe;
    constexpr bool PF = !STAGED && Prefetch<P>::on;
    const DevTables& T = c_tab[P - 1];
    bool done;
    const double dt = stage_dt(a, ti, done);
    if (ti.first && threadIdx.x == 0) {
        if (a.stage == 0 && !done) {
            ti.st->ws[1 - a.slot] = 0ull;
            ti.st->res[0] = 0ull;
            ti.st->res[1] = 0ull;
            ti.st->dt[a.slot] = dt;
        }
        if (a.stage == 2) {
            ti.st->t[1 - a.slot] = done ? ti.st->t[a.slot] : ti.st->t[a.slot] + dt;
            if (done) ti.st->ws[1 - a.slot] = ti.st->ws[a.slot];
            else ti.st->nsteps += 1;
        }
    }
    if (done) return;

    const long long ld = a.ld;
    const int e = ti.tile * blockDim.x + threadIdx.x;
    const double g = a.gamma;
    double ws_l = 0.0, r0_l = 0.0, r1_l = 0.0;

    const int* nb_ids = STAGED ? a.nloc : a.nbr;
    double Uo[4][Np];
    if (e < a.E) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int i = 0; i < Np; ++i) Uo[v][i] = a.U[(v * Np + i) * ld + e];
    }
    if (STAGED) {
        const int S = a.lds_stride;
        if (e < a.E) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int i = 0; i < Np; ++i) sU[(v * Np + i) * S + threadIdx.x] = Uo[v][i];
        }
        const int h0 = a.hoff[ti.tile], nh = a.hoff[ti.tile + 1] - h0;
        for (int h = threadIdx.x; h < nh; h += blockDim.x) {
            const int n = a.hids[h0 + h];
#pragma unroll
            for (int r = 0; r < 4 * Np; ++r) sU[r * S + blockDim.x + h] = a.U[r * ld + n];
        }
        __syncthreads();
    }

    if (e < a.E) {
        const int w0 = nb_ids[e], w1 = nb_ids[ld + e], w2 = nb_ids[2 * ld + e];
        double R[4][Np], Nb[4][Np];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int i = 0; i < Np; ++i) R[v][i] = 0.0;
        const double area = a.geo[4 * ld + e];
        const double xix = a.geo[0 * ld + e], xiy = a.geo[1 * ld + e];
        const double etx = a.geo[2 * ld + e], ety = a.geo[3 * ld + e];
        double lk_n = a.geo[5 * ld + e], nx_n = a.geo[8 * ld + e], ny_n = a.geo[9 * ld + e];
        if (PF && w0 >= 0) load_modes<Np>(a.U, ld, w0 >> 2, Nb);

        // ---- volume integral: sum_q w_q A (F . grad psi_i) -------------------------
        // The point loop stays rolled: each iteration reads one row of the reference
        // operators into scalar registers instead of keeping every table live.
#pragma unroll 1
        for (int q = 0; q < Nqv; ++q) {
            double uq[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double s = Uo[v][0] * T.Vv[q][0];
#pragma unroll
                for (int i = 1; i < Np; ++i) s = fma(Uo[v][i], T.Vv[q][i], s);
                uq[v] = s;
            }
            // contravariant fluxes w A F.(xi_x, xi_y) and w A F.(eta_x, eta_y)
            const double ir = frcp(uq[0]);
            const double pr = pressure(g, uq, ir);
            const double wa = T.wv[q] * area;
            double Fx[4], Fe[4];
            euler_dir(g, uq, ir, pr, wa * xix, wa * xiy, Fx);
            euler_dir(g, uq, ir, pr, wa * etx, wa * ety, Fe);
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int i = 1; i < Np; ++i) R[v][i] = fma(Fx[v], T.Dxi[q][i], fma(Fe[v], T.Deta[q][i], R[v][i]));
        }

        // ---- surface integral, element-owned -------------------------------------
#pragma unroll 1
        for (int k = 0; k < 3; ++k) {
            const int w = k == 0 ? w0 : (k == 1 ? w1 : w2);
            const double lk = lk_n, nx = nx_n, ny = ny_n;
            double uR[Nqe][4], mR[4];
            if (w >= 0) {
                // neighbour trace at the mirrored points
                const double* tb = sVe + (w & 3) * (Nqe * Np);
                if (STAGED) {
                    const double* src = sU + (w >> 2);
                    const int S = a.lds_stride;
#pragma unroll
                    for (int i = 0; i < Np; ++i) {
                        double c[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) c[v] = src[(v * Np + i) * S];
                        if (i == 0) {
#pragma unroll
                            for (int v = 0; v < 4; ++v) mR[v] = c[v];
                        }
#pragma unroll
                        for (int j = 0; j < Nqe; ++j) {
                            const double t = tb[(Nqe - 1 - j) * Np + i];
#pragma unroll
                            for (int v = 0; v < 4; ++v) uR[j][v] = i == 0 ? c[v] * t : fma(c[v], t, uR[j][v]);
                        }
                    }
                } else {
                    if (!PF) load_modes<Np>(a.U, ld, w >> 2, Nb);
#pragma unroll
                    for (int v = 0; v < 4; ++v) mR[v] = Nb[v][0];
#pragma unroll
                    for (int i = 0; i < Np; ++i)
#pragma unroll
                        for (int j = 0; j < Nqe; ++j) {
                            const double t = tb[(Nqe - 1 - j) * Np + i];
#pragma unroll
                            for (int v = 0; v < 4; ++v)
                                uR[j][v] = i == 0 ? Nb[v][0] * t : fma(Nb[v][i], t, uR[j][v]);
                        }
                }
            } else if (w == -RKDG_BC_FARFIELD) {
                // kernel_boundaryCondition CELL_FARFIELD (src/kernelfunctions.cu:184-201)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    mR[v] = a.fs[v];
#pragma unroll
                    for (int j = 0; j < Nqe; ++j) uR[j][v] = a.fs[v];
                }
            } else {
                // CELL_REFLECTION (232-268) mirrors the velocity about the edge normal;
                // CELL_OUTFLOW (203-230) copies the interior trace
                const bool refl = w == -RKDG_BC_REFLECTION;
                const double cxx = refl ? ny * ny - nx * nx : 1.0, cyy = refl ? nx * nx - ny * ny : 1.0;
                const double cxy = refl ? 2.0 * nx * ny : 0.0;
#pragma unroll
                for (int j = 0; j < Nqe; ++j) {
                    double uL[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        double s = Uo[v][0] * T.Ve[k][j][0];
#pragma unroll
                        for (int i = 1; i < Np; ++i) s = fma(Uo[v][i], T.Ve[k][j][i], s);
                        uL[v] = s;
                    }
                    uR[j][0] = uL[0];
                    uR[j][3] = uL[3];
                    uR[j][1] = uL[1] * cxx - cxy * uL[2];
                    uR[j][2] = uL[2] * cyy - cxy * uL[1];
                }
                mR[0] = Uo[0][0];
                mR[3] = Uo[3][0];
                mR[1] = Uo[1][0] * cxx - cxy * Uo[2][0];
                mR[2] = Uo[2][0] * cyy - cxy * Uo[1][0];
            }
            // next edge's geometry and neighbour modes, in flight during this edge's fluxes
            if (k < 2) {
                const int wn = k == 0 ? w1 : w2;
                lk_n = a.geo[(6 + k) * ld + e];
                nx_n = a.geo[(10 + 2 * k) * ld + e];
                ny_n = a.geo[(11 + 2 * k) * ld + e];
                if (PF && wn >= 0) load_modes<Np>(a.U, ld, wn >> 2, Nb);
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
#pragma unroll
            for (int j = 0; j < Nqe; ++j) {
                double uL[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double s = Uo[v][0] * T.Ve[k][j][0];
#pragma unroll
                    for (int i = 1; i < Np; ++i) s = fma(Uo[v][i], T.Ve[k][j][i], s);
                    uL[v] = s;
                }
                double Fh[4];
                if (FLUX == RKDG_FLUX_LF) {
                    // 0.5 (F(uL).n + F(uR).n - alpha (uR - uL)), kernel_calculateFlux's formula
                    double FL[4], FR[4];
                    const double irl = frcp(uL[0]), irr = frcp(uR[j][0]);
                    euler_dir(g, uL, irl, pressure(g, uL, irl), nx, ny, FL);
                    euler_dir(g, uR[j], irr, pressure(g, uR[j], irr), nx, ny, FR);
#pragma unroll
                    for (int v = 0; v < 4; ++v) Fh[v] = 0.5 * (FL[v] + FR[v] - alpha * (uR[j][v] - uL[v]));
                } else {
                    roe_flux(g, uL, uR[j], nx, ny, Fh);
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

        // ---- SSP-RK3 stage (src/kernelfunctions.cu:640-748) -----------------------
        const double ia = frcp(area);
        double mean[4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int i = 0; i < Np; ++i) {
                const long long idx = (v * Np + i) * ld + e;
                const double r = R[v][i] * (T.minv[i] * ia);
                double out;
                if (a.stage == 0) {
                    out = Uo[v][i] + dt * r;
                } else if (a.stage == 1) {
                    out = 0.75 * a.U0[idx] + 0.25 * Uo[v][i] + 0.25 * dt * r;
                } else {
                    const double o0 = a.U0[idx];
                    out = (1.0 / 3.0) * o0 + (2.0 / 3.0) * Uo[v][i] + (2.0 / 3.0) * dt * r;
                    if (i == 0 && a.dtflag[e]) {
                        if (v == 0) r0_l = fabs(out - o0);
                        if (v == 3) r1_l = fabs(out - o0);
                    }
                }
                if (i == 0) mean[v] = out;
                a.Uout[idx] = out;
            }

        if (a.stage == 2) {
            // next step's CFL wave speed on the new means (kernel_getTimeStep's formula,
            // src/kernelfunctions.cu:19-25)
            const double rho = mean[0], u = mean[1] / rho, v = mean[2] / rho;
            const double p = (g - 1.0) * (mean[3] - 0.5 * rho * (u * u + v * v));
            if (!(rho > 0.0) || !(p > 0.0)) atomicMin(&ti.st->err_pos, e);
            const double c = sqrt(g * p / rho);
            const double per = a.geo[5 * ld + e] + a.geo[6 * ld + e] + a.geo[7 * ld + e];
            const double s = (sqrt(u * u + v * v) + c) * per / area;
            if (a.dtflag[e] && s > 0.0) ws_l = s;
        }
    }
    if (a.stage == 2)
        block_max3_atomic(ws_l, r0_l, r1_l, &ti.st->ws[1 - a.slot], &ti.st->res[0], &ti.st->res[1], true);
}

// WPE > 0 asks the compiler for at least WPE waves per SIMD (fewer registers, possibly
// spills); chosen at run time by measurement.  The staged form takes its LDS image of the
// tile and halo as dynamic shared memory: 4 Np rows of lds_stride doubles.
template <int P, int FLUX, int WPE, bool STAGED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1, 8))) k_stage(StageArgs a) {
    using D = Deg<P>;
    constexpr int Np = D::Np, Nqe = D::Nqe;
    const DevTables& T = c_tab[P - 1];
    __shared__ double sVe[3 * Nqe * Np];
    extern __shared__ double sU[];
    for (int t = threadIdx.x; t < 3 * Nqe * Np; t += blockDim.x) {
        const int k = t / (Nqe * Np), r = t - k * (Nqe * Np), j = r / Np, i = r - j * Np;
        sVe[t] = T.Ve[k][j][i];
    }
    __syncthreads();
    const int nit = iter_count(a);
    for (int it = blockIdx.x; it < nit; it += gridDim.x) {
        TileInfo ti;
        if (!tile_of_iter(a, it, nit, ti)) continue;  // uniform across the workgroup
        stage_tile<P, FLUX, STAGED>(a, ti, sVe, sU);
        __syncthreads();  // LDS and the reduction scratch are reused by the next tile
    }
}

#include "stage_mx.inc"

// CFL wave speed of the current state, max-reduced into ws[slot]
template <int P>
__global__ void __launch_bounds__(256) k_wavespeed(StageArgs a) {
    constexpr int Np = Deg<P>::Np;
    const long long ld = a.ld;
    const int nit = iter_count(a);
    for (int it = blockIdx.x; it < nit; it += gridDim.x) {
        TileInfo ti;
        if (!tile_of_iter(a, it, nit, ti)) continue;
        const int e = ti.tile * blockDim.x + threadIdx.x;
        double s_l = 0.0;
        if (e < a.E) {
            const double g = a.gamma;
            const double rho = a.U[(0 * Np) * ld + e];
            const double u = a.U[(1 * Np) * ld + e] / rho, v = a.U[(2 * Np) * ld + e] / rho;
            const double p = (g - 1.0) * (a.U[(3 * Np) * ld + e] - 0.5 * rho * (u * u + v * v));
            if (!(rho > 0.0) || !(p > 0.0)) atomicMin(&ti.st->err_pos, e);
            const double c = sqrt(g * p / rho);
            const double area = a.geo[4 * ld + e];
            const double s = (sqrt(u * u + v * v) + c) * (a.geo[5 * ld + e] + a.geo[6 * ld + e] + a.geo[7 * ld + e]) / area;
            if (a.dtflag[e] && s > 0.0) s_l = s;
        }
        block_max3_atomic(s_l, 0.0, 0.0, &ti.st->ws[a.slot], nullptr, nullptr, false);
        __syncthreads();
    }
}

__global__ void k_state_init(DevState* st, int n, int slot, double t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        st[i].ws[0] = 0ull; st[i].ws[1] = 0ull;
        st[i].t[slot] = t; st[i].t[1 - slot] = t;
        st[i].dt[0] = 0.0; st[i].dt[1] = 0.0;
        st[i].res[0] = 0ull; st[i].res[1] = 0ull;
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
    const long long ld = a.ld;
    double c[4][Np], nm[3][4];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int i = 0; i < Np; ++i) c[v][i] = a.U[(v * Np + i) * ld + e];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int w = a.nbr[k * ld + e];
        if (w >= 0) {
            const int n = w >> 2;
#pragma unroll
            for (int v = 0; v < 4; ++v) nm[k][v] = a.U[(v * Np) * ld + n];
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) nm[k][v] = c[v][0];
        }
    }
    double al[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) al[r] = a.lim[r * ld + e];
    const double tol = a.M * a.lim[6 * ld + e];
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
            for (int i = 1; i < Np; ++i) a.U[(v * Np + i) * ld + e] = c[v][i];
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

static int stage_wpe() {
    static int w = [] {
        const char* e = std::getenv("RKDG_WPE");
        return e ? std::atoi(e) : 0;
    }();
    return w;
}

template <int P, int FLUX>
static void launch_stage_pf(int block, const StageArgs& a, hipStream_t s) {
    dim3 grid(grid_blocks(a)), blk(block);
    if (a.staged) {
        const size_t lds = (size_t)4 * Deg<P>::Np * a.lds_stride * sizeof(double);
        static size_t lds_set = 65536;  // above 64 KiB the kernel must opt in
        if (lds > lds_set) {
            hipFuncSetAttribute((const void*)k_stage<P, FLUX, 0, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            lds_set = lds;
        }
        hipLaunchKernelGGL((k_stage<P, FLUX, 0, true>), grid, blk, lds, s, a);
        return;
    }
    switch (stage_wpe()) {
        case 2: hipLaunchKernelGGL((k_stage<P, FLUX, 2, false>), grid, blk, 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_stage<P, FLUX, 3, false>), grid, blk, 0, s, a); break;
        default: hipLaunchKernelGGL((k_stage<P, FLUX, 0, false>), grid, blk, 0, s, a); break;
    }
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

int launch_stage(int P, int flux, int prec, int block, const StageArgs& a, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (prec == RKDG_PREC_FP32_MFMA) {
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
