/*
 * librkdg — MI355X (gfx950) RKDG solver for the 2-D compressible Euler equations on
 * unstructured triangles.  C ABI: plain pointers and sizes, int status codes, no
 * exception crosses the boundary.  The caller owns host buffers, the library owns
 * device buffers.  One HIP stream per context; a context is not thread-safe.
 *
 * Reference paths below are relative to the reference checkout's multi_gpuRkdg/.
 *
 * Conventions shared with the reference:
 *   - triangles are CCW, edge k lies opposite vertex k, neighbour k lies across edge k
 *     (src/triangleinfo.cpp:171-189, src/unstructuredgrid.cpp:193-288);
 *   - state arrays are [4 conserved variables][Np modes][n_elements], variables
 *     (rho, rho*u, rho*v, rho*E) — the reference's freedom_* rows
 *     (src/cudaarrays.cpp:72-152), element order is the caller's (original) order;
 *   - boundary tags are the reference's cell_tags (inc/defines.h:49-55).
 */
#ifndef RKDG_H
#define RKDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define RKDG_OK               0
#define RKDG_ERR_ARG         -1  /* invalid argument or size */
#define RKDG_ERR_HIP         -2  /* HIP runtime error (message in rkdg_last_error) */
#define RKDG_ERR_NONPHYSICAL -3  /* non-positive density or pressure (cell id reported) */
#define RKDG_ERR_MESH        -4  /* inconsistent connectivity */
#define RKDG_ERR_IO          -5  /* file missing or malformed */
#define RKDG_ERR_STATE       -6  /* call out of order (e.g. no state set) */
#define RKDG_ERR_NOMEM       -7

/* ---- enumerations ------------------------------------------------------------------ */
/* boundary tags: inc/defines.h:49-55 (CELL_FARFIELD, CELL_OUTFLOW, CELL_REFLECTION) */
#define RKDG_BC_FARFIELD    1
#define RKDG_BC_OUTFLOW     2   /* transmissive */
#define RKDG_BC_REFLECTION  3   /* slip wall */

#define RKDG_FLUX_LF   0   /* Lax-Friedrichs with alpha from averaged means, src/kernelfunctions.cu:354-390,510-582 */
#define RKDG_FLUX_ROE  1   /* Roe (1981), Harten-Hyman entropy fix on the acoustic waves */

#define RKDG_LIMITER_NONE 0
#define RKDG_LIMITER_TVB  1   /* TVB minmod, Cockburn-Shu, on the P1 part, component-wise */

#define RKDG_DT_REFERENCE 0   /* CFL over the reduction subset of src/kernelfunctions.cu:3-81 */
#define RKDG_DT_MAX       1   /* CFL over every element */
#define RKDG_DT_FIXED     2

#define RKDG_PREC_FP64      0   /* plain form: every operation in FP64 */
#define RKDG_PREC_FP32_MFMA 1   /* basis contractions on MFMA in FP32, FP64 state and update */
#define RKDG_PREC_BF16_MFMA 2   /* basis contractions on MFMA in BF16 with FP32 accumulation */

#define RKDG_BASIS_ELEMENT 0  /* the reference's per-element Gram-Schmidt basis (src/triangleinfo.cpp:433-602) */
#define RKDG_BASIS_NATIVE  1  /* this library's reference-element orthonormal basis */

#define RKDG_IC_FREESTREAM 0  /* src/cudarkdgsolver.cu:326-368; p: gamma, alpha_rad, mach, rhoref, pref */
#define RKDG_IC_VORTEX     1  /* isentropic vortex; p: gamma, beta, uinf, vinf, xc, yc, Lx, Ly, t (Lx=Ly=0: no images) */
#define RKDG_IC_RIEMANN4   2  /* four quadrants (cell averages); p: gamma, xs, ys, then (rho,u,v,p) for UR, UL, LL, LR */

/* ---- mesh and parameters ----------------------------------------------------------- */
typedef struct rkdg_mesh {
    int32_t n_vertices;
    int32_t n_elements;
    const double*  vx;          /* [n_vertices] */
    const double*  vy;          /* [n_vertices] */
    const int32_t* tri;         /* [n_elements][3] vertex ids, CCW */
    const int32_t* neighbour;   /* [n_elements][3] element across edge k, -1 = boundary */
    const int32_t* shared_edge; /* [n_elements][3] neighbour's local index of the shared edge; NULL: derived */
    const int32_t* boundary;    /* [n_elements][3] RKDG_BC_* of boundary edges; NULL: all RKDG_BC_OUTFLOW */
    const int32_t* order;       /* optional [n_elements] device order (locality permutation); NULL: identity */
    const int32_t* dt_id;       /* optional [n_elements] ids for the RKDG_DT_REFERENCE subset; NULL: element index */
} rkdg_mesh;

typedef struct rkdg_params {
    int32_t degree;            /* 1, 2 or 3 */
    int32_t flux;              /* RKDG_FLUX_* */
    int32_t limiter;           /* RKDG_LIMITER_* */
    int32_t dt_policy;         /* RKDG_DT_* */
    int32_t reduction_threads; /* 'reductionthreads' key of main.conf (RKDG_DT_REFERENCE); default 512 */
    int32_t precision;         /* RKDG_PREC_* */
    int32_t device;            /* HIP device ordinal */
    int32_t block;             /* elements per workgroup (0: library default / autotuned) */
    double  gamma;
    double  cfl;
    double  tvb_M;
    double  dt_fixed;
    double  freestream[4];     /* conserved far-field state (src/rkdgadvance.cpp:31-34) */
} rkdg_params;

typedef struct rkdg_ctx rkdg_ctx;
typedef struct rkdg_batch rkdg_batch;

/* Fills p with the reference's defaults (src/cudarkdgsolver.cu:3-32): gamma 1.4,
 * cfl 0.18, P2, LF, no limiter, reference dt subset with 512 reduction threads. */
void rkdg_default_params(rkdg_params* p);

/* ---- single mesh ------------------------------------------------------------------- */
/* Replaces CCUDARkdgSolver::runNext's set-up: initializeTriangleInfos,
 * markLocalBoundaryTriangles, CCUDAArrays::allocateMemory, copyTriangleInfosToGPU
 * (src/cudarkdgsolver.cu:192-230,266-324). */
int rkdg_create(const rkdg_mesh* mesh, const rkdg_params* params, rkdg_ctx** out);
int rkdg_destroy(rkdg_ctx* ctx);

/* Host <-> device state, [4][Np][n_elements] in the caller's element order.
 * Replaces the freedom_* H2D copies of initRKDG and copyFreedomToHost
 * (src/cudarkdgsolver.cu:253-264,357-367). */
int rkdg_set_state(rkdg_ctx* ctx, int32_t basis, const double* U);
int rkdg_get_state(rkdg_ctx* ctx, int32_t basis, double* U);

/* Initial condition: RKDG_IC_FREESTREAM restates CCUDARkdgSolver::initRKDG
 * (src/cudarkdgsolver.cu:326-368); the others are L2 projections. */
int rkdg_project_ic(rkdg_ctx* ctx, int32_t kind, const double* p, int32_t np);

/* CFL time step of the current state; replaces getTimeStep + the D2H of ddt
 * (src/cudarkdgsolver.cu:370-380, src/rkdgadvance.cpp:165-168). */
int rkdg_compute_dt(rkdg_ctx* ctx, double* dt);

/* One SSP-RK3 stage s = 0, 1, 2 with a caller dt; stage 0 keeps the step-start state
 * (the U -> U_old copy of src/rkdgadvance.cpp:104-107).  Replaces the per-stage chain
 * calculateConVars ... calculateEdgeRHS + rkdgStepOne/Two/Three
 * (src/rkdgadvance.cpp:132-195, src/cudarkdgsolver.cu:382-599). */
int rkdg_stage(rkdg_ctx* ctx, int32_t s, double dt);

/* One full step with the dt computed and kept on the device; dt_used may be NULL
 * (reading it synchronises).  Replaces one iteration of RkdgAdvance::advance
 * (src/rkdgadvance.cpp:91-243). */
int rkdg_step(rkdg_ctx* ctx, double* dt_used);

/* nsteps steps, or fewer when t_end > 0 is reached (the last step is clipped).
 * No host synchronisation inside; steps_done and t may be NULL.  Replaces the
 * RkdgAdvance::advance loop (src/rkdgadvance.cpp:91-243, which stops after 11 steps). */
int rkdg_advance(rkdg_ctx* ctx, int32_t nsteps, double t_end, int32_t* steps_done, double* t);

/* max |change| of the rho and rho*E cell means over the last step on the dt subset;
 * replaces calculateResidual + its D2H (src/cudarkdgsolver.cu:601-611,
 * src/rkdgadvance.cpp:201-213). */
int rkdg_residual(rkdg_ctx* ctx, double res[2]);

/* One SSP-RK3 step of length dt > 0 on the context's state that allocates nothing: no
 * reference counterpart, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  It runs the kernels
 * of rkdg_stage(0, 1, 2, dt) (and the limiter) on `stream` (NULL: the context's stream)
 * without synchronising; their scratch — the step counters and one pair of residual maxima
 * per workgroup tile, which rkdg_create otherwise allocates per context — is a static
 * device array of the library (one per GPU, kept until the process ends) of
 * RKDG_NOALLOC_MAX_TILES tiles (16.7M elements at 256 per tile, 8.4M at 128): a larger mesh
 * is rejected with RKDG_ERR_ARG.  res_dev (device memory, may be NULL) receives the step's
 * residual pair, as rkdg_residual reports it.  The context's simulated time and step count
 * do not advance and non-physical states are not reported (nothing is read back).  Calls
 * share their GPU's scratch: calls that overlap on one GPU race. */
#define RKDG_NOALLOC_MAX_TILES 65536
int rkdg_step_noalloc(rkdg_ctx* ctx, double dt, double* res_dev, void* stream);

/* Simulated time and steps taken so far (synchronises). */
int rkdg_time(rkdg_ctx* ctx, double* t, int64_t* steps);

/* Per-element limiter flags of the last stage (the .plt FLAG column,
 * src/cudarkdgsolver.cu:984-992); [n_elements] in caller order. */
int rkdg_get_flags(rkdg_ctx* ctx, int32_t* flags);

/* Waits for all work of the context. */
int rkdg_sync(rkdg_ctx* ctx);

/* Times `reps` launches of one stage kernel (stage s) with HIP events on the
 * context's stream; returns the average per launch in ms.  For measurement only. */
int rkdg_time_stage(rkdg_ctx* ctx, int32_t s, int32_t reps, double* avg_ms);

/* Times `nsteps` steps launched stage by stage, with HIP events on the context's stream
 * around the whole loop and around every stage (and limiter) launch.  out[0]: total ms;
 * out[1]: mean ms of a stage launch as bracketed by its own events (includes the dispatch
 * of each launch); out[2]: stage launches; out[3]: mean ms of a limiter launch (0 without
 * limiter).  For measurement only (no reference counterpart). */
int rkdg_time_advance(rkdg_ctx* ctx, int32_t nsteps, double* out);

/* Times `nsteps` steps of rkdg_advance (device dt, hipGraph-replayed step pairs) with one
 * HIP event before and one after, on the context's stream; *total_ms is the elapsed time.
 * For measurement only (no reference counterpart). */
int rkdg_time_steps(rkdg_ctx* ctx, int32_t nsteps, double* total_ms);

/* Error of the current state against the exact solution `kind` (RKDG_IC_VORTEX with its
 * t parameter set, or RKDG_IC_FREESTREAM), on the 36-point rule per element:
 * out[0..3] relative L2 per variable, out[4..7] max abs error at the rule's points. */
int rkdg_error_vs_exact(rkdg_ctx* ctx, int32_t kind, const double* p, int32_t np, double out[8]);

/* Measurement helper: times `reps` launches of a plain device-to-device copy kernel
 * (k_stream_copy) of `bytes` bytes (read once, written once) on `device`; returns the
 * copy bandwidth (bytes read + written per second, GB/s).  Used to quote the measured
 * HBM bandwidth next to the vendor peak and to calibrate memory counters. */
int rkdg_measure_copy(int32_t device, int64_t bytes, int32_t reps, double* gbs);

/* Sizes the library derives: Np, Nqv, Nqe, n_elements, device bytes. */
int rkdg_info(rkdg_ctx* ctx, int64_t out[5]);

const char* rkdg_last_error(rkdg_ctx* ctx);   /* NULL ctx: last global error */
const char* rkdg_strerror(int32_t code);

/* ---- batching form: many independent meshes in one launch per stage --------------- */
/* No reference counterpart (the reference runs one mesh per process; the batch replaces n
 * runs of RkdgAdvance::advance, src/rkdgadvance.cpp:91-243, one per mesh).
 * Per instance: the mesh, the initial state and cfl.  Every other rkdg_params field is one
 * per launch and must agree across the instances (RKDG_ERR_ARG otherwise): degree, flux,
 * gamma, dt_policy (and dt_fixed under RKDG_DT_FIXED, the reduction_threads subset rule
 * under RKDG_DT_REFERENCE), precision, limiter (RKDG_LIMITER_NONE only), device and block;
 * freestream must agree among the instances that have far-field edges. */
int rkdg_batch_create(int32_t n, const rkdg_mesh* meshes, const rkdg_params* params, rkdg_batch** out);
int rkdg_batch_destroy(rkdg_batch* b);
int rkdg_batch_project_ic(rkdg_batch* b, int32_t i, int32_t kind, const double* p, int32_t np);
int rkdg_batch_set_state(rkdg_batch* b, int32_t i, int32_t basis, const double* U);
int rkdg_batch_get_state(rkdg_batch* b, int32_t i, int32_t basis, double* U);
/* nsteps steps of every instance; the stages are captured once in a hipGraph
 * (use_graph != 0) and replayed.  Each instance has its own CFL dt. */
int rkdg_batch_advance(rkdg_batch* b, int32_t nsteps, int32_t use_graph);
int rkdg_batch_time(rkdg_batch* b, int32_t i, double* t);
int rkdg_batch_sync(rkdg_batch* b);
/* Autotuner: times every tile configuration the batch admits (rkdg_batch_set_tiles) on a
 * repeatable stage and keeps the fastest; best_block may be NULL. */
int rkdg_batch_autotune(rkdg_batch* b, int32_t* best_block);
/* Selects one tile configuration: staged != 0 for the staged 256-element tiles (FP64 form:
 * P1/P2; precision forms: P1-P3; when every tile's halo fits its LDS image), else
 * global-gather tiles of block = 64, 128 or 256 elements (the FP32-MFMA form: 64; the
 * BF16-MFMA form has staged tiles only); per_cu 0 launches one workgroup per tile, 4 or 8 a
 * persistent grid of that many 256-thread workgroups per CU.  RKDG_ERR_ARG when the batch
 * does not admit it.  rkdg_batch_get_tiles reports the configuration in use as
 * {staged, block, per_cu}. */
int rkdg_batch_set_tiles(rkdg_batch* b, int32_t staged, int32_t block, int32_t per_cu);
int rkdg_batch_get_tiles(rkdg_batch* b, int32_t out[3]);
int rkdg_batch_info(rkdg_batch* b, int64_t out[4]);   /* instances, total elements, block, device bytes */
const char* rkdg_batch_last_error(rkdg_batch* b);     /* NULL b: last batch error of this thread */

/* ---- host mesh utilities (no GPU) ------------------------------------------------ */
/* Owning mesh container filled by the library. */
typedef struct rkdg_grid {
    int32_t  n_vertices;
    int32_t  n_elements;
    double*  vx;
    double*  vy;
    int32_t* tri;
    int32_t* neighbour;
    int32_t* shared_edge;
    int32_t* boundary;
    int32_t* partition;   /* ElemDepart ids when read from the reference format, else NULL */
} rkdg_grid;

/* nx x ny quads on [x0,x0+lx] x [y0,y0+ly], each split into two triangles; diagonal
 * 0: all along (i,j)-(i+1,j+1), 1: seed-chosen per quad; jitter: interior vertices
 * moved by U(-jitter*h, +jitter*h) per coordinate; periodic != 0 pairs opposite
 * sides, else boundary edges get tag bc. */
int rkdg_grid_splitquad(int32_t nx, int32_t ny, double x0, double y0, double lx, double ly,
                        int32_t periodic, int32_t diagonal, double jitter, uint64_t seed,
                        int32_t bc, rkdg_grid* out);

/* Reads the reference mesh format (mesh.conf + Vertice/ElemVertice/ElemNeigh/ElemEdge/
 * EdgeVertice, headers tolerated) and an optional ElemDepart partition file; boundary
 * tags follow markLocalBoundaryTriangles' barycentre box (+-box; far field outside,
 * wall inside).  Replaces CUnstructuredGrid::initializeGrid/initializeGridNext/
 * markLocalBoundaryTriangles (src/unstructuredgrid.cpp:14-80,193-305,729-760,934-976). */
int rkdg_grid_read(const char* meshconf, const char* partition_file, double box, rkdg_grid* out);

/* Writes g in the reference mesh format into directory dir (which must exist): mesh.conf
 * (the keys CUnstructuredGrid::initializeGrid reads, src/unstructuredgrid.cpp:19-36, with
 * file names dir/<file>), Vertice.dat, ElemVertice.dat, ElemNeigh.dat, ElemEdge.dat,
 * EdgeVertice.dat, and ElemDepart.dat when g->partition is set — the layouts the readers
 * of src/unstructuredgrid.cpp:193-305 parse: one header line per file except ElemNeigh
 * (its header is not skipped, :220) and ElemDepart, 0-based ids, neighbour k across edge k
 * (-1 on a boundary), edge k opposite vertex k, coordinates to 17 significant digits.
 * Boundary tags are not part of the format (rkdg_grid_read derives them). */
int rkdg_grid_write(const char* dir, const rkdg_grid* g);

/* Locality order: by partition id then index (reuse of ElemDepart), or Morton order of
 * barycentres when partition is NULL.  order_out: [n_elements]. */
int rkdg_grid_locality_order(const rkdg_grid* g, int32_t* order_out);

void rkdg_grid_free(rkdg_grid* g);
/* Message of the last failed rkdg_grid_* call on this thread. */
const char* rkdg_grid_last_error(void);
/* Fills a non-owning rkdg_mesh view of g. */
void rkdg_grid_as_mesh(const rkdg_grid* g, rkdg_mesh* m);

/* Tecplot .plt of the cell means, byte-compatible with CCUDARkdgSolver::outputSolution
 * (src/cudarkdgsolver.cu:860-1001).  U: [4][Np][n_elements]; flags may be NULL (all 1). */
int rkdg_write_plt(const char* path, const rkdg_grid* g, int32_t Np, const double* U,
                   const int32_t* flags, double gamma);

/* Host-only conversion of a state [4][Np][n] between the reference's per-element basis
 * and the native basis; ex, ey: [n][3] element vertex coordinates.  No GPU needed. */
int rkdg_convert_basis(int32_t degree, int32_t n, const double* ex, const double* ey, int32_t from_basis,
                       const double* U_in, double* U_out);

/* Raw FP64 state dump / restore (checkpoint): header + [4][Np][E] doubles. */
int rkdg_dump_state(const char* path, int32_t Np, int32_t n_elements, double t, const double* U);
int rkdg_load_state(const char* path, int32_t Np, int32_t n_elements, double* t, double* U);

/* The reference's 'key value # comment' config parser (src/config.cpp:57-112).
 * Returns the value of key into out (empty string if absent). */
int rkdg_config_get(const char* path, const char* key, char* out, int32_t outlen);

#ifdef __cplusplus
}
#endif
#endif /* RKDG_H */
