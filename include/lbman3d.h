/*
 * lbman3d.h — C ABI of the MI355X-native lbman3d solver step.
 *
 * One handle owns one simulation on one GPU: D3Q15 populations, density,
 * velocity, body force and the five independent Q-tensor components, all in
 * HBM as structure-of-arrays with cell = (z*ny + y)*nx + x (x fastest).
 * Every entry point returns LBM_OK (0) or a negative LBM_ERR_* code; the
 * message of the last failure on the calling thread is lbm_last_error().
 *
 * The entry points replace the reference's compile-time C++ device API:
 *   DeviceSolver<BC>::Initialize / QTensorStep / LBMStep   (src/device_solver.h:9-18)
 *   DeviceFields::DeviceFields / Initialize / CopyToHost    (src/device_fields.h:23-80)
 *   InitializeComputeBackend                                (src/device_fields.h:21)
 *   SimIO::Log reductions and TotalNematicFreeEnergy        (src/sim_io.cc:78-165,
 *                                                            src/analysis_fields.h:39-72)
 * Boundary conditions, physical constants and the advection scheme, which the
 * reference fixes at compile time (src/boundary.h, src/params.h,
 * src/sim_config.h:99), are runtime fields of lbm_config here, so one library
 * serves every BCConfig and the batched form can mix them.
 *
 * Threading: a handle is not thread-safe; distinct handles may be driven from
 * distinct threads. Each handle owns one HIP stream at the default priority.
 * Download, diagnostics and timing calls synchronise that stream.
 */
#ifndef LBMAN3D_H_
#define LBMAN3D_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
#define LBM_OK               0
#define LBM_ERR_INVALID     -1  /* bad argument, size or configuration        */
#define LBM_ERR_UNSUPPORTED -2  /* well-formed but outside what is supported  */
#define LBM_ERR_HIP         -3  /* HIP runtime failure (allocation, launch)   */
#define LBM_ERR_STATE       -4  /* call not valid in the handle's state       */

/* ---- boundary tags (src/boundary.h:31-50) ------------------------------ */
/* Q-tensor wall policy: Periodic, Neumann, Anchoring<S,theta,phi>. */
#define LBM_Q_PERIODIC   0
#define LBM_Q_NEUMANN    1
#define LBM_Q_ANCHORING  2
/* Velocity wall policy: Periodic, NoSlip, MovingWall<Ux,Uy,Uz>, SpecularReflection. */
#define LBM_U_PERIODIC   0
#define LBM_U_NOSLIP     1
#define LBM_U_MOVING     2
#define LBM_U_SPECULAR   3

/* Face order of lbm_config.walls (src/boundary.h:82-97, XLo..ZHi). */
#define LBM_XLO 0
#define LBM_XHI 1
#define LBM_YLO 2
#define LBM_YHI 3
#define LBM_ZLO 4
#define LBM_ZHI 5

/* Advection scheme for u.grad(Q) (src/qtensor_types.h:22). */
#define LBM_ADV_CENTRED  0
#define LBM_ADV_UPWIND   1

/* Storage/arithmetic precision of the step. */
#define LBM_FP64          0  /* plain FP64, reference-parity path             */
#define LBM_FP32_SHIFTED  1  /* FP32, populations stored as f_i - w_i*rho0    */

/* Body-force model applied after the Q update.
 * FULL: Ericksen force + active stress + passive stress divergence - mu*u
 *       (QTensorSolver::SetActiveStressAndComputeBodyForce, src/qtensor_solver.tpp:155-217).
 * ERICKSEN_ONLY: the override with an empty body used by the reference's
 *       zero-activity tests (tests/integration/test_qtensor_relaxation.cc:21-31),
 *       so the force is the Ericksen force alone. */
#define LBM_FORCE_FULL           0
#define LBM_FORCE_ERICKSEN_ONLY  1

/* One face: WallSpec<QBC, UBC> (src/boundary.h:72-76). */
typedef struct lbm_wall {
    int32_t qbc;            /* LBM_Q_*                                   */
    int32_t ubc;            /* LBM_U_*                                   */
    double  anchor_s;       /* Anchoring<S, theta, phi> (radians)        */
    double  anchor_theta;
    double  anchor_phi;
    double  wall_u[3];      /* MovingWall<Ux, Uy, Uz>                    */
} lbm_wall;

/* Physical constants (namespace Params, src/params.h:29-112). */
typedef struct lbm_params {
    double dt;       /* DT                                   */
    double density;  /* kDensity (rho0)                      */
    double tauf;     /* TAUF                                 */
    double L;        /* Frank elasticity                     */
    double A, B, C;  /* Landau-de Gennes coefficients        */
    double lambda;   /* LAMBDA, flow alignment               */
    double gamma;    /* GAMMA, inverse rotational viscosity  */
    double alpha;    /* ALPHA, activity zeta                 */
    double mu;       /* MU, linear friction                  */
} lbm_params;

typedef struct lbm_config {
    int32_t    nx, ny, nz;
    int32_t    advection;    /* LBM_ADV_*   */
    int32_t    precision;    /* LBM_FP64 or LBM_FP32_SHIFTED */
    int32_t    force_model;  /* LBM_FORCE_* */
    lbm_wall   walls[6];     /* LBM_XLO .. LBM_ZHI */
    lbm_params p;
} lbm_config;

/* FP64 domain reductions, SimIO::Log (src/sim_io.cc:78-165) plus the
 * nematic free energy (src/analysis_fields.h:39-72). */
typedef struct lbm_diag {
    double mass;          /* sum rho                                   */
    double px, py, pz;    /* sum rho*u                                 */
    double kinetic;       /* sum rho*|u|^2 / 2                         */
    double du2;           /* sum |u - u_prev|^2 against the previous call */
    double u2;            /* sum |u|^2                                 */
    double umax;          /* max |u|                                   */
    double free_energy;   /* total Landau-de Gennes free energy        */
    int32_t nonfinite;    /* 1 if mass or momentum is NaN/Inf           */
    int32_t _pad;
} lbm_diag;

typedef struct lbm_solver* lbm_handle;

/* ---- configuration ------------------------------------------------------ */
/* Fill cfg with the params.h defaults (src/params.h:14-112) on an
 * nx x ny x nz FullyPeriodicConfig grid, centred advection, FP64. */
int lbm_default_config(lbm_config* cfg, int32_t nx, int32_t ny, int32_t nz);
/* Apply a named preset (src/boundary.h:105-129, src/sim_config.h:19-45):
 * "FullyPeriodic", "Channel", "SlitFreeSlip", "SlitNoSlip". */
int lbm_set_preset(lbm_config* cfg, const char* name);

/* ---- lifetime (DeviceFields ctor / dtor, InitializeComputeBackend) ------ */
int  lbm_create(const lbm_config* cfg, lbm_handle* out);
int  lbm_destroy(lbm_handle h);
const char* lbm_last_error(void);
int  lbm_device_info(char* buf, size_t len);
int  lbm_get_config(lbm_handle h, lbm_config* out);

/* ---- state transfer ----------------------------------------------------- */
/* Host arrays are FP64 SoA of V = nx*ny*nz cells; q holds Qxx,Qxy,Qxz,Qyy,Qyz
 * as 5 consecutive blocks of V. Any pointer may be NULL to leave that field
 * unchanged. lbm_upload_state sets the populations to Feq(rho, u)
 * (LbmSolver::Initialize, src/lbm_solver.tpp:59-87) and leaves the body
 * force as it is (zero after lbm_create, as FluidFields starts it,
 * src/fluid_fields.cc:6-17). */
int lbm_upload_state(lbm_handle h, const double* rho, const double* ux,
                     const double* uy, const double* uz, const double* q);
/* Populations f_i(x) as 15 blocks of V (i*V + cell, the reference device
 * layout, src/local_grid.h:64). They are taken as the post-streaming state. */
int lbm_upload_populations(lbm_handle h, const double* f);
int lbm_upload_force(lbm_handle h, const double* fx, const double* fy, const double* fz);

int lbm_download_state(lbm_handle h, double* rho, double* ux, double* uy,
                       double* uz, double* q);
int lbm_download_populations(lbm_handle h, double* f);
int lbm_download_force(lbm_handle h, double* fx, double* fy, double* fz);

/* ---- stepping ------------------------------------------------------------ */
/* Q update + body force (ActiveNematicSim::QTensorStep, src/active_nematic.h:84-94). */
int lbm_qtensor_step(lbm_handle h);
/* Collide-stream with the current body force (ActiveNematicSim::LBMStep). */
int lbm_lbm_step(lbm_handle h);
/* n full steps, QTensorStep then LBMStep each (ActiveNematicSim::Step). */
int lbm_step(lbm_handle h, int32_t n);
int lbm_synchronize(lbm_handle h);

/* ---- diagnostics --------------------------------------------------------- */
int lbm_diagnostics(lbm_handle h, lbm_diag* out);
/* Order parameter (largest eigenvalue of Q) and director per cell
 * (QtensorToOrderDirector, src/analysis_fields.cc:59-93); director is
 * [V][3] interleaved. Computed on the device in FP64. */
int lbm_order_director(lbm_handle h, double* order, double* director);
/* lbm_order_director without any device allocation. This entry point has no
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc). The fields
 * are staged in a fixed 16 MiB device array per GPU, so grids are limited to
 * 524,288 cells (larger ones return LBM_ERR_INVALID); calls that overlap on
 * one GPU race on that array. Results equal lbm_order_director's. */
int lbm_order_director_noalloc(lbm_handle h, double* order, double* director);

/* Plaquettes pierced by a disclination line (DefectFinder::ComputeWindingNumbers,
 * src/analysis/defect_finder.tpp:12-162): 1 where the director of
 * lbm_order_director, carried around the plaquette with a sign flip whenever
 * n1 . n2 <= 0 (FlipN, src/analysis/defect_fields.cc:15-24), returns flipped,
 * else 0. Layouts as DefectFields (src/analysis/defect_fields.h:17-30):
 * def_x: nx*(ny-1)*(nz-1) plaquettes normal to x, index (z*(ny-1) + y)*nx + x;
 * def_y: (nx-1)*ny*(nz-1), index (z*ny + y)*(nx-1) + x;
 * def_z: (nx-1)*(ny-1)*nz, index (z*(ny-1) + y)*(nx-1) + x. */
int lbm_defect_faces(lbm_handle h, uint8_t* def_x, uint8_t* def_y, uint8_t* def_z);

/* ---- timing and kernel selection ------------------------------------------ */
/* Run n full steps and return the elapsed device time (ms) measured with HIP
 * events on the handle's stream. */
int lbm_time_steps(lbm_handle h, int32_t n, double* ms);
/* Per-kernel mean duration (ms) over n full steps, events bracketing every
 * launch on the handle's stream. names is a ';'-separated list of kernel
 * labels in the order of ms_out; at most max entries are written. */
int lbm_profile_kernels(lbm_handle h, int32_t n, double* ms_out, int32_t max,
                        char* names, size_t names_len, int32_t* count);
/* Kernel schedule. 0 = automatic (autotuned on first use), 1 = two-kernel
 * split, 2.. = a fixed fused tile shape. FP32 tiles: lbm_tile_count /
 * lbm_tile_name. FP64: fully periodic domains also have fused single-kernel
 * tiles, whose results equal the two-kernel split bit for bit; other domains
 * run the two-kernel split. */
int lbm_set_schedule(lbm_handle h, int32_t schedule);
int lbm_get_schedule(lbm_handle h, int32_t* schedule);
int lbm_tile_count(void);
int lbm_tile_name(int32_t schedule, char* buf, size_t len);
/* Name of a schedule of this handle's precision ("auto:...", "two-kernel",
 * "fused:..."). */
int lbm_schedule_name(lbm_handle h, int32_t schedule, char* buf, size_t len);
/* Time every schedule of the handle's precision for n steps from a snapshot
 * of the state (restored after each trial), then the three fastest again for
 * 8n steps each, and keep the fastest of those for the automatic schedule;
 * results per schedule (1, 2, ...) in ms_out[count] (ms per step, the
 * longer trial's for the three). */
int lbm_autotune(lbm_handle h, int32_t n, double* ms_out, int32_t max, int32_t* count);

/* HBM calibration: device-to-device copy of `bytes` (16-byte accesses),
 * repeated `reps` times; returns read+write GB/s and ms per copy. */
int lbm_bandwidth_probe(size_t bytes, int32_t reps, double* gbs, double* ms_per_copy);

/* ---- host helpers -------------------------------------------------------- */
/* Seeded random-director initial Q: per cell a director uniform on the unit
 * sphere and S = s0 + U(-noise, noise), Q = S (n n - I/3), from a
 * counter-based splitmix64 stream keyed on (seed, cell, k). q is 5 blocks
 * of V as in lbm_upload_state. Host-only, FP64. */
int lbm_seed_director_field(uint64_t seed, int32_t nx, int32_t ny, int32_t nz,
                            double s0, double noise, double* q);

/* Pull-slot table the library derives from cfg's walls (64 node classes x 15
 * slots; bits [1:0] kind 0 stream / 1 bounce-back / 2 mirror, [5:2] source
 * direction, [8:6] face). Host-only; for inspection and tests. Returns
 * LBM_ERR_UNSUPPORTED for wall combinations whose reference result depends on
 * write order or keeps stale populations. */
int lbm_debug_slot_table(const lbm_config* cfg, uint16_t* table);
/* The cells within w cells of an open face of cfg's grid, in the order the
 * wall-shell passes of a walled step enumerate them (x, y, z per cell, each
 * cell once). With xyz NULL only *count is set. Host-only; for tests. No
 * counterpart in the reference: the shell passes are how this library splits
 * a step on a grid with walls. */
int lbm_debug_shell_region(const lbm_config* cfg, int32_t w, int32_t* xyz, int64_t max, int64_t* count);

/* How the fused step maps a coordinate t outside [0, n) back into the domain:
 * kind 0 for z planes (requested from two cells below a face to one above
 * it), kind 1 for the x and y columns of a tile (any distance); a periodic
 * side wraps, any other side clamps. Host-only; for tests. No counterpart in
 * the reference: the tiles of the fused step are this library's own. */
int lbm_debug_tile_map(int32_t kind, int32_t t, int32_t n, int32_t lo_per, int32_t hi_per, int32_t* out);

/* ---- batched form: many independent simulations in one launch per step ---- */
typedef struct lbm_batch* lbm_batch_handle;
/* All configs must be LBM_FP32_SHIFTED. */
int lbm_batch_create(const lbm_config* cfgs, int32_t count, lbm_batch_handle* out);
int lbm_batch_destroy(lbm_batch_handle b);
int lbm_batch_upload_state(lbm_batch_handle b, int32_t sim, const double* rho,
                           const double* ux, const double* uy, const double* uz,
                           const double* q);
int lbm_batch_download_state(lbm_batch_handle b, int32_t sim, double* rho,
                             double* ux, double* uy, double* uz, double* q);
/* n steps of every simulation; with use_graph the per-step launch sequence
 * is captured once in a hipGraph and replayed. */
int lbm_batch_step(lbm_batch_handle b, int32_t n, int32_t use_graph);
int lbm_batch_time_steps(lbm_batch_handle b, int32_t n, int32_t use_graph, double* ms);
int lbm_batch_diagnostics(lbm_batch_handle b, int32_t sim, lbm_diag* out);
/* Tile shapes of the batched kernel: count, select, query, and an autotuner
 * that times every shape on scratch copies of the states and keeps the
 * fastest (ms per step in ms_out[count]). */
int lbm_batch_tile_count(void);
int lbm_batch_tile_name(int32_t tile, char* buf, size_t len);
int lbm_batch_set_tile(lbm_batch_handle b, int32_t tile);
int lbm_batch_get_tile(lbm_batch_handle b, int32_t* tile);
int lbm_batch_autotune(lbm_batch_handle b, int32_t n, double* ms_out, int32_t max, int32_t* count);
/* Cap the persistent grid of the batched kernels at max_blocks blocks (0, the
 * default: as many as fit on the device at once). With a grid smaller than
 * the work list, blocks take several tiles and the final round of the work
 * list is split; results do not depend on the cap. For tests and tuning. */
int lbm_batch_set_grid_limit(lbm_batch_handle b, int32_t max_blocks);

/* ---- learned closure for a 2x-coarsened FP32 simulation -------------------
 * No counterpart in the reference. A periodic 3x3x3 convolutional network
 * 8 -> hidden -> hidden -> 5 (ReLU between layers) maps the coarse state
 * x = (Q 5, u 3) to a correction dQ of the coarse step's Q. BF16 matrix-core
 * kernels with FP32 accumulation; FP32 master weights and Adam state.
 * Fields on the host are channels-first FP64: x is 8 blocks of V (Q xx, xy,
 * xz, yy, yz, then ux, uy, uz), y is 5 blocks of V. nx must be a multiple of
 * 16 and ny, nz multiples of 4 (the convolutions work on 16 x 4 x 4 cell
 * tiles); hidden must be 32. Parameters, layer by layer: W[27][cin][cout] with
 * stencil offset (dz+1)*9 + (dy+1)*3 + (dx+1), then b[cout]. */
typedef struct lbm_closure* lbm_closure_handle;
int lbm_closure_create(int32_t nx, int32_t ny, int32_t nz, int32_t hidden, uint64_t seed,
                       lbm_closure_handle* out);
int lbm_closure_destroy(lbm_closure_handle c);
int lbm_closure_param_count(lbm_closure_handle c, int64_t* n);
int lbm_closure_get_params(lbm_closure_handle c, double* w);
/* Replaces the weights and resets the optimizer state. */
int lbm_closure_set_params(lbm_closure_handle c, const double* w);
/* Network input = x * in_scale[8]; network output = y * out_scale[5]. */
int lbm_closure_set_normalization(lbm_closure_handle c, const double* in_scale, const double* out_scale);
/* Training samples are kept on the device (BF16 inputs, FP32 targets). */
int lbm_closure_add_sample(lbm_closure_handle c, const double* x, const double* y, int32_t* index);
int lbm_closure_clear_samples(lbm_closure_handle c);
/* iters Adam steps (beta 0.9/0.999, eps 1e-8) on the mean-squared error of
 * one sample each, drawn from the seeded stream; mean_loss (may be NULL)
 * is the mean normalised loss over the iterations. */
int lbm_closure_train(lbm_closure_handle c, int32_t iters, double lr, uint64_t seed, double* mean_loss);
/* Normalised loss and its gradient (param_count values, may be NULL) for
 * one sample; does not change the weights. */
int lbm_closure_loss_and_grad(lbm_closure_handle c, const double* x, const double* y, double* loss,
                              double* grad);
/* y = network(x) in physical units (the output scale undone). */
int lbm_closure_predict(lbm_closure_handle c, const double* x, double* y);
/* Adds alpha * network(Q, u) to the Q of an FP32 simulation of the same grid.
 * Asynchronous: enqueued on the simulation's stream after the closure's
 * pending work (later closure calls wait for it); calls on one closure must
 * not overlap from several threads. */
int lbm_closure_apply(lbm_closure_handle c, lbm_handle sim, double alpha);
/* n coarse steps of sim, each followed by the closure correction (alpha as
 * in lbm_closure_apply), launched back to back on the simulation's stream;
 * ms (optional) receives their device time from HIP events. */
int lbm_closure_run(lbm_closure_handle c, lbm_handle sim, int32_t n, double alpha, double* ms);
/* Mean time of one forward pass over the whole grid (ms), n passes. */
int lbm_closure_time_forward(lbm_closure_handle c, int32_t n, double* ms);
/* D = A (16x32, row-major) * B (32x16) through the BF16 matrix-core
 * fragment mapping the closure kernels use (inputs rounded to BF16). */
int lbm_debug_mfma_bf16(const float* a, const float* b, float* d);

#ifdef __cplusplus
}
#endif

#endif /* LBMAN3D_H_ */
