/*
 * zfv.h — C-ABI of libzfv_hip.so, the MI355X (gfx950) rate-of-change engine.
 *
 * The reference (ZisaFVM) exposes the hot path through two C++ plugin points:
 *
 *   RateOfChange::compute(AllVariables &tendency, const AllVariables &state, double t)
 *       include/zisa/ode/rate_of_change.hpp:37-39
 *   TimeIntegration::compute_step(shared_ptr<AllVariables> u0, double t, double dt)
 *       include/zisa/ode/time_integration.hpp:42-44
 *
 * plus the per-step CFL hook (include/zisa/model/cfl_condition.hpp:17) and the
 * per-stage boundary hook BoundaryCondition::apply
 * (include/zisa/boundary/boundary_condition.hpp:17).  The entry points below
 * replace exactly those, on plain pointers and sizes.  The C++ adapters in
 * zisafvm_amd/csrc/adapters.hpp wrap them back into the reference's class
 * interfaces.
 *
 * Conventions
 *   - State in and out is the reference's AllVariables::cvars layout:
 *     n_cells x 5 row-major FP64 (rho, rho u, rho v, rho w, E); cell order is the
 *     caller's (src/zisa/model/all_variables.cpp:25-30).  Internally the
 *     library renumbers cells along a Hilbert curve and stores structure-of-
 *     arrays; the boundary transposes.
 *   - Every function returns 0 on success and a negative code on failure;
 *     zfv_last_error() gives the message (thread-local).  No exception crosses
 *     the boundary.  (The reference aborts via LOG_ERR instead.)
 *   - A handle owns one HIP stream; use a handle from one host thread at a time.
 */
#ifndef ZFV_H
#define ZFV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZFV_MAX_STENCILS 5

enum { ZFV_CWENO_AO = 0, ZFV_WENO_AO = 1 };
enum { ZFV_FLUX_HLLC = 0, ZFV_FLUX_RUSANOV = 1 };
enum { ZFV_EQ_NONE = 0, ZFV_EQ_ISENTROPIC = 1, ZFV_EQ_ISOTHERMAL = 2 };
enum { ZFV_GRAVITY_NONE = 0, ZFV_GRAVITY_AXIAL = 1, ZFV_GRAVITY_RADIAL = 2, ZFV_GRAVITY_POLYTROPE = 3 };
/* flux-bc.mode "none" / "constant" (FluxBC) / "isentropic" (EquilibriumFluxBC)
 * (include/zisa/experiments/euler_experiment_impl.hpp:277-299) */
enum { ZFV_FLUX_BC_NONE = 0, ZFV_FLUX_BC_CONSTANT = 1, ZFV_FLUX_BC_ISENTROPIC = 2 };

enum {
  ZFV_OK = 0,
  ZFV_ERR_INVALID = -1,   /* bad argument / unsupported configuration */
  ZFV_ERR_HIP = -2,       /* HIP runtime error */
  ZFV_ERR_NUMERIC = -3,   /* implausible state (rho <= 0, E_int <= 0, NaN) */
  ZFV_ERR_NO_DEVICE = -4, /* no gfx950 device visible */
  ZFV_ERR_MEMORY = -5     /* admission / allocation failure */
};

/* The reference's JSON config keys for the path, flattened:
 * reconstruction.{mode,orders,biases,overfit_factors,linear_weights,
 * smoothness_indicator.{epsilon,exponent},steps_per_recompute,recompute_threshold},
 * well-balancing.mode, euler.{eos,gravity}, quadrature.{edge,volume,moments},
 * flux-bc.mode (include/zisa/experiments/euler_experiment_impl.hpp:83-92,411-440;
 * src/zisa/experiments/numerical_experiment.cpp:193-214). */
typedef struct zfv_params {
  int32_t n_dims;                 /* 2 (triangles) or 3 (tetrahedra) */
  int32_t n_stencils;             /* <= ZFV_MAX_STENCILS */
  int32_t orders[ZFV_MAX_STENCILS];
  int32_t biases[ZFV_MAX_STENCILS];  /* 0 = central "c", 1 = one-sided "b" */
  double overfit[ZFV_MAX_STENCILS];
  double linear_weights[ZFV_MAX_STENCILS];
  double epsilon;                 /* smoothness_indicator.epsilon */
  double exponent;                /* smoothness_indicator.exponent */
  int32_t reconstruction;         /* ZFV_CWENO_AO / ZFV_WENO_AO */
  int32_t flux;                   /* ZFV_FLUX_* */
  int32_t equilibrium;            /* ZFV_EQ_* (well-balancing.mode) */
  int32_t gravity;                /* ZFV_GRAVITY_* */
  double gravity_g;               /* constant gravity strength */
  double gravity_axis[3];         /* axial alignment axis */
  double polytrope_rhoC, polytrope_K, polytrope_G;
  double gamma;                   /* ideal gas */
  double gas_constant;
  int32_t edge_deg, volume_deg, moments_deg;  /* quadrature degrees, -1 = max */
  int32_t flux_bc;                /* ZFV_FLUX_BC_*: 0 none, 1 FluxBC (the cell average's
                                     physical flux), 2 EquilibriumFluxBC (the flux of the
                                     cell's isentropic equilibrium at the face points;
                                     with ZFV_EQ_ISENTROPIC only) */
  int32_t steps_per_recompute;    /* >= 1; with k > 1 every cell keeps its equilibrium (theta
                                     and scale) and recomputes it on every k-th rate
                                     evaluation since the last recompute, or earlier when its
                                     (rho, E_int) moves from the kept equilibrium's cell
                                     average by recompute_threshold or more in the kept scale
                                     (include/zisa/reconstruction/local_reconstruction.hpp:87-100);
                                     autotuning and zfv_time_kernels launches recompute every
                                     cell and leave the kept state alone */
  int32_t scaling;                /* 0 EulerScaling (the experiments' choice), 1 UnityScaling
                                     (include/zisa/model/characteristic_scale.hpp:14-46) */
  double recompute_threshold;
} zfv_params;

/* A simplex mesh as the reference's .msh.h5 carries it (vertex_indices), plus
 * per-cell corner coordinates.  Corners are per cell so that periodic meshes
 * can carry unwrapped coordinates; neighbours are found through shared vertex
 * ids (src/zisa/grid/grid.cpp:486-540). */
typedef struct zfv_mesh {
  int32_t n_dims;
  int32_t n_cells;
  const int32_t *vertex_ids;  /* n_cells * (n_dims+1) */
  const double *corners;      /* n_cells * (n_dims+1) * 3 */
  const uint8_t *ghost_mask;  /* n_cells (nonzero = ghost cell, FrozenBC), or NULL */
} zfv_mesh;

typedef struct zfv_handle zfv_handle;

/* Geometry, stencils, LSQ pseudo-inverses and the device image of one mesh.
 * Replaces the setup half of make_reconstruction_array + FluxLoop ctor
 * (include/zisa/reconstruction/global_reconstruction_decl.hpp:105-147,
 *  include/zisa/fvm_loops/flux_loop.hpp:29-90). */
int zfv_create(const zfv_mesh *mesh, const zfv_params *params, zfv_handle **out);
int zfv_destroy(zfv_handle *h);

/* AllVariables::cvars in / out (host memory, caller's cell order). */
int zfv_set_state(zfv_handle *h, const double *cvars);
int zfv_get_state(zfv_handle *h, double *cvars);

/* Snapshot for output while the run goes on (the state TimeLoop::write_output
 * hands to its visualization, src/zisa/fvm_loops/time_loop.cpp:83-85,163-167;
 * AllVariables::cvars, src/zisa/model/all_variables.cpp:74-90).  begin: the
 * state as of all work issued on the handle so far is copied on the device,
 * then to pinned host memory on a second stream, while later steps proceed.
 * end: waits for that copy and writes it like zfv_get_state.  One snapshot
 * at a time (begin while one is pending is an error).  The first begin
 * allocates a device and a pinned host copy of the state.  A batch member's
 * state is complete whenever zfv_batch_run has returned. */
int zfv_snapshot_begin(zfv_handle *h);
int zfv_snapshot_end(zfv_handle *h, double *cvars);

/* FrozenBC steady state (src/zisa/boundary/frozen_boundary_condition.cpp:11-37). */
int zfv_set_steady_state(zfv_handle *h, const double *cvars);

/* RateOfChange::compute on host buffers: dudt += L(u).  The sum of terms is
 * FluxLoop + GravitySourceLoop + FluxBC, in that order
 * (src/zisa/ode/rate_of_change.cpp:21-27). */
int zfv_rate(zfv_handle *h, const double *u, double *dudt, double t);

/* TimeIntegration::compute_step for SSP3 on the device-resident state, with the
 * frozen BC after every stage sum (src/zisa/ode/runge_kutta.cpp:87-112). */
int zfv_step(zfv_handle *h, double t, double dt);

/* n_steps fixed-dt SSP3 steps replayed from one captured HIP graph. */
int zfv_steps(zfv_handle *h, int32_t n_steps, double t, double dt);

/* TimeLoop's inner loop on the device (src/zisa/fvm_loops/time_loop.cpp:36-81):
 * from t0, SSP3 steps with dt = cfl_number * LocalCFL (computed on the device,
 * include/zisa/model/local_cfl_condition_impl.hpp:25-40) clamped to t_end,
 * until t_end or max_steps (<= 0: no limit), or until a step would not advance
 * the time (dt not positive or below the resolution of t).  One captured graph per step;
 * the host reads the loop state back every check_every steps (< 1: 16), so a
 * finished run may replay up to check_every - 1 steps of dt 0 (no change).
 * drho_crit_rel > 0 turns on RejectLargeDensityChange
 * (src/zisa/ode/step_rejection.cpp:17-57): a step that changes the density of
 * any cell by more than drho_crit_rel times its old value is undone and retried
 * with half the factor times its dt (TimeLoop::reject_step,
 * src/zisa/fvm_loops/time_loop.cpp:182-196); the factor doubles after two good
 * steps in a row and stays within [2^-max_exponent, 1].  It starts at 1 on every
 * call.  drho_crit_rel <= 0 rejects nothing (RejectNothing).
 * Writes the final time, the number of accepted steps and the number of
 * rejected steps (each pointer may be null). */
int zfv_advance(zfv_handle *h, double cfl_number, double t0, double t_end, int32_t max_steps, int32_t check_every,
                double drho_crit_rel, int32_t max_exponent, double *t_final, int32_t *steps_done,
                int32_t *rejected);

/* The device time loop of zfv_advance (same arguments, same steps, same
 * results) for callers that cannot allocate: code that must not call
 * hipMalloc or instantiate graphs (inside a capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.  It
 * allocates nothing: the loop state and, with step rejection, the backup of
 * the state live in a file-scope device array of 16 MiB that each GPU the
 * library runs on holds from module load until the process ends; steps are
 * launched one by one (no graph) with the handle's tuned launch configuration
 * (one wavefront per block if it was never tuned).  Limit: with step
 * rejection (drho_crit_rel > 0) the state must fit 16 MiB - 256 B
 * (nvars * n_cells * 8 bytes: 524,280 triangles or 419,424 tetrahedra),
 * else ZFV_ERR_MEMORY.  The handle's device must be current.  Calls on one
 * GPU share the array: calls that overlap on one GPU (from several host
 * threads) race and must not be issued. */
int zfv_advance_noalloc(zfv_handle *h, double cfl_number, double t0, double t_end, int32_t max_steps,
                        int32_t check_every, double drho_crit_rel, int32_t max_exponent, double *t_final,
                        int32_t *steps_done, int32_t *rejected);

/* LocalCFL::operator() on the device-resident state
 * (include/zisa/model/local_cfl_condition_impl.hpp:25-40). */
int zfv_cfl_dt(zfv_handle *h, double cfl_number, double *dt);

int zfv_synchronize(zfv_handle *h);

typedef struct zfv_info {
  int64_t n_cells;
  int64_t n_faces;            /* interior faces evaluated by the face kernel */
  int64_t device_bytes;       /* HBM held by the handle */
  double bytes_per_cell_stage;/* algorithmic bytes, all kernels of one stage */
  double recon_bytes_per_cell;/* algorithmic bytes of the reconstruction kernel */
  int32_t max_stencil_size;   /* largest combined stencil */
  int32_t n_points_face;      /* quadrature points per face */
  int64_t n_far_ghost_cells;  /* ghost cells with no interior neighbour: once a
                                 steady state is set, steps leave them frozen
                                 and do not reconstruct them */
} zfv_info;
int zfv_get_info(zfv_handle *h, zfv_info *info);

/* Combined stencil (local2global) of cell i, caller's numbering, for parity
 * checks against an independent stencil selection.  Writes up to cap ids; the
 * per-stencil member lists follow in `members` as
 * [n_stencils][order, size, id_0 .. id_{size-1}] flattened.  Returns the
 * number of ints written to `members` (or < 0 on error). */
int zfv_get_stencils(zfv_handle *h, int32_t i, int32_t *l2g, int32_t cap, int32_t *l2g_size,
                     int32_t *members, int32_t members_cap);

/* Runs n_steps SSP3 steps of fixed dt launch by launch (the state advances)
 * with HIP events on the handle's stream around every kernel; returns the
 * milliseconds summed over the launches of each kernel. */
typedef struct zfv_kernel_times {
  double reconstruct_ms, face_flux_ms, update_ms;
  int32_t n_launches;
} zfv_kernel_times;
int zfv_time_kernels(zfv_handle *h, int32_t n_steps, double dt, zfv_kernel_times *out);

/* Launch configuration of the reconstruction kernel.  No counterpart in the
 * reference.  A cell is served by a group of nvars lanes, 64 / nvars cells per
 * wavefront; a block holds 1, 2 or 4 wavefronts (64, 128, 256 threads).  The
 * first rate or step call tunes it on the instance's own data by timing the
 * three candidates with HIP events; results never depend on it.  zfv_tune
 * re-tunes now (rec_threads = 0), fixes the configuration (rec_threads 64,
 * 128 or 256) or only reports it (-1). */
typedef struct zfv_tuning {
  int32_t rec_threads;      /* threads per reconstruction block in use */
  int32_t cells_per_block;  /* cells per reconstruction block in use */
  int32_t tuned;            /* 1 once a configuration is chosen */
  int32_t reserved;
  double rec_ms[3];         /* average ms per launch at 64, 128, 256 threads */
} zfv_tuning;
int zfv_tune(zfv_handle *h, int32_t rec_threads, zfv_tuning *out);

/* ---- Batches: many instances of one kernel shape advanced together.
 * Has no counterpart in the reference (which runs one experiment per
 * process); each member behaves exactly like its own TimeLoop
 * (src/zisa/fvm_loops/time_loop.cpp:36-75): per step, dt = cfl * LocalCFL
 * clamped to t_end, one SSP3 step with the frozen BC.  All members of a batch
 * advance in one captured HIP graph per step (block tables map launched
 * blocks to members; no per-member launches).  An instance is a member of
 * at most one batch; zfv_destroy of a member removes it from its batch first,
 * zfv_set_steady_state on a member reaches the batch, and a snapshot
 * (zfv_snapshot_begin) of a member is ordered before the batch's later
 * steps.  Read a member's state with zfv_get_state after zfv_batch_run
 * returns. */
typedef struct zfv_batch zfv_batch;
/* Capacity: total cells / interior faces / members the batch can hold.
 * Batches above 65,536 cells or 8 members are refused (ZFV_ERR_INVALID):
 * those are the largest run on the GPU with checked results, and a batched
 * run of 64 members and 4.0M cells ended in a GPU memory fault whose cause
 * has not been found. */
int zfv_batch_create(const zfv_params *shape_params, int64_t cell_capacity, int64_t face_capacity,
                     int32_t max_instances, zfv_batch **out);
int zfv_batch_destroy(zfv_batch *b);
/* Joins at the next step boundary.  The state is the handle's current state
 * (zfv_set_state).  max_steps <= 0: no step limit. */
int zfv_batch_add(zfv_batch *b, zfv_handle *h, double cfl, double t0, double t_end, int32_t max_steps,
                  int32_t *slot);
int zfv_batch_remove(zfv_batch *b, int32_t slot);
/* n_steps graph replays; members that reach t_end / max_steps stop. */
int zfv_batch_run(zfv_batch *b, int32_t n_steps);
/* zfv_batch_run in two halves, for callers that keep several batches in
 * flight at once (each batch has its own stream, so batches of different
 * shapes overlap on the device): launch enqueues n_steps steps and returns,
 * wait blocks until they have run.  Every other call on the batch or on one
 * of its members waits for steps in flight first. */
int zfv_batch_launch(zfv_batch *b, int32_t n_steps);
int zfv_batch_wait(zfv_batch *b);
int zfv_batch_query(zfv_batch *b, int32_t slot, double *t, double *dt, int32_t *steps_done, int32_t *active);
/* Block-to-XCD mapping: 1 (default) gives each XCD a contiguous eighth of the
 * cells so stencil gathers share one L2; 0 is the hardware round robin. */
/* Reconstruction launch geometry of a batch, as zfv_tune; tuned on the
 * members present at the first zfv_batch_run. */
int zfv_batch_tune(zfv_batch *b, int32_t rec_threads, zfv_tuning *out);
int zfv_batch_set_xcd_chunks(zfv_batch *b, int32_t on);

/* Device bytes zfv_create would allocate for this mesh (host-only estimate
 * from the kernel shape and the cell count; used for HBM admission). */
int zfv_estimate_device_bytes(const zfv_mesh *mesh, const zfv_params *params, int64_t *bytes);

/* ---- Serving: continuous batching of solve requests (no counterpart in the
 * reference).  A request is a mesh, parameters, initial state (and FrozenBC
 * steady state) and an end time; it runs its own time loop (CFL every step,
 * clamped to t_end).  Requests are admitted while their estimated device
 * bytes fit `hbm_budget_bytes`; setup workers build admitted instances while
 * the GPU keeps stepping; built instances join the batch of their kernel
 * shape at the next step boundary and leave it when they finish, returning
 * their memory to the budget. */
typedef struct zfv_server zfv_server;
enum {
  ZFV_REQ_QUEUED = 0,  /* waiting for HBM budget */
  ZFV_REQ_SETUP = 1,   /* admitted; host setup + upload in progress */
  ZFV_REQ_RUNNING = 2, /* stepping on the device */
  ZFV_REQ_DONE = 3,
  ZFV_REQ_FAILED = 4
};
typedef struct zfv_request_status {
  int32_t state;
  int32_t steps;        /* SSP3 steps taken */
  int64_t n_cells;
  double t_final;       /* simulation time reached */
  double admit_s;       /* seconds from submit to admission (-1: not yet) */
  double start_s;       /* seconds from submit to joining a batch */
  double latency_s;     /* seconds from submit to result (-1: not finished) */
} zfv_request_status;
typedef struct zfv_server_stats_t {
  int64_t reserved_bytes; /* HBM estimate of admitted, unfinished requests */
  int32_t pending;        /* waiting for HBM */
  int32_t in_setup;       /* admitted, instance being built or queued for a setup worker */
  int32_t running;        /* members of a batch */
  int32_t ready;          /* built, joining at the next step boundary */
  int64_t cell_stages;  /* cell-stage updates completed so far */
  int64_t steps_run;    /* step boundaries passed by the GPU loop */
} zfv_server_stats_t;
/* batch_cell_capacity and max_members size the server's batches (one per
 * kernel shape, created on first use): within zfv_batch_create's limits. */
int zfv_server_create(int64_t hbm_budget_bytes, int64_t batch_cell_capacity, int32_t max_members,
                      int32_t setup_workers, zfv_server **out);
/* Copies the request; returns at once with its id. */
int zfv_server_submit(zfv_server *s, const zfv_mesh *mesh, const zfv_params *params, const double *u0,
                      const double *steady, double t_end, double cfl, int32_t max_steps, int64_t *id);
int zfv_server_poll(zfv_server *s, int64_t id, zfv_request_status *status);
/* Final state (n_cells x 5, caller order) of a finished request. */
int zfv_server_result(zfv_server *s, int64_t id, double *u_out);
int zfv_server_release(zfv_server *s, int64_t id);
int zfv_server_stats(zfv_server *s, zfv_server_stats_t *out);
int zfv_server_destroy(zfv_server *s);

/* Streaming copy of `bytes` bytes between two device buffers, `iters` times,
 * timed with HIP events: the measured HBM copy rate (read + write bytes per
 * second) used to calibrate the roofline and the PMC byte counters.  Has no
 * counterpart in the reference. */
int zfv_bandwidth_probe(int64_t bytes, int32_t iters, double *gbytes_per_s);

const char *zfv_last_error(void);

/* ---- Synthetic workloads (mesh generation is not in the reference, which loads
 * meshes from files; these produce the seeded jittered meshes of the
 * benchmark configurations). */
typedef struct zfv_mesh_buf zfv_mesh_buf;
/* 2D: n x n jittered lattice on [0,1]^2, each square split along its Delaunay
 * diagonal (2 n^2 triangles).  periodic_x / periodic_y wrap the lattice; a
 * non-periodic direction gets `ghost_layers` layers of ghost cells at each end
 * (cells whose centroid lies outside [0,1] after the domain is extended). */
int zfv_mesh_lattice2d(int32_t n, double jitter, uint64_t seed, int32_t periodic_x,
                       int32_t periodic_y, int32_t ghost_layers, zfv_mesh_buf **out);
/* 3D: jittered cubic lattice of the ball of radius `radius` (Kuhn subdivision,
 * 6 tets per cube, cells kept when their centroid is inside the ball); cells
 * whose centroid radius exceeds ghost_radius are ghosts. */
int zfv_mesh_ball3d(int32_t n, double radius, double ghost_radius, double jitter, uint64_t seed,
                    zfv_mesh_buf **out);
int zfv_mesh_view(const zfv_mesh_buf *buf, zfv_mesh *view);
int zfv_mesh_free(zfv_mesh_buf *buf);

#ifdef __cplusplus
}
#endif
#endif /* ZFV_H */
