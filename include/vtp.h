/*
 * vtp.h - C ABI of the MI355X particle-fluid step (PBF, DFSPH, IMM, IMM-CT).
 *
 * Plain C: opaque context, plain pointers and sizes, status codes, no
 * exceptions and no C++ or torch types across the boundary.  Every host-side
 * array handed in or out is in ATTACH order (the order particles were added),
 * exactly as the reference keeps its host-visible arrays.
 *
 * One context owns the device state of one solver instance (all particle
 * arrays, the uniform grid, captured graphs and one HIP stream).  Contexts
 * share nothing, so several may live in one process (the reference's
 * function-static locals made its solvers per-process singletons; this ABI
 * has no global mutable state).  A context is used by one host thread at a
 * time.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   Simulator/Runtime/Include/Framework/Solver.hpp:27-62        (Solver virtuals)
 *   Simulator/Runtime/Source/Manager/SolverManager.cpp:19-40    (createSolver)
 *   Simulator/Runtime/Source/Solvers/PBF/PBFCudaApi.cuh:15-56   (pbf:: launchers)
 *   Simulator/Runtime/Source/Solvers/DFSPH/DFSPHCudaApi.cuh     (dfsph:: launchers)
 *   Simulator/Runtime/Source/Solvers/IMM/IMMCudaApi.cuh         (imm:: launchers)
 *   Simulator/Runtime/Source/Solvers/IMM-CT/IMMCTCudaApi.cuh    (immct:: launchers)
 *   Simulator/Runtime/Include/Modules/NeighborSearch/UGNS/UniformGridNeighborSearch.hpp:65-92
 */
#ifndef VTP_H
#define VTP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vtp_ctx vtp_ctx;

/* Status codes.  The reference reports failure as `bool false` plus a log
 * line and a sticky crash flag (PBFSolver.cpp:370-374); here every call
 * returns a code and vtp_last_error() holds the message. */
typedef enum vtp_status {
    VTP_OK = 0,
    VTP_EINVAL = 1,   /* bad argument / missing config key / size mismatch  */
    VTP_ENOMEM = 2,   /* device or host allocation failed                   */
    VTP_EHIP = 3,     /* HIP runtime error (launch, copy, ...)              */
    VTP_ECRASH = 4,   /* solver crashed: iteration cap, NaN, out-of-grid    */
    VTP_ESTATE = 5,   /* call out of order (e.g. step before initialize)    */
    VTP_ENODEV = 6,   /* no usable GPU / library built without device code  */
    VTP_ENBRCAP = 7   /* a neighbour list reached maxNeighborNum where the call
                         needs complete (symmetric) lists: the PBF adjoint    */
} vtp_status;

/* Solver.hpp:20-25 (eSolverType). */
typedef enum vtp_solver_type {
    VTP_PBF = 0,
    VTP_DFSPH = 1,
    VTP_IMM = 2,
    VTP_IMMCT = 3
} vtp_solver_type;

/* Material.hpp:14-22 */
enum { VTP_MAT_FLUID = 0, VTP_MAT_BOUNDARY = 1 };

enum { VTP_MAX_PHASES = 8, VTP_MAX_OBJECTS = 1024 };

/* Crash-word bits (vtp_step_stats.crash_flags). */
enum {
    VTP_CRASH_DIV_ITER_CAP = 1u << 0,    /* DFSPHCudaImpl.cu:499-502 */
    VTP_CRASH_INCOMP_ITER_CAP = 1u << 1, /* DFSPHCudaImpl.cu:607-610 */
    VTP_CRASH_PHASE_ITER_CAP = 1u << 2,  /* IMMCudaImpl.cu:929-932   */
    VTP_CRASH_NAN = 1u << 3,             /* non-finite position       */
    VTP_CRASH_OUT_OF_GRID = 1u << 4      /* UGNSCudaImpl.cu:122-123   */
};

/*
 * Solver parameters.  Field names follow the JSON keys of the reference's
 * templates (ConfigTemplates/<Solver>/<Solver>ConfigTemplate.json); the
 * derived constants (h = 4r, V0 = (2r)^3, 1/dt, ...) are computed inside the
 * library exactly where the reference computes them (PBFSolver.cpp:43-58).
 */
typedef struct vtp_params {
    int32_t solver;                  /* vtp_solver_type                           */
    float animation_time;            /* Required.animationTime                    */
    float time_step;                 /* Required.timeStep                         */
    float particle_radius;           /* Required.particleRadius                   */
    float sim_space_lb[3];           /* Required.simSpaceLB                       */
    float sim_space_size[3];         /* Required.simSpaceSize                     */
    uint32_t max_neighbor_num;       /* Required.maxNeighborNum (self included)   */
    float gravity[3];                /* (0,-9.8,0) unless Optional.enable         */
    /* PBF (PBFSolver.hpp:18-29) */
    uint32_t iteration_num;          /* Required.iterationNum                     */
    float xsph_k;                    /* Required.XSPH_k                           */
    float f_rest_density;            /* Required.fPartRestDensity                 */
    float b_rest_density;            /* Required.bPartRestDensity                 */
    /* DFSPH (DFSPHSolver.hpp:18-31) */
    float f_rest_viscosity;          /* Required.fPartRestViscosity               */
    float div_free_threshold;        /* Required.divFreeThreshold                 */
    float incomp_threshold;          /* Required.incompThreshold                  */
    float surface_tension_coefficient; /* Required.surfaceTensionCoefficient      */
    /* IMM / IMM-CT (IMMSolver.hpp:17-32, IMMCTSolver.hpp:17-38) */
    uint32_t phase_num;              /* len(Required.phaseRestDensity)            */
    float phase_rest_density[VTP_MAX_PHASES];
    float phase_rest_viscosity[VTP_MAX_PHASES]; /* IMM: phaseRestViscosity; IMM-CT:
                                        solventViscosity for every phase unless
                                        the extension key phaseRestViscosity is set */
    float phase_rest_color[VTP_MAX_PHASES][3];
    float cf;                        /* Required.diffusionCoefficientCf           */
    float cd;                        /* momentumExchangeCoefficientCd (IMM) / Cd0 */
    float solution_vis_base;         /* IMM-CT solutionBasicViscosity (pre x1000) */
    float solution_vis_max;          /* IMM-CT solutionMaxViscosity (pre x1000)   */
    float relaxation_time;           /* IMM-CT relaxationTime                     */
    float thinning_exp0;             /* IMM-CT shearThinningBasicFactor           */
    float rheology_threshold;        /* IMM-CT rheologicalThreshold               */
    int32_t polymer_phase;           /* extension: polymer phase index (ref: 1)   */
    /* extensions (no reference counterpart) */
    int32_t cfl_enable;              /* adaptive dt: dt = min(dt_max, c*2r/|v|max) */
    float cfl_number;
    float dt_max;
    uint32_t max_solver_iterations;  /* iteration cap, reference: 100             */
    int32_t mass_weighted;           /* 1 = neighbour mass m_j in the PBF constraint
                                        gradient and the IMM-CT polymer-stress term
                                        (the reference omits it); default 0       */
} vtp_params;

/* Readable / writable per-particle fields (attach order). */
typedef enum vtp_field {
    VTP_FIELD_POS = 0,          /* float[n][3]                                   */
    VTP_FIELD_VEL = 1,          /* float[n][3]                                   */
    VTP_FIELD_MAT = 2,          /* int32[n]                                      */
    VTP_FIELD_MASS = 3,         /* float[n]                                      */
    VTP_FIELD_VOLUME = 4,       /* float[n]                                      */
    VTP_FIELD_DENSITY = 5,      /* float[n]  PBF density_sph                     */
    VTP_FIELD_LAMBDA = 6,       /* float[n]  PBF lamb                            */
    VTP_FIELD_ERROR = 7,        /* float[n]  PBF error (C_i)                     */
    VTP_FIELD_REST_DENSITY = 8, /* float[n]                                      */
    VTP_FIELD_COMPRESSION_RATIO = 9,  /* float[n]                                */
    VTP_FIELD_DF_ALPHA = 10,    /* float[n]                                      */
    VTP_FIELD_DELTA_CR = 11,    /* float[n]  delta_compression_ratio             */
    VTP_FIELD_KAPPA_DIV = 12,   /* float[n]                                      */
    VTP_FIELD_KAPPA_INCOMP = 13,/* float[n]                                      */
    VTP_FIELD_VEL_ADV = 14,     /* float[n][3]                                   */
    VTP_FIELD_ACC = 15,         /* float[n][3]                                   */
    VTP_FIELD_SURFACE_NORMAL = 16, /* float[n][3]                                */
    VTP_FIELD_VIS = 17,         /* float[n]                                      */
    VTP_FIELD_VOL_FRAC = 18,    /* float[n][P]                                   */
    VTP_FIELD_VEL_PHASE = 19,   /* float[n][P][3]                                */
    VTP_FIELD_COLOR = 20,       /* float[n][3]                                   */
    VTP_FIELD_CD = 21,          /* float[n]   IMM-CT per-particle Cd             */
    VTP_FIELD_Q = 22,           /* float[n][9] conformation tensor, row-major
                                   (held as Q - I on the device)                 */
    VTP_FIELD_VEL_GRAD = 23,    /* float[n][9]                                   */
    VTP_FIELD_VISE_STRESS = 24, /* float[n][9]                                   */
    VTP_FIELD_VEL_DRIFT = 25,   /* float[n][P][3]                                */
    VTP_FIELD_SCENE = 26,       /* uint32[n] scene index                         */
    VTP_FIELD_COUNT_ = 27
} vtp_field;

typedef struct vtp_step_stats {
    uint64_t steps;                 /* ticks done since initialize               */
    float sim_time;                 /* Data::cur_simTime                         */
    float dt;                       /* last dt used                              */
    uint32_t div_iterations;        /* last step (DFSPH/IMM/IMM-CT)              */
    uint32_t incomp_iterations;     /* last step                                 */
    uint32_t phase_iterations;      /* last step (IMM/IMM-CT negative-check loop)*/
    uint32_t pad_;
    double div_residual;            /* residual at exit of the last div solve    */
    double incomp_residual;         /* residual at exit of the last incomp solve */
    uint64_t total_div_iterations;
    uint64_t total_incomp_iterations;
    uint32_t dead_particles;        /* particles that left the grid              */
    uint32_t crash_flags;           /* VTP_CRASH_* bits; sticky                  */
} vtp_step_stats;

/* ---- configuration -------------------------------------------------------
 * Parse a reference solver config (the whole document: {"PBF": {"Required":
 * {...}, "Optional": {...}}, "EXPORT": {...}}) into vtp_params.  Performs the
 * reference's checkConfig (PBFSolver.cpp:378-405 and siblings): a missing
 * Required/Optional key is VTP_EINVAL with the key named in `err`.
 * Host-only; needs no GPU. */
vtp_status vtp_params_from_json(int32_t solver, const char *json_text,
                                vtp_params *out, char *err, size_t err_len);

/* Default parameters = the solver's config template values. Host-only. */
vtp_status vtp_params_default(int32_t solver, vtp_params *out);

/* The solver's config template as JSON text (JsonHandler::load<Solver>
 * ConfigTemplateJson, JSONHandler.cpp:65-67 and siblings; values of
 * ConfigTemplates/<Solver>/<Solver>ConfigTemplate.json).  Static storage;
 * NULL for an unknown solver.  Host-only. */
const char *vtp_config_template(int32_t solver);
/* The solver's object-component template (Solver::
 * getSolverObjectComponentConfigTemplate, JSONHandler.cpp:69-71 and
 * siblings).  Static storage; NULL for an unknown solver.  Host-only. */
const char *vtp_object_config_template(int32_t solver);

/* ---- lifecycle (Solver.hpp:27-62) ---------------------------------------- */
/* SolverManager::createSolver + Solver::setConfig(json) */
vtp_status vtp_create(vtp_ctx **out, int32_t solver, const char *json_text);
/* Same, from an already-filled parameter block. */
vtp_status vtp_create_with_params(vtp_ctx **out, const vtp_params *params);

/* Solver::attachObject: appends one object of `n` particles of material `mat`
 * (VTP_MAT_*).  `xyz` is float[n][3]; `vel` float[n][3] or NULL (zero);
 * `vol_frac` float[n][P] or NULL (zeros) for IMM/IMM-CT.  `scene` selects an
 * independent scene (batched mode): particles of different scenes never
 * interact.  The object's attach range [start, end) feeds the residual mean
 * exactly like exportObjectStartIndex/EndIndex (DFSPHSolver.cpp:341-346). */
vtp_status vtp_add_particles(vtp_ctx *ctx, uint32_t scene, int32_t mat,
                             const float *xyz, const float *vel,
                             const float *vol_frac, uint32_t n);

/* Solver::initialize: allocate, upload, first neighbour search, init kernels
 * (PBFSolver.cpp:73-125). */
vtp_status vtp_initialize(vtp_ctx *ctx);

/* Solver::tickNsteps(n): n ticks, then one synchronisation; returns
 * VTP_ECRASH once any crash bit is set (sticky, like m_isCrashed). */
vtp_status vtp_step(vtp_ctx *ctx, uint32_t nsteps);

/* Solver::run(): ticks until sim_time >= animationTime. */
vtp_status vtp_run(vtp_ctx *ctx);

/* Solver::destroy */
void vtp_destroy(vtp_ctx *ctx);

/* Message of the last failing call on this context (or of the last failing
 * vtp_create when ctx is NULL). */
const char *vtp_last_error(const vtp_ctx *ctx);

/* ---- data access --------------------------------------------------------- */
uint32_t vtp_particle_count(const vtp_ctx *ctx);
/* Copies a field to host memory in attach order; `bytes` must equal the
 * field's size.  Synchronises the context's stream. */
vtp_status vtp_read(vtp_ctx *ctx, int32_t field, void *host_dst, size_t bytes);
/* Overwrites VTP_FIELD_POS or VTP_FIELD_VEL (attach order) after initialize. */
vtp_status vtp_write(vtp_ctx *ctx, int32_t field, const void *host_src, size_t bytes);
/* vtp_read / vtp_write of VTP_FIELD_POS or VTP_FIELD_VEL on DEVICE memory of
 * the context's GPU (float[n][3], attach order), with no host copy: the same
 * cudaMemcpy of exportData / the state upload, for callers whose arrays are
 * already on the GPU (the torch autograd wrapper).  The caller's writes to
 * d_src must be complete; both return after the context's stream. */
vtp_status vtp_read_device(vtp_ctx *ctx, int32_t field, float *d_dst, size_t bytes);
vtp_status vtp_write_device(vtp_ctx *ctx, int32_t field, const float *d_src, size_t bytes);
vtp_status vtp_stats(vtp_ctx *ctx, vtp_step_stats *out);

/* UGNS state of the last neighbour search, in the reference's own layout
 * (UniformGridNeighborSearch.hpp:28-63): neighbors[p*M + t] in attach ids
 * with entry 0 = p and UINT32_MAX padding, neighborNum[p] (self included),
 * particleIndices[i] (sorted order), cellStart[c]/cellEnd[c] (inclusive,
 * UINT32_MAX = empty).  Any pointer may be NULL.  Sizes: N*M, N, N, C, C. */
vtp_status vtp_debug_neighbors(vtp_ctx *ctx, uint32_t *neighbors,
                               uint32_t *neighbor_num, uint32_t *sorted_ids,
                               uint32_t *cell_start, uint32_t *cell_end);
/* grid[3] = ceil(simSpaceSize / h) per axis (UniformGridNeighborSearch.cpp:94-98);
 * *cells = cells per scene. */
vtp_status vtp_grid_info(const vtp_ctx *ctx, uint32_t grid[3], uint32_t *cells);

/* ---- checkpoint / resume --------------------------------------------------
 * No counterpart in the reference: its exporter writes positions and colour
 * only (ExportUtil.cpp:28-88), so a run cannot be resumed.  A state image is
 * this header followed by the carried state in the solver's current slot
 * order, each array packed (no padding), little-endian:
 *   uint32 orig[n]                 attach id of each slot
 *   float  x[n][4]                 position, w = mass
 *   float  v[n][4]                 velocity, w = volume
 *   IMM / IMM-CT only:
 *   float  vol_frac[P][n]          volume fraction per phase
 *   float  vel_phase[P][n][4]      phase velocity (w unused)
 *   float  rest_density[n]
 *   float  color[n][4]             (w unused)
 *   IMM-CT only:
 *   float  Q_minus_I[9][n]         conformation tensor minus the identity
 *                                  (the device's representation), row-major
 * Loading into a context built with the same solver type and the same
 * attached objects (count, material, scene and size of each) continues the run
 * bit for bit: the ticks after vtp_load_state equal the ticks the saved
 * context would have done. */
typedef struct vtp_state_header {
    char magic[8];              /* "VTPSTATE"                                   */
    uint32_t version;           /* 1                                            */
    int32_t solver;             /* vtp_solver_type                              */
    uint32_t n;                 /* particles                                    */
    uint32_t phases;            /* P (0 for PBF and DFSPH)                      */
    uint32_t objects;           /* attached objects                             */
    uint32_t crash_flags;       /* VTP_CRASH_* bits at save time                */
    uint64_t layout_hash;       /* FNV-1a of solver, n, P and every object's
                                   (start, end, scene, material)               */
    uint64_t steps;
    double sim_time;
    uint64_t total_div_iterations;
    uint64_t total_incomp_iterations;
    uint32_t iter_hint_div, iter_hint_incomp;   /* iterations of the last solves */
    uint32_t nbr_at_cap;        /* particles whose list reached the cap in the
                                   last neighbour build (selects the next
                                   build's walk, which orders the lists)      */
    uint32_t reserved;          /* 0                                            */
    uint64_t payload_bytes;     /* bytes that follow the header                 */
} vtp_state_header;

/* Bytes of the state image of an initialized context (header included). */
vtp_status vtp_state_size(vtp_ctx *ctx, size_t *bytes);
/* Writes the state image to `host_dst`; `bytes` must be at least
 * vtp_state_size.  Synchronises the context's stream. */
vtp_status vtp_save_state(vtp_ctx *ctx, void *host_dst, size_t bytes);
/* Replaces the carried state of an initialized context by a state image.
 * VTP_EINVAL (nothing changed) when the magic, version, solver type, layout
 * hash or size do not match.  Neighbour lists are rebuilt by the next tick. */
vtp_status vtp_load_state(vtp_ctx *ctx, const void *host_src, size_t bytes);

/* ---- execution options ---------------------------------------------------
 * Keys: "graph" (1 = capture the PBF substep in a hipGraph and replay it),
 * "autotune" (time launch shapes of the neighbour passes of the context's
 * solver, at initialize or immediately when already initialized, and keep
 * the fastest; each candidate runs from a state image that is loaded back
 * afterwards, so the simulation is not advanced.  1 = the workgroup size of
 * each pass (64, 128 or 256; results bit-identical); 2 = also 1, 2 or 4
 * lanes per particle ("lanes"), which changes the FP32 rounding of
 * neighbour sums.  The chosen shapes read back through vtp_get_option.  As
 * after vtp_load_state, a recorded PBF rollout is dropped, the lists count
 * as stale until the next tick, and recorded stage times are cleared),
 * "block" (threads per block for the neighbour passes: 64, 128, 192 or 256;
 * 0 = default 256),
 * "stage_timing" (1 = record hipEvents around every stage; see
 * vtp_stage_times), "solver_chunk" (iterations launched per host check),
 * "lanes" (1, 2 or 4 lanes per particle in the PBF neighbour passes and in
 * the DFSPH / IMM / IMM-CT pressure-solve passes; default 1),
 * "neighbor_walk" (0 = adaptive, default: the list build walks cell rows and
 * switches to the reference's cell order once many lists reach the cap;
 * 1 = always rows; 2 = always the reference's order.  The neighbour sets are
 * the reference's either way; the order inside a list, and so the FP32
 * rounding of neighbour sums, depends on the walk),
 * "static_lists" (PBF and DFSPH: 0 = default, a tick's search builds the
 * lists of fluid particles only, since no PBF or DFSPH kernel reads a static
 * particle's list after initialize; 1 = every list, as the reference, which
 * vtp_debug_neighbors needs after a tick.  IMM and IMM-CT always build every
 * list: their boundary rest density is a maximum over each wall particle's
 * neighbours every tick),
 * "radix_sort" (0 = default, the search groups particles by cell with a
 * counting sort; 1 = a stable LSD radix sort of the cell keys.  Both give
 * the same (cellId, attach id) order and bit-identical results),
 * "full_flux_retry" (IMM / IMM-CT: 0 = default, a negative-fraction retry
 * recomputes the fluxes of the particles next to newly flagged ones; 1 =
 * every retry recomputes all fluxes, as the reference.  Bit-identical),
 * "build_lambda" (PBF with "lanes" 1: 1 = the neighbour build also
 * computes the first constraint iteration's density and lambda from the
 * neighbours it lists; 0 = default, a separate pass does.  Bit-identical). */
vtp_status vtp_set_option(vtp_ctx *ctx, const char *key, double value);
/* Reads an option back: the keys above except "stage_timing" and
 * "solver_chunk", and "block.<pass>" for the workgroup size a neighbour pass
 * launches with (pbf_lambda, pbf_dx, pbf_xsph, df_precompute, df_delta_kappa,
 * df_update_vadv, df_vis_normal, imm_precompute, imm_vis_normal,
 * imm_tension_phase_vel, imm_fluxes, immct_conformation, immct_stress_acc). */
vtp_status vtp_get_option(vtp_ctx *ctx, const char *key, double *value);

/* Per-stage device time (ms, summed over launches) and launch counts,
 * recorded with hipEvents on the context's stream while "stage_timing" is 1.
 * names receives a '\n'-separated list.  Returns the number of stages. */
int32_t vtp_stage_times(vtp_ctx *ctx, char *names, size_t names_len,
                        double *ms, uint64_t *launches, int32_t max_stages);
vtp_status vtp_reset_stage_times(vtp_ctx *ctx);

/* ---- differentiable PBF rollout (no reference counterpart) ---------------
 * Forward: runs nsteps PBF ticks from the current state, storing the state
 * every ckpt_every steps (recomputation checkpoints).  The loss used by the
 * backward pass is supplied as its gradient with respect to the final
 * positions and velocities (float[n][3] each, attach order, may be NULL).
 * Backward: returns dL/dv0 (float[n][3]), dL/dx0 (float[n][3], may be NULL)
 * and dL/dXSPH_k (one float). */
vtp_status vtp_pbf_rollout_fwd(vtp_ctx *ctx, uint32_t nsteps, uint32_t ckpt_every);
vtp_status vtp_pbf_rollout_bwd(vtp_ctx *ctx, const float *grad_x, const float *grad_v,
                               float *grad_v0, float *grad_x0, float *grad_xsph_k);
/* vtp_pbf_rollout_bwd with every array (and dL/dXSPH_k, one float) in DEVICE
 * memory of the context's GPU; no host round trip.  Returns after the
 * context's stream. */
vtp_status vtp_pbf_rollout_bwd_device(vtp_ctx *ctx, const float *d_grad_x, const float *d_grad_v,
                                      float *d_grad_v0, float *d_grad_x0, float *d_grad_xsph_k);
/* Mean squared density deviation of the fluid at the current positions over
 * the last neighbour lists, L = mean_f (rho_f/rho0 - 1)^2 (unclamped SPH
 * density), and dL/dx (float[n][3], attach order; may be NULL). */
vtp_status vtp_pbf_density_loss(vtp_ctx *ctx, double *loss, float *grad_x);
/* vtp_pbf_density_loss with the loss (one double) and dL/dx (may be NULL) in
 * DEVICE memory of the context's GPU; its scratch is allocated per call. */
vtp_status vtp_pbf_density_loss_device(vtp_ctx *ctx, double *d_loss, float *d_grad_x);
/* vtp_pbf_density_loss without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under hipGraph
 * capture of their own, tight loops, harnesses that forbid hipMalloc).
 * Outputs are DEVICE pointers on the context's GPU: d_loss one double,
 * d_grad_x float[n][3] in attach order (may be NULL).  Its scratch (the loss
 * accumulator and dL/drho per particle) is a fixed array of 16 MiB + 256 B on
 * each GPU the library runs on, held until the process ends, so contexts of
 * at most 4,194,304 particles are accepted (VTP_EINVAL above that).  Every
 * call on one GPU shares that array: calls that overlap on one GPU (from
 * several host threads or contexts) race.  Returns after the context's stream
 * has finished the work. */
vtp_status vtp_pbf_density_loss_noalloc(vtp_ctx *ctx, double *d_loss, float *d_grad_x);
/* Sets XSPH_k (the learnable coefficient) on an existing context. */
vtp_status vtp_pbf_set_xsph(vtp_ctx *ctx, float xsph_k);

/* Library version string. */
const char *vtp_version(void);

#ifdef __cplusplus
}
#endif
#endif /* VTP_H */
