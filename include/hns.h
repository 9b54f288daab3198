/*
 * hns.h — C ABI of the MI355X (gfx950) sparse smoke/fire step.
 *
 * Drop-in boundary for the launch functions HNanoSolver's SOP nodes call. Every entry point
 * names the reference interface it replaces (paths relative to the reference repository).
 * The reference exports `extern "C"` functions that take C++ references
 * (HNS::GridIndexedData&, nanovdb::GridHandle&, std::vector&) and throw std::runtime_error;
 * here every argument is a plain pointer or size and every call returns an hns_status.
 * include/hns.hpp restores the reference's C++ signatures and exceptions on top of this.
 *
 * Data model (src/Utils/GridData.hpp:16-166, src/Utils/GridBuilder.hpp:164-174):
 *   - a payload of N voxels, N = 512 x leaves ("leaf-dense"): coords[3*i .. 3*i+2] is the
 *     (x, y, z) index-space coordinate of voxel i; each leaf's 512 voxels are consecutive in
 *     offset order (x<<6 | y<<3 | z, local to the 8^3 leaf) and leaves come in the order
 *     NanoVDB's voxelsToGrid sorts them (the order OpenVDB's LeafManager gives);
 *   - named value blocks, each N floats (HNS_FLOAT) or N x 3 floats AoS (HNS_VEC3F).
 * Device state lives in HBM for the lifetime of a context; nothing is re-uploaded per step.
 *
 * Threading: one host thread per context at a time. Every call that launches work ends with
 * a stream synchronisation, as the reference's do (src/Cuda/HNanoSolver.cu:135, :220, :455, :494).
 */
#ifndef HNS_H_
#define HNS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HNS_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------------------- */
typedef int32_t hns_status;
#define HNS_OK 0
#define HNS_ERR_INVALID_ARGUMENT (-1) /* bad size, pointer or parameter */
#define HNS_ERR_EMPTY (-2)            /* N == 0 (HNanoSolver.cu:61) */
#define HNS_ERR_NULL_CONTEXT (-3)     /* null context handle */
#define HNS_ERR_TOPOLOGY (-4)         /* coords are not leaf-dense or do not match the topology */
#define HNS_ERR_LEAF_ORDER (-5)       /* leaves are not in voxelsToGrid order (see hns_leaf_order) */
#define HNS_ERR_NO_VELOCITY (-6)      /* payload has no Vec3f block (HNanoSolver.cu:101) */
#define HNS_ERR_HIP (-7)              /* a HIP runtime call failed (CUDA_CHECK, Utils.cuh:10-18) */
#define HNS_ERR_OUT_OF_MEMORY (-8)    /* device allocation failed (Utils.cuh:287-296) */
#define HNS_ERR_NOT_FOUND (-9)        /* a named block is missing */
#define HNS_ERR_NO_DEVICE (-10)       /* no HIP device is visible */
#define HNS_ERR_NOT_CONVERGED (-11)   /* a to-tolerance solve missed tol: it hit max_iters or its
                                         update broke down (results still written) */
#define HNS_ERR_NON_FINITE (-12)      /* a to-tolerance solve met a residual or right-hand side that
                                         is not finite (results still written) */

/* Message of the last failing call on this thread (never null). */
const char* hns_last_error(void);
/* HNS_ABI_VERSION of the loaded library. */
int32_t hns_abi_version(void);

/* ---- payload (HNS::GridIndexedData as plain data) ----------------------------------- */
#define HNS_FLOAT 1
#define HNS_VEC3F 3
typedef struct hns_value_block {
  const char* name; /* e.g. "density", "velocity", "src_density", "collision_sdf" */
  int32_t type;     /* HNS_FLOAT or HNS_VEC3F */
  void* values;     /* N or 3N floats, caller-owned */
} hns_value_block;

typedef struct hns_grid_data {
  size_t size;              /* N voxels (GridIndexedData::size, GridData.hpp:111) */
  const int32_t* coords;    /* 3N int32, (x,y,z) per voxel (pCoords, GridData.hpp:91) */
  size_t num_blocks;
  hns_value_block* blocks;  /* insertion order, as GridData.hpp:73 keeps them */
} hns_grid_data;

/* CombustionParams, src/Cuda/Kernels.cuh:5-24 (and the identical copy in
 * src/SOP/HNanoSolver/SOP_HNanoSolver.hpp:23-43): 72 bytes, same field order. */
typedef struct hns_params {
  float expansionRate;
  float temperatureRelease;
  float buoyancyStrength;
  float ambientTemp;
  float vorticityScale;
  float factorScale;
  float disturbanceEnable;
  float disturbanceStrength;
  float disturbanceSwirl;
  float disturbanceThreshold;
  float disturbanceGain;
  float disturbanceFrequency;
  int32_t substeps;
  float gravity;
  float maskDensityMin;
  float maskDensityMax;
  int32_t downloadVel;
  int32_t simReset;
} hns_params;

/* ---- host ingestion (the SOP's domain and IndexGridBuilder, SOP_HNanoSolver.cpp:276-301,
 *      src/Utils/GridBuilder.hpp:93-175, :231-251) ------------------------------------------
 * A source grid as its leaves, the form OpenVDB and NanoVDB hold them in. Tiles are not
 * represented (the reference's domain ignores them: dilateVoxels(..., IGNORE_TILES) and a
 * LeafManager over leaves). */
#define HNS_ROLE_VALUE 0 /* a value grid (feedback or source field) */
#define HNS_ROLE_SDF 1   /* the collision SDF: its leaves join the domain, missing leaves read 0x01 bytes */
typedef struct hns_leaf_grid {
  const char* name;
  int32_t type;            /* HNS_FLOAT or HNS_VEC3F */
  int32_t role;            /* HNS_ROLE_VALUE or HNS_ROLE_SDF */
  size_t num_leaves;
  const int32_t* origins;  /* 3 int32 per leaf, multiples of 8, no leaf repeated */
  const float* values;     /* per leaf 512 floats (HNS_FLOAT) or 512 x (x,y,z) (HNS_VEC3F), offset order
                              x<<6 | y<<3 | z: the leaf buffer */
  const uint64_t* active;  /* per leaf 8 words, bit (o & 63) of word (o >> 6) = voxel o active; NULL = all */
} hns_leaf_grid;
typedef struct hns_payload hns_payload;
/* The leaf-dense payload of one cook. Domain: the active voxels of grids[domain_grid] (the primary
 * velocity) dilated by `padding` voxels over the 26-neighbourhood, united with the active voxels of
 * the first HNS_ROLE_SDF grid; every voxel of every domain leaf, leaves in voxelsToGrid order.
 * Blocks: one per grid name in the grids' order (a repeated name keeps the first grid, as
 * GridIndexedData::addValueBlock does); a grid's leaf buffer is copied where it has a leaf at the
 * domain leaf's origin, else the leaf is 0 (memset(..., 0, ...)), or for the SDF the byte pattern of
 * memset(..., 1, ...) (every float 0x01010101, about 2.37e-38; GridBuilder.hpp:116). The payload is in
 * pinned host memory when a HIP device is visible. */
hns_status hns_payload_build(const hns_leaf_grid* grids, size_t num_grids, size_t domain_grid, int32_t padding,
                             hns_payload** out);
/* The payload as an hns_grid_data (pointers into the payload, valid until it is destroyed);
 * *pinned (nullable) = 1 when it lives in pinned host memory. */
hns_status hns_payload_view(hns_payload* payload, hns_grid_data* view, int32_t* pinned);
/* The domain's active voxels of each payload leaf, the topology IndexGridBuilder::writeIndexGrid
 * gives its grids (topologyUnion with the domain grid, GridBuilder.hpp:201-206): the dilated domain
 * grid's active voxels united with the SDF's, 8 words per leaf in payload order (bit (o & 63) of
 * word (o >> 6) = voxel o). num_leaves must be the payload's leaf count (view.size / 512). */
hns_status hns_payload_domain_mask(hns_payload* payload, uint64_t* active, size_t num_leaves);
hns_status hns_payload_destroy(hns_payload** payload);
/* The output scatter of SOP_HNanoSolver.cpp:345-388 (scatterScalar / scatterVector): the field
 * `name` of a leaf-dense payload as a grid whose voxels with mask[i] != 0 are set on with their
 * value (every other voxel at `background`, inactive), returned as leaves: each payload leaf with a
 * masked voxel, in payload order, gives its origin (3 int32), 512 values (x3 for a Vec3f block)
 * and 8 active-mask words. mask NULL = all voxels (HNS_GetActiveMask's mask is identically 1).
 * *num_leaves receives the output leaf count; the arrays (nullable, to query the count) hold
 * `capacity` leaves. */
hns_status hns_scatter_leaves(const hns_grid_data* data, const char* name, const uint8_t* mask, float background,
                              int32_t* origins, float* values, uint64_t* active, size_t capacity,
                              size_t* num_leaves);
/* IndexGridBuilder::writeIndexGrid (GridBuilder.hpp:177-226): the field `name` of a leaf-dense
 * payload as a grid of every payload leaf (payload order), each leaf's buffer the payload's 512
 * values (x3 for a Vec3f block; memcpy of the leaf, active or not) and its active voxels
 * domain_active's (8 words per leaf, as hns_payload_domain_mask gives them; NULL = every voxel).
 * The arrays (nullable) hold view.size / 512 leaves: 3 int32, 512 (x3) floats and 8 words each. */
hns_status hns_write_leaf_grid(const hns_grid_data* data, const char* name, const uint64_t* domain_active,
                               int32_t* origins, float* values, uint64_t* active);

/* ---- topology (replaces nanovdb::GridHandle<nanovdb::cuda::DeviceBuffer>) ------------ */
typedef struct hns_topology hns_topology;

/* CreateIndexGrid(GridIndexedData&, GridHandle<cuda::DeviceBuffer>&, float voxelSize),
 * src/Cuda/HNanoSolver.cu:498-510. Builds the device-resident leaf topology (leaf origins,
 * 27-entry neighbour table per leaf, leaf hash) from data->coords. Validates that the payload
 * is leaf-dense and in voxelsToGrid order (HNS_ERR_TOPOLOGY / HNS_ERR_LEAF_ORDER otherwise).
 * The reference rebuilds its grid on every cook (SOP_HNanoSolver.cpp:310-317); here an existing
 * *out built from the same leaf origins, voxel size and device is kept as it is (the payload's
 * leaf origins are compared on the host), and any other *out is released and replaced. */
hns_status hns_create_index_grid(const hns_grid_data* data, float voxel_size, hns_topology** out);
/* Number of device builds behind a handle (1 + the rebuilds avoided = calls that kept it). */
hns_status hns_topology_builds(const hns_topology* topo, int64_t* builds, int64_t* reuses);
/* Releases a topology handle (contexts built from it keep their own reference). */
hns_status hns_topology_destroy(hns_topology** topo);
/* Number of voxels / leaves described by a handle. */
hns_status hns_topology_info(const hns_topology* topo, size_t* n_voxels, size_t* n_leaves);
/* Copies the device neighbour table (27 int32 per leaf, -1 = absent; entry (dx+1)*9+(dy+1)*3+(dz+1)). */
hns_status hns_topology_neighbours(const hns_topology* topo, int32_t* out, size_t capacity);

/* Host helper: permutation of L leaf origins (3L int32) into voxelsToGrid order
 * (nanovdb/tools/cuda/PointsToGrid.cuh:596-651): perm[r] = input index of the r-th leaf. */
hns_status hns_leaf_order(const int32_t* origins, size_t n_leaves, uint64_t* perm);

/* ---- simulation context (SimContext, src/Cuda/HNanoSolver.cu:27-48) ------------------ */
typedef struct hns_context hns_context;

/* HNS_EnsureContext(void** ctx, GridIndexedData&, const GridHandle&, float voxelSize,
 * bool hasCollision), src/Cuda/HNanoSolver.cu:58-155. Creates the context when *ctx is null or
 * when N or the voxel size changed (uploading velocity, every float block except
 * "collision_sdf" and "src_*", and the collision SDF); otherwise keeps the device state and only
 * refreshes the topology and the collision SDF. */
hns_status hns_ensure_context(hns_context** ctx, const hns_grid_data* data, const hns_topology* topo,
                              float voxel_size, int32_t has_collision);

/* HNS_UploadSources(void*, GridIndexedData&), HNanoSolver.cu:158-221: replaces the per-cook
 * sources ("src_vel" Vec3 or "src_vel_x/y/z", and every "src_*" float block). */
hns_status hns_upload_sources(hns_context* ctx, const hns_grid_data* data);

/* HNS_Advance(void*, int iteration, float dt, const CombustionParams&), HNanoSolver.cu:223-456.
 * `iteration` red-black Gauss-Seidel iterations per pressure solve, as the reference. */
hns_status hns_advance(hns_context* ctx, int32_t iteration, float dt, const hns_params* params);

/* Pressure-solver selection for hns_advance_ex (no reference counterpart: the reference always
 * runs a fixed iteration count, HNanoSolver.cu:366-373). */
#define HNS_SOLVER_RBGS_FIXED 0  /* the reference: `iteration` RBGS iterations */
#define HNS_SOLVER_RBGS_TOL 1    /* RBGS until ||div - L p|| / ||div|| <= tol (or max_iters) */
#define HNS_SOLVER_CG 2          /* conjugate gradients on the same operator, to tol */
#define HNS_SOLVER_PCG_LEARNED 3 /* flexible PCG with the learned preconditioner, to tol */
#define HNS_SOLVER_MGCG 4        /* CG preconditioned by a geometric multigrid V-cycle, to tol */
typedef struct hns_solve_options {
  uint32_t struct_size;  /* sizeof(hns_solve_options) */
  int32_t solver;        /* HNS_SOLVER_* */
  float tol;             /* relative residual target */
  int32_t max_iters;     /* cap for the to-tolerance solvers */
  int32_t check_every;   /* residual check period for HNS_SOLVER_RBGS_TOL (0 = default) */
} hns_solve_options;
typedef struct hns_solve_stats {
  uint32_t struct_size;          /* sizeof(hns_solve_stats) */
  int32_t iterations;            /* iterations of the last pressure solve */
  double rel_residual;           /* ||div - L p||_2 / ||div||_2 after the last solve (FP64 sums) */
  double solve_ms;               /* device time of the last solve */
  int32_t fallback_iterations;   /* RBGS iterations run after a learned PCG missed tol (0 if none) */
  int32_t reserved;
} hns_solve_stats;
hns_status hns_advance_ex(hns_context* ctx, int32_t iteration, float dt, const hns_params* params,
                          const hns_solve_options* opts, hns_solve_stats* stats);

/* HNS_DownloadForOutput(void*, GridIndexedData&), HNanoSolver.cu:471-495: writes velocity into
 * the "velocity" block (else the first Vec3f block) and every state field into the float block of
 * the same name, where present. */
hns_status hns_download_for_output(hns_context* ctx, hns_grid_data* data);

/* HNS_GetActiveMask(void*, std::vector<unsigned char>&), HNanoSolver.cu:459-468. The
 * reference's mask is identically 1 (Kernel.cu:74-75). Writes min(N, capacity) bytes; *n_out = N
 * (0 for a null context, which the reference answers with an empty mask). */
hns_status hns_get_active_mask(hns_context* ctx, uint8_t* mask, size_t capacity, size_t* n_out);

/* HNS_DestroyContext(void** ctx), HNanoSolver.cu:51-56. Null-safe; sets *ctx to null. */
hns_status hns_destroy_context(hns_context** ctx);

/* Runs the context's work on a caller stream (a hipStream_t; null = the context's own stream). */
hns_status hns_set_stream(hns_context* ctx, void* stream);
/* Context sizes and whether a named state field exists. */
hns_status hns_context_info(const hns_context* ctx, size_t* n_voxels, size_t* n_leaves, int32_t* n_state);
int32_t hns_has_field(const hns_context* ctx, const char* name);

/* Per-kernel device timing with HIP events (for measurement; off by default). */
#define HNS_K_ADVECT_VELOCITY 0
#define HNS_K_VORTICITY_FORCES 1
#define HNS_K_DIVERGENCE_COMBUSTION 2
#define HNS_K_RBGS 3
#define HNS_K_RESIDUAL 4
#define HNS_K_GRADIENT 5
#define HNS_K_ADVECT_SCALARS 6
#define HNS_K_CG_PRODUCT 7 /* CG/PCG: direction update + operator product (+ d.q partials) */
#define HNS_K_CG_UPDATE 8  /* CG/PCG: x, r update (+ r.r partials) */
#define HNS_K_PRECOND 9    /* PCG: the learned preconditioner (+ r.z partials) */
#define HNS_K_OTHER 10
#define HNS_K_COUNT 11
hns_status hns_set_timing(hns_context* ctx, int32_t enable);
/* Accumulated time (ms) and launch count for kernel class `which` since the last reset. */
hns_status hns_get_timing(hns_context* ctx, int32_t which, double* total_ms, int64_t* launches);
hns_status hns_reset_timing(hns_context* ctx);

/* ---- standalone entry points (the HNanoAdvect / HNanoAdvectVelocity /
 *      HNanoProjectNonDivergent SOP nodes); each builds its topology from data->coords,
 *      updates the payload in place and frees its device memory before returning ---------- */

/* AdvectIndexGrid(GridIndexedData&, float dt, float voxelSize, const cudaStream_t&),
 * src/Cuda/Advection.cu:13-112, :169-171: every float block is advected by the single Vec3f
 * block with advect_scalar (Kernel.cu:294-344; misses read 0). */
hns_status hns_advect_index_grid(hns_grid_data* data, float dt, float voxel_size, void* stream);
/* AdvectIndexGridVelocity(...), Advection.cu:114-166, :173-175: the Vec3f block advects itself
 * (advect_vector, Kernel.cu:349-426). */
hns_status hns_advect_index_grid_velocity(hns_grid_data* data, float dt, float voxel_size, void* stream);
/* Divergence(GridIndexedData&, float voxelSize, const cudaStream_t&), PressureProjection.cu:81-129:
 * writes the divergence of the Vec3f block into the "divergence" float block. */
hns_status hns_divergence(hns_grid_data* data, float voxel_size, void* stream);
/* ProjectNonDivergent(GridIndexedData&, size_t iterations, float voxelSize, const cudaStream_t&),
 * PressureProjection.cu:9-78, :132-135: divergence, `iterations` RBGS iterations from p = 0 with
 * omega = 2/(1+sin(3.14159*voxelSize)), gradient subtraction, in place on the Vec3f block. */
hns_status hns_project_non_divergent(hns_grid_data* data, size_t iterations, float voxel_size, void* stream);

/* ---- device-array kernel API ----------------------------------------------------------
 * Each call launches one kernel of the step on caller-owned device arrays (pointers from
 * hipMalloc or a framework allocator) and synchronises `stream` unless async != 0.
 * Layouts: scalars leaf-dense (length N); velocity SoA (x plane, y plane, z plane; 3N);
 * pressure and divergence for the RBGS sweep "colour-split" (see hnsk_to_colour_split).
 * Missing neighbours read 0 except in hnsk_advect_scalars, which reads element 0 as
 * the reference's advect_scalars does (Kernel.cu:187-188, :229-230, :256-257). */

/* advect_vector, Kernel.cu:349-426. sdf may be null. */
hns_status hnsk_advect_velocity(const hns_topology* topo, const float* vel, float* out, float dt, float inv_dx,
                                const float* sdf, void* stream, int32_t async);
/* advect_scalars, Kernel.cu:163-289: m fields share one back-trace with the velocity `vel`. */
hns_status hnsk_advect_scalars(const hns_topology* topo, const float* vel, int32_t m, const float* const* in,
                               float* const* out, float dt, float inv_dx, const float* sdf, void* stream,
                               int32_t async);
/* advect_scalar, Kernel.cu:294-344 (misses read 0). */
hns_status hnsk_advect_scalar(const hns_topology* topo, const float* vel, const float* in, float* out, float dt,
                              float inv_dx, const float* sdf, void* stream, int32_t async);
/* vorticityConfinement (Kernel.cu:682-727, out of place) followed by the velocity source
 * (vel += src*dt, Kernel.cu:739-755; src_vel SoA or null) and gravity (Kernel.cu:834-838). */
hns_status hnsk_vorticity_forces(const hns_topology* topo, const float* vel, float* out, float dt, float inv_dx,
                                 float vorticity_scale, float factor_scale, const float* src_vel, float gravity,
                                 void* stream, int32_t async);
/* divergence (Kernel.cu:432-453); colour_split != 0 writes the colour-split layout. */
hns_status hnsk_divergence(const hns_topology* topo, const float* vel, float* div, float inv_dx,
                           int32_t colour_split, void* stream, int32_t async);
/* One red-black Gauss-Seidel colour sweep (Kernel.cu:459-489) on colour-split p and div. */
hns_status hnsk_rbgs_sweep(const hns_topology* topo, const float* div_cs, float* p_cs, float dx, int32_t colour,
                           float omega, void* stream, int32_t async);
/* ||div - L p||^2 and ||div||^2 in FP64 (colour-split p and div); synchronous. */
hns_status hnsk_residual(const hns_topology* topo, const float* p_cs, const float* div_cs, float dx,
                         double* r2, double* d2, void* stream);
/* hnsk_residual without allocating. No counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Writes ||div - L p||^2 and ||div||^2 (the same bits as hnsk_residual) to out_dev[0]
 * and out_dev[1], a caller-owned device buffer, and returns without synchronising `stream`. The
 * per-leaf partial sums live in a static device array of the library (2 MiB per device, kept until
 * the process ends; its address is looked up when a topology is built on the device, where the
 * runtime may back it), so topologies above HNS_NOALLOC_MAX_LEAVES leaves are rejected with
 * HNS_ERR_INVALID_ARGUMENT. Every call on a device shares that array: calls that overlap on one
 * device (different streams, or host threads) race and must be ordered by the caller. */
#define HNS_NOALLOC_MAX_LEAVES 131072
hns_status hnsk_residual_noalloc(const hns_topology* topo, const float* p_cs, const float* div_cs, float dx,
                                 double* out_dev, void* stream);
/* A whole pressure solve from p = 0 with one of the HNS_SOLVER_* methods on colour-split div and
 * p (RBGS uses `iterations` and `omega`; the to-tolerance solvers use opts). Synchronous. */
hns_status hnsk_pressure_solve(const hns_topology* topo, const float* div_cs, float* p_cs, float dx, float omega,
                               int32_t iterations, const hns_solve_options* opts, hns_solve_stats* stats,
                               void* stream);
/* subtractPressureGradient, Kernel.cu:495-541 (colour-split p; sdf may be null). */
hns_status hnsk_subtract_gradient(const hns_topology* topo, const float* vel, const float* p_cs, float* out,
                                  float inv_dx, const float* sdf, void* stream, int32_t async);
/* enforceCollisionBoundaries, Kernel.cu:132-158, in place on SoA velocity. */
hns_status hnsk_enforce_collision(const hns_topology* topo, float* vel, const float* sdf, float voxel_size,
                                  void* stream, int32_t async);
/* combustion_oxygen (Kernel.cu:641-677, in place, div read-modify-write in leaf-dense layout)
 * followed by temperature_buoyancy (Kernel.cu:546-563) from vel to vel_out. */
hns_status hnsk_combustion_buoyancy(const hns_topology* topo, float* fuel, float* waste, float* temperature,
                                    float* flame, float* div, const float* vel, float* vel_out, float dt,
                                    float temp_gain, float expansion, float ambient, float buoyancy,
                                    void* stream, int32_t async);
/* Layout converters: leaf-dense <-> colour-split (colour c = (x+y+z)&1 first for c = 0; inside a
 * colour, leaf l, local (x,y,z) sits at c*N/2 + 256*l + 32*x + 4*y + (z>>1)). */
hns_status hnsk_to_colour_split(const hns_topology* topo, const float* in, float* out, void* stream, int32_t async);
hns_status hnsk_from_colour_split(const hns_topology* topo, const float* in, float* out, void* stream,
                                  int32_t async);
/* AoS (x,y,z per voxel) <-> SoA velocity. */
hns_status hnsk_aos_to_soa(size_t n, const float* in, float* out, void* stream, int32_t async);
hns_status hnsk_soa_to_aos(size_t n, const float* in, float* out, void* stream, int32_t async);

/* ---- learned pressure preconditioner (no reference counterpart: the reference projects with a
 *      fixed red-black Gauss-Seidel count, HNanoSolver.cu:364-373) ------------------------------
 * A sparse 3-level 3x3x3 convolutional network over the active leaves (16/32/64 channels, BF16
 * weights and activations, FP32 accumulation) maps a residual r to a correction z; it is the
 * preconditioner of the flexible PCG behind HNS_SOLVER_PCG_LEARNED. Weights are one FP32 vector,
 * layer after layer (c0, c1, c2, c3, u2, c5, u1, c6, out), each W[tap][cin][cout]. A handle keeps
 * the bf16 MFMA operands and activation scratch; calls sharing a handle must not overlap. */
typedef struct hns_precond hns_precond;
int64_t hns_precond_num_params(void);
/* From host weights (copied to the device). */
hns_status hns_precond_create(const float* params, int64_t n, hns_precond** out);
/* Over caller-owned device weights (training); call hns_precond_refresh after changing them. */
hns_status hns_precond_create_device(float* params_dev, int64_t n, hns_precond** out);
hns_status hns_precond_refresh(hns_precond* p, void* stream);
hns_status hns_precond_destroy(hns_precond** p);
/* The context's HNS_SOLVER_PCG_LEARNED solves use `p` (not owned; null detaches). */
hns_status hns_set_preconditioner(hns_context* ctx, hns_precond* p);
/* z = M(r) on colour-split vectors; keep != 0 keeps the activations for hnsk_precond_backward. */
hns_status hnsk_precond_apply(const hns_topology* topo, hns_precond* p, const float* r_cs, float* z_cs, int32_t keep,
                              void* stream);
/* grad (device, hns_precond_num_params() floats) += d/dparams sum(dz . z) of the last kept apply. */
hns_status hnsk_precond_backward(const hns_topology* topo, hns_precond* p, const float* r_cs, const float* dz_cs,
                                 float* grad, void* stream);
/* One Adam step (bias-corrected, step >= 1) on n device parameters. */
hns_status hnsk_precond_adam(int64_t n, float* params, const float* grad, float* m, float* v, float lr, float beta1,
                             float beta2, float eps, int32_t step, void* stream);
/* out = a*y + b*(A x) on colour-split vectors, A = 6 I - (6 face neighbours, 0 outside the domain)
 * (the pressure operator scaled by -dx^2); work: device buffer of HNS_POISSON_WORK_DOUBLES
 * doubles, which receives ||out||^2, ||y||^2 and out.x in work[0..2] (FP64). */
#define HNS_POISSON_WORK_DOUBLES 6148
hns_status hnsk_poisson_apply(const hns_topology* topo, const float* x_cs, const float* y_cs, float a, float b,
                              float* out_cs, double* work, void* stream);
/* z = V(r): one geometric multigrid V-cycle of the HNS_SOLVER_MGCG preconditioner on colour-split
 * vectors (approximately A^-1 r, A as above). flags: HNS_MG_GRAPH replays the V-cycle captured in
 * a HIP graph (else its kernels are launched one by one); HNS_MG_COARSE_LAUNCHES runs the coarsest
 * level as one launch per colour sweep instead of one workgroup; HNS_MG_UNFUSED runs level 0 as
 * separate sweep, restriction and prolongation launches instead of two fused tile passes. All
 * combinations compute the same bits. No reference counterpart (the reference runs a fixed RBGS count, HNanoSolver.cu:364-373).
 * Calls sharing a topology must not overlap. */
#define HNS_MG_GRAPH 1
#define HNS_MG_COARSE_LAUNCHES 2
#define HNS_MG_UNFUSED 4
hns_status hnsk_mg_apply(const hns_topology* topo, const float* r_cs, float* z_cs, int32_t flags, void* stream);
/* ---- reverse mode of the differentiable step (no reference counterpart; the reference is
 *      forward only). Step: u1 = advect_vector(u) (Kernel.cu:349-426), div = divergence(u1),
 *      p = K RBGS iterations from 0 (HNanoSolver.cu:364-373), u2 = u1 - grad p, rho' =
 *      advect_scalars(rho; u2) (Kernel.cu:163-289), u' = u2. Velocity SoA (3N), density N, device
 *      memory. The adjoints are the exact transposes of those kernels. -------------------------- */
typedef struct hns_diff_tape hns_diff_tape;
/* Runs `steps` steps from (vel0, rho0), keeping the state every `checkpoint_every` steps in a new
 * tape; vel_out/rho_out (nullable) receive the final state. */
hns_status hnsk_diff_forward(const hns_topology* topo, const float* vel0, const float* rho0, int32_t steps, float dt,
                             float dx, int32_t iterations, float omega, int32_t checkpoint_every, float* vel_out,
                             float* rho_out, hns_diff_tape** tape, void* stream);
/* Gradients of sum(gvel_out . u_S + grho_out . rho_S) w.r.t. vel0 and rho0 (gvel_out/grho_out
 * nullable = 0); recomputes each segment from its checkpoint. Synchronous. */
hns_status hnsk_diff_backward(hns_diff_tape* tape, const float* gvel_out, const float* grho_out, float* gvel0,
                              float* grho0, void* stream);
/* Checkpoints held, device bytes held by the tape, and elements where a recomputed segment end
 * differed from the stored state in the last backward pass (0 expected). */
hns_status hns_diff_tape_info(const hns_diff_tape* tape, int32_t* checkpoints, int64_t* bytes,
                              int64_t* recompute_mismatches);
hns_status hns_diff_tape_destroy(hns_diff_tape** tape);
/* Single-kernel transposes (accumulate into gin/gvel/gdiv; each allocates its scratch):
 * advect_scalars for one field, advect_vector, and K RBGS iterations from 0 (g_cs overwritten). */
hns_status hnsk_advect_scalars_adj(const hns_topology* topo, const float* vel, const float* in, const float* gout,
                                   float* gin, float* gvel, float dt, float inv_dx, void* stream);
hns_status hnsk_advect_vector_adj(const hns_topology* topo, const float* vel, const float* gout, float* gvel, float dt,
                                  float inv_dx, void* stream);
hns_status hnsk_rbgs_adj(const hns_topology* topo, float* g_cs, float* gdiv_cs, float dx, int32_t iterations,
                         float omega, void* stream);

/* hnsk_pressure_solve with the learned preconditioner (flexible PCG to opts->tol). Synchronous. */
hns_status hnsk_pressure_solve_learned(const hns_topology* topo, hns_precond* p, const float* div_cs, float* p_cs,
                                       float dx, const hns_solve_options* opts, hns_solve_stats* stats,
                                       void* stream);

#ifdef __cplusplus
}
#endif
#endif /* HNS_H_ */
