/*
 * dsmc_c.h - C ABI of the MI355X DSMC engine.
 *
 * This is the drop-in boundary for the GPU path of denglide/CudaDsmc: every
 * entry point replaces one member of the reference's solver interface
 * (include/dsmc_base.h:40-89, include/gpu_dsmc.h:33-78), of its settings
 * reader (include/sim_settings.h:32-69) or of its kernel host API
 * (kernels/dsmc_kernel.cuh:29-78). The C++ class mirror that keeps the
 * reference's method names lives in dsmc_solver.hpp.
 *
 * Conventions
 *   - Every call returns DSMC_OK (0) or a negative DSMC_ERR_* code; nothing
 *     throws across the ABI. dsmc_last_error() gives the message of the last
 *     failing call on a handle (dsmc_global_last_error() for calls without one).
 *   - Handles own all device memory; one handle per simulation instance and
 *     any number of handles per process (the reference's module globals,
 *     dsmc_kernel.cu:117-137, allowed one).
 *   - Arrays passed in or out are host arrays with caller-owned storage,
 *     unless the name starts with d_ (device pointer) in the dsmc_op_* calls.
 *   - A handle is used by one host thread at a time.
 */
#ifndef DSMC_C_H_
#define DSMC_C_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSMC_OK 0
#define DSMC_ERR_INVALID (-1)     /* bad argument or settings */
#define DSMC_ERR_IO (-2)          /* file missing or unparsable */
#define DSMC_ERR_DEVICE (-3)      /* HIP runtime error or no device */
#define DSMC_ERR_STATE (-4)       /* call out of order (e.g. step before init) */
#define DSMC_ERR_CAPACITY (-5)    /* particle buffer too small */
#define DSMC_ERR_UNSUPPORTED (-6) /* solver type or feature not built */
#define DSMC_ERR_NUMERIC (-7)     /* NaN/inf found by dsmc_validate */

/* dsmc::SolverType (include/dsmc_base.h:82-87) */
typedef enum {
  DSMC_SOLVER_GPU = 0,
  DSMC_SOLVER_GPU_VBO = 1, /* visualisation variant: not built, returns UNSUPPORTED */
  DSMC_SOLVER_CPU = 2      /* reference returns no solver (dsmc_base.cu:225-226): UNSUPPORTED */
} dsmc_solver_type;

/* Physics of the step. REFERENCE reproduces the reference GPU step: hard
 * spheres, specular walls with first hit in list order, periodic/bounce
 * box faces, FP32 sampling every 100th step. PHYSICS adds VHS collisions,
 * diffuse walls at wall_temp, open (freestream) faces with inflow injection,
 * FP64 sampling and drag. */
typedef enum { DSMC_MODE_REFERENCE = 0, DSMC_MODE_PHYSICS = 1 } dsmc_mode;

/* Box face kinds in PHYSICS mode (order of face_bc[]: -x,+x,-y,+y,-z,+z). */
typedef enum {
  DSMC_FACE_SPECULAR = 0,
  DSMC_FACE_PERIODIC = 1,
  DSMC_FACE_FREESTREAM = 2 /* particles leaving are removed; freestream flux injected */
} dsmc_face;

/* settings_t (include/sim_settings.h:32-69) plus the physics-mode extensions.
 * Field order is part of the ABI. */
typedef struct dsmc_settings {
  int32_t benchmark;
  int32_t cpu;
  int32_t geometry;
  int32_t mt;
  uint32_t num_particles;
  uint32_t part_per_cell;
  float max_run_time;
  float dt;
  float molmass;  /* gas_props_t::molmass */
  float diameter; /* gas_props_t::diameter (d_ref for VHS) */
  float Tref;     /* gas_props_t::Tref */
  char gas_name[64];
  float density;
  float fnum;
  float temp;
  float stream_velocity[3];
  char periodic_condition; /* 'x','y','z' or 'n' */
  char pad_[3];
  float bmin[3];
  float bmax[3];
  float cell_size[3];
  int32_t grid_dim[3];
  /* ---- extensions ---- */
  int32_t mode;              /* dsmc_mode */
  float omega;               /* VHS viscosity index (0.5 = hard sphere) */
  float wall_temp;           /* PHYSICS: diffuse wall temperature [K]; <= 0: specular */
  int32_t face_bc[6];        /* PHYSICS: dsmc_face per box face */
  uint32_t target_particles; /* PHYSICS: if > 0, fnum = density*V_fluid/target */
  uint32_t seed;             /* first MT seed; stream s is seeded with seed + s */
  uint32_t sample_every;     /* sampling cadence (0: every 100th step as the reference) */
  uint32_t capacity;         /* particle buffer capacity (0: automatic; at most 2^30) */
  uint32_t n_streams;        /* RNG streams (0: automatic, multiple of 256) */
  char rng_param_file[256];  /* optional MT parameter rows (16-byte records:
                                matrix_a, mask_b, mask_c, unused), e.g. the
                                reference's data/MersenneTwister.dat */
} dsmc_settings;

/* Per-cell sums of the sampled moments (statistics_t, include/statistics.h:31-37,
 * widened to double; velMag is sum |v|^2). */
typedef struct dsmc_cell_stats {
  double num_in_sample;
  double xyz_sum[3];
  double vel_mag;
} dsmc_cell_stats;

/* The counters of simulation_state_t (include/draw_helper.h:40-78) that the
 * statistics output reads: simulated time (FP32 sum of dt per step, as
 * advanceSimTime), wall time, steps and samples. */
typedef struct dsmc_sim_state {
  float sim_time;
  float real_time;
  uint32_t step_count;
  uint32_t samples_count;
} dsmc_sim_state;

typedef struct dsmc_solver dsmc_solver;

/* ---------------------------------------------------------------- settings */
const char* dsmc_version(void);
const char* dsmc_global_last_error(void);
/* settings_t::settings_t (src/sim_settings.cpp:34-44) */
void dsmc_settings_default(dsmc_settings* s);
/* settings_t::readFromFile (src/sim_settings.cpp:46-132); the gases file named
 * by <gas file=...> is opened relative to the working directory first, as the
 * reference does, then relative to the config file's directory. */
int dsmc_settings_read_file(dsmc_settings* s, const char* path);
/* settings_t::readFromCmdLine (src/sim_settings.cpp:134-186) */
int dsmc_settings_read_cmdline(dsmc_settings* s, int argc, const char* const* argv);
/* main's particle count (kernels/dsmc.cu:324-330): ppc = (uint)(density *
 * cell_volume / fnum) in FP32, num_particles = ppc * cells. */
int dsmc_settings_derive_particles(dsmc_settings* s);

/* ------------------------------------------------------------------ solver */
/* dsmc::CreateSolver (kernels/dsmc_base.cu:221-233) */
int dsmc_create(int solver_type, dsmc_solver** out);
/* GPUDSMCSolver::~GPUDSMCSolver (include/gpu_dsmc.inl:40-66) */
void dsmc_destroy(dsmc_solver* h);
const char* dsmc_last_error(const dsmc_solver* h);
/* DSMCSolver::Init -> GPUDSMCSolver::InitData (kernels/dsmc_base.cu:43-47,
 * include/gpu_dsmc.inl:342-378): RNG tables, buffers, NTC state. */
int dsmc_init(dsmc_solver* h, const dsmc_settings* s);
/* readFromFile + particle derivation + Init in one call. */
int dsmc_init_from_config(dsmc_solver* h, const char* config_path);
/* DSMCSolver::SetModel + Model::CreateGrid (kernels/dsmc_base.cu:49-55,
 * src/model.cpp:733-769): triangle mesh (3 floats per vertex, 3 indices per
 * triangle; face normals optional, NULL = computed) and the uniform
 * collision grid of gx*gy*gz cells over the mesh bounding box. */
int dsmc_set_geometry(dsmc_solver* h, const float* vertices, uint32_t n_vertices,
                      const uint32_t* triangles, uint32_t n_triangles,
                      const float* face_normals, uint32_t gx, uint32_t gy, uint32_t gz);
/* dsmc_set_geometry followed by Model::MoveModel (src/model.cpp:709-731),
 * the order of the reference's driver (kernels/dsmc.cu:403-405): the grid is
 * built over the given vertices, then vertices and grid bounds shift by
 * move[3] while the cell size and triangle lists stay. */
int dsmc_set_geometry_moved(dsmc_solver* h, const float* vertices, uint32_t n_vertices,
                            const uint32_t* triangles, uint32_t n_triangles, const float* face_normals,
                            uint32_t gx, uint32_t gy, uint32_t gz, const float move[3]);
/* GPUDSMCSolver::InitSystemState (include/gpu_dsmc.inl:69-92). */
int dsmc_init_state(dsmc_solver* h, float ux, float uy, float uz, uint32_t* n_out);
/* Warm start: load local Maxwellians from per-cell fields (n [m^-3], u,v,w
 * [m/s], T [K]; arrays of grid cells). Replaces dsmc_init_state. */
int dsmc_init_state_from_fields(dsmc_solver* h, const float* n, const float* u, const float* v,
                                const float* w, const float* T, uint32_t n_cells, uint32_t* n_out);
/* GPUDSMCSolver::RunBenchmarkSimulationStep (include/gpu_dsmc.inl:221-274). */
int dsmc_step(dsmc_solver* h, float dt);
/* DSMCSolver::DoBenchmark (kernels/dsmc_base.cu:64-84); steps are enqueued
 * without host synchronisation (HIP graph replay where possible). */
int dsmc_run(dsmc_solver* h, uint32_t steps, float dt);
/* Enable/disable per-step sampling into the FP64 accumulators (PHYSICS) and
 * clear them; REFERENCE mode keeps its every-100th-step cadence. */
int dsmc_set_sampling(dsmc_solver* h, int enabled, int reset);
/* SampleCellStatistics (kernels/dsmc_kernel.cu:892-910): one extra sample of
 * the current state. */
int dsmc_sample(dsmc_solver* h);
/* SampleStatistics' per-cell array (include/gpu_dsmc.inl:277-282). */
int dsmc_get_cell_stats(dsmc_solver* h, dsmc_cell_stats* out, uint32_t n_cells);
/* Per-cell moments from the samples: n [m^-3], u, v, w [m/s], T [K]
 * (ProcessAndOutputStatistics, kernels/dsmc_base.cu:104-163). */
int dsmc_get_fields(dsmc_solver* h, float* n, float* u, float* v, float* w, float* T,
                    uint32_t n_cells);
/* GetStatistics (kernels/dsmc_kernel.cu:840-854): {selections, collisions,
 * sum of relative speeds of collisions, 0}, totals since init. */
int dsmc_get_collision_stats(dsmc_solver* h, double out[4]);
/* Body force {Fx, Fy, Fz} [N] and heat flux [W] averaged over the sampling
 * window (PHYSICS mode; zero otherwise). */
int dsmc_get_drag(dsmc_solver* h, double out[4]);
/* DSMCSolver::DumpStatistics (kernels/dsmc_base.cu:165-174): one more,
 * uncounted sample, then the reference's statistics.txt table with the
 * engine's counters (real time 0, as the reference's benchmark path leaves
 * it). */
int dsmc_dump_statistics(dsmc_solver* h, const char* path);
/* GPUDSMCSolver::SampleStatistics (include/gpu_dsmc.inl:277-282): the
 * collision totals as the float4 of GetStatistics, one more sample that the
 * samples counter does not include, and per cell the statistics_t record
 * (include/statistics.h:31-37) as 8 floats {numInSample, xyzSum[3], velMag,
 * pad[3]} (stats may be NULL). */
int dsmc_sample_statistics(dsmc_solver* h, float coll_out[4], float* stats, uint32_t n_cells);
/* ProcessAndOutputStatistics (kernels/dsmc_base.cu:104-163) of the current
 * sums as text, with the counters of st (NULL: the engine's own, see
 * dsmc_get_sim_state). *len receives the text length; buf (cap bytes, NUL
 * included) may be NULL to query it. Does not sample. */
int dsmc_format_statistics(dsmc_solver* h, const dsmc_sim_state* st, char* buf, uint64_t cap, uint64_t* len);
/* The engine's counters in simulation_state_t form (real_time 0). */
int dsmc_get_sim_state(dsmc_solver* h, dsmc_sim_state* out);
/* Counters: steps taken, live particles, samples taken. */
int dsmc_get_counts(dsmc_solver* h, uint64_t* steps, uint32_t* n_particles, uint32_t* samples);
/* Wait for all queued work of the handle. */
int dsmc_synchronize(dsmc_solver* h);
/* Scan the particle state for NaN/inf or particles outside the box. */
int dsmc_validate(dsmc_solver* h, uint32_t* n_bad);

/* ----------------------------------------------------- state (tests, I/O) */
/* Particles in storage order: pos/vel as x[n],y[n],z[n] blocks (3*cap floats),
 * keys = cell index per particle (0xFFFFFFFF: removed). */
int dsmc_get_particles(dsmc_solver* h, float* pos, float* vel, uint32_t* keys, uint32_t cap,
                       uint32_t* n_out);
int dsmc_set_particles(dsmc_solver* h, const float* pos, const float* vel, uint32_t n);
/* Cell offsets of the last step's ordering (n_cells + 1 entries). */
int dsmc_get_cell_start(dsmc_solver* h, uint32_t* out, uint32_t n);
/* Last step's stable permutation: sorted position -> previous storage index. */
int dsmc_get_perm(dsmc_solver* h, uint32_t* out, uint32_t n);
/* Per particle (sorted order) number of wall hits in the last step. */
int dsmc_get_hit_counts(dsmc_solver* h, uint32_t* out, uint32_t n);
/* NTC state per cell: {(sigma*c_r)max, remainder} pairs. */
int dsmc_get_ntc_state(dsmc_solver* h, float* out, uint32_t n_cells);
/* RNG parameters of each stream: twist vector a and tempering masks b, c. */
int dsmc_get_rng_params(dsmc_solver* h, uint32_t* a, uint32_t* b, uint32_t* c, uint32_t n_streams);
/* RNG stream state: 20 words per stream (the last 19 untempered words,
 * oldest first, and a zero), stream-major. */
int dsmc_get_rng_state(dsmc_solver* h, uint32_t* out, uint32_t n_words);
int dsmc_set_rng_state(dsmc_solver* h, const uint32_t* in, uint32_t n_words);
/* Cells per tile of the cell-tile kernel (a power of two in [32, 256]),
 * before loading. The default keeps a tile's expected particles near half
 * the LDS staging capacity (2048). Changes the RNG assignment (a run is
 * reproducible for a given value), not the physics. No reference
 * counterpart: the reference launches one thread per cell. */
int dsmc_set_tile_cells(dsmc_solver* h, uint32_t tile_cells);
/* Hybrid-Taus collision RNG state, 4 words {x, y, z, w} per cell (REFERENCE
 * mode with mt = 0: ComputeBirdCollisionsNoMTDev, kernels/dsmc_kernel.cu:
 * 912-1000, seeded per cell as CreateRandomSeedsUint4 does at :1846-1886). */
int dsmc_get_taus_state(dsmc_solver* h, uint32_t* out, uint32_t n_words);
int dsmc_set_taus_state(dsmc_solver* h, const uint32_t* in, uint32_t n_words);
/* PHYSICS with a body: near-body code per cell. 255 when no collision-grid
 * cell holding part of the body lies within the cell's box widened by one
 * cell per axis; else the distance from the cell's box to the nearest such
 * collision-grid cell in units of (smallest cell size) / 64, rounded down
 * (at most 254). A particle goes through the wall pass when its flight
 * |v| dt reaches that distance or it moves a cell or more along an axis. */
int dsmc_get_near_cells(dsmc_solver* h, uint8_t* out, uint32_t n_cells);
/* Per-cell fluid volume fraction used by PHYSICS mode (1 = no body). */
int dsmc_get_fluid_fraction(dsmc_solver* h, float* out, uint32_t n_cells);
/* Collision grid of the geometry: flags (0 empty, 1 triangles, 2 filled
 * interior) for gx*gy*gz cells. */
int dsmc_get_collision_grid(dsmc_solver* h, uint32_t* flags, uint32_t n);
/* Average milliseconds per step of each kernel class {scan, scatter, cells,
 * inject, total} over the steps timed since dsmc_set_timing; out_ms[5]
 * (when n > 5) receives the number of timed steps and out_ms[6] (n > 6) the
 * wall pass of PHYSICS runs with a body. */
int dsmc_get_timings(dsmc_solver* h, float* out_ms, uint32_t n);
/* Enable hipEvent timing of each kernel class (off by default). Events are
 * recorded on the solver's stream without host synchronisation and resolved
 * by dsmc_get_timings; enabling resets the averages. */
int dsmc_set_timing(dsmc_solver* h, int enabled);
/* Workgroups of the scatter kernel (grid-stride; does not change results). */
int dsmc_set_scatter_grid(dsmc_solver* h, uint32_t blocks);
/* dsmc_run replays pairs of steps from captured HIP graphs (default on);
 * off: every step is launched eagerly. Results are identical either way. */
int dsmc_set_graphs(dsmc_solver* h, int enabled);
/* Debug outputs of each step: bit 2 keeps the sorted permutation readable
 * by dsmc_get_perm, bit 4 records wall hits for dsmc_get_hit_counts. */
int dsmc_set_debug(dsmc_solver* h, uint32_t flags);
/* Internal step structure: tile cells, threads per tile, RNG streams. */
int dsmc_get_layout(dsmc_solver* h, uint32_t* tile_cells, uint32_t* threads, uint32_t* n_streams,
                    uint32_t* n_vwg);

/* Scalars derived at init/loading: {fnum, vhs_k, vhs_exp, (sigma c_r)max
 * start, cell_volume, sigma_hs, inj_lambda[6]} (12 floats). */
int dsmc_get_derived(dsmc_solver* h, float* out, uint32_t n);
/* The collision grid the solver built (as dsmc_cgrid_build). */
int dsmc_get_collision_grid_full(dsmc_solver* h, float bounds[9], uint32_t dims[3], uint32_t* flags,
                                 uint32_t* start, uint32_t* end, uint32_t* entries, uint32_t cap_cells,
                                 uint32_t cap_entries, uint32_t* n_entries);

/* ---------------------------------------------------- geometry (host only) */
/* Icosphere: icosahedron with subdiv midpoint subdivisions (20*4^subdiv
 * triangles), outward normals. Pass NULL buffers to query the sizes. */
int dsmc_mesh_icosphere(float cx, float cy, float cz, float r, uint32_t subdiv, float* verts, uint32_t vcap,
                        uint32_t* tris, uint32_t tcap, float* normals, uint32_t* nv, uint32_t* nt);
/* Closed sphere-cone along +x: nose tip at (x0, cy, cz), nose radius rn,
 * cone half-angle, base radius rb, n_az azimuthal and ~n_profile meridian
 * segments. */
int dsmc_mesh_sphere_cone(float x0, float cy, float cz, float rn, float half_angle_deg, float rb, uint32_t n_az,
                          uint32_t n_profile, float* verts, uint32_t vcap, uint32_t* tris, uint32_t tcap,
                          float* normals, uint32_t* nv, uint32_t* nt);
/* Model::LoadRaw (src/model.cpp:348-401): vertices, indices, face normals. */
int dsmc_mesh_load_raw(const char* path, float* verts, uint32_t vcap, uint32_t* tris, uint32_t tcap,
                       float* normals, uint32_t* nv, uint32_t* nt);
/* Triangle/box overlap of the grid builder (AABB_t::tri_overlaps ->
 * triBoxOverlap, src/aabb.cpp:38-50): 1 when they overlap. */
int dsmc_tri_box_overlap(const float center[3], const float half[3], const float a[3], const float b[3],
                         const float c[3]);
/* Model::CreateGrid (src/model.cpp:733-769) + reg_grid_t::encloseGrid
 * (src/reggrid.cpp:49-79) for mode REFERENCE; padded grid with flood-filled
 * interior for PHYSICS. bounds = min[3], max[3], cell[3]. */
int dsmc_cgrid_build(int mode, const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                     const float* normals, uint32_t gx, uint32_t gy, uint32_t gz, float bounds[9],
                     uint32_t dims[3], uint32_t* flags, uint32_t* start, uint32_t* end, uint32_t* entries,
                     uint32_t cap_cells, uint32_t cap_entries, uint32_t* n_entries);

/* ------------------------------------------------ batched ensembles (C4) */
/* Several independent instances (each its own handle: grid, particle count,
 * freestream, body, seeds, dt) advanced by the same kernel launches, one
 * launch per step kernel for the whole ensemble (grid.y = instance), with
 * pairs of steps replayed from captured HIP graphs. Members must be loaded
 * and at the same step; results are identical to stepping each member alone.
 * The batch form of the reference's DoBenchmark loop (dsmc_base.cu:64-84). */
typedef struct dsmc_batch dsmc_batch;
int dsmc_batch_create(dsmc_solver* const* solvers, uint32_t n, dsmc_batch** out);
void dsmc_batch_destroy(dsmc_batch* b);
const char* dsmc_batch_last_error(const dsmc_batch* b);
int dsmc_batch_run(dsmc_batch* b, uint32_t steps);
int dsmc_batch_set_sampling(dsmc_batch* b, int enabled, int reset);
int dsmc_batch_set_graphs(dsmc_batch* b, int enabled);
int dsmc_batch_synchronize(dsmc_batch* b);

/* --------------------------------------- kernel-level operators (d_ = device) */
/* MT(19,9) streams: seeds[s] with parameters (a[s], b[s], c[s]); writes
 * n_draws outputs per stream, d_out[d * n_streams + s]. Mirrors
 * InitTwisterFast + GetRandomIntegerFast (kernels/mtwister.cu:70-116). */
int dsmc_op_mt_generate(const uint32_t* d_a, const uint32_t* d_b, const uint32_t* d_c,
                        const uint32_t* d_seeds, uint32_t n_streams, uint32_t n_draws,
                        uint32_t* d_out, void* stream);
/* CalculateCellIDs (kernels/dsmc_kernel.cu:1701-1732) on SoA positions. */
int dsmc_op_cell_ids(const dsmc_settings* s, const float* d_x, const float* d_y, const float* d_z,
                     uint32_t n, uint32_t* d_keys, void* stream);
/* SortParticlesIndices + FindCellBoundaries bounds (dsmc_kernel.cu:1734-1808,
 * 2033-2051): stable sort of keys into a permutation (d_perm[n]: indices of
 * the keys < n_cells in key order, ties in index order, then the indices of
 * the dropped keys >= n_cells) and cell offsets (d_cell_start[n_cells + 1]).
 * An LSD radix sort, 8 bits per pass over the bits of n_cells, with the
 * workspace d_work of dsmc_op_sort_work_words(n, n_cells) uint32 (the
 * reference's thrust call allocated its temporary storage itself). */
uint64_t dsmc_op_sort_work_words(uint32_t n, uint32_t n_cells);
int dsmc_op_sort(const uint32_t* d_keys, uint32_t n, uint32_t n_cells, uint32_t* d_perm,
                 uint32_t* d_cell_start, uint32_t* d_work, void* stream);
/* dsmc_op_sort without the d_work argument. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops). Allocates nothing: the workspace is a
 * file-scope device array of 17.3 MB, held on each GPU the library runs on
 * until the process ends. n and n_cells at most 2^20 (DSMC_ERR_CAPACITY
 * beyond). All calls on one GPU share its array, so calls that overlap on one
 * GPU (several streams or host threads) race. */
int dsmc_op_sort_static(const uint32_t* d_keys, uint32_t n, uint32_t n_cells, uint32_t* d_perm,
                        uint32_t* d_cell_start, void* stream);

/* ------------------------------------ field operator (warm-start loading) */
/* The reference has no counterpart: these kernels train and evaluate the 3-D
 * convolutional operator that predicts steady fields (n, u, v, w, T) on the
 * cell grid from the case parameters; dsmc_init_state_from_fields then loads
 * particles from the prediction instead of the uniform freestream of
 * GenerateInitialVelocities (kernels/dsmc_kernel.cu:629-680).
 * Layouts: activations BF16 channels-last [B][Z][Y][X][C]; weights BF16
 * [taps][Cout][Cin] (taps 27 = 3x3x3 with dilation dil and zero padding, or
 * 1); biases, gradients and the optimiser state FP32. All pointers are
 * device pointers; calls are asynchronous on `stream`. */

/* y = act(conv(x, w) + bias) (+ res): relu != 0 applies ReLU before the
 * residual add; h (optional) receives the activation before the residual;
 * y (BF16) and/or y32 (FP32) receive the result. Cin % 16 == 0,
 * Cout % 32 == 0, Cout <= 128. */
int dsmc_no_conv_fwd(const void* x, const void* w, const float* bias, const void* res, void* y, float* y32,
                     void* h, int B, int Z, int Y, int X, int Cin, int Cout, int taps, int dil, int relu,
                     void* stream);
/* wb[mirror(t)][ci][co] = w[t][co][ci]: conv_fwd with wb computes the data
 * gradient of conv_fwd with w. */
int dsmc_no_weight_flip(const void* w, void* wb, int taps, int Cout, int Cin, void* stream);
/* dw[t][co][ci] (+)= sum over voxels m of dy[m][co] * x[m + offset(t)][ci],
 * through per-chunk partials (part: ceil(M / chunk) * taps * Cout * Cin
 * floats) reduced in a fixed order. chunk % 64 == 0; Cin, Cout multiples of
 * 32, at most 128. */
int dsmc_no_conv_bwd_w(const void* x, const void* dy, float* dw, float* part, long part_floats, int B, int Z,
                       int Y, int X, int Cin, int Cout, int taps, int dil, int chunk, int accumulate, void* stream);
/* db[co] (+)= sum over voxels of dy[m][co]. */
int dsmc_no_bias_grad(const void* dy, long M, int Cout, float* db, float* part, long part_floats, int accumulate,
                      void* stream);
/* dz = dy where h > 0, else 0. */
int dsmc_no_relu_bwd(const void* dy, const void* h, void* dz, long n, void* stream);
/* c = a + b (BF16). */
int dsmc_no_add(const void* a, const void* b, void* c, long n, void* stream);
/* Masked mean-squared error of pred[M][C] (first Cf channels) against
 * target[M][Cf]: dpred (BF16 [M][C]) = 2 (pred - target) mask * inv_count,
 * loss_part[block] = partial sums of (pred - target)^2 mask * inv_count over
 * blocks of 256 elements. */
int dsmc_no_mse(const float* pred, const float* target, const float* mask, long M, int C, int Cf, float inv_count,
                void* dpred, float* loss_part, void* stream);
/* Adam with bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t and L2
 * weight decay wd; refreshes the BF16 copy wbf when given. */
int dsmc_no_adam(float* w, const float* g, float* m1, float* m2, void* wbf, long n, float lr, float b1, float b2,
                 float eps, float c1, float c2, float wd, void* stream);
int dsmc_no_f32_to_bf16(const float* a, void* b, long n, void* stream);
/* Operator inputs [B][Z][Y][X][C]: cell-centre coordinates in [-1, 1],
 * signed distance to the sphere sphere[4b..4b+3] = (cx, cy, cz, r) in units
 * of the box height, and the five case parameters p[5b..5b+4]; channels 9..C-1
 * zero. box = (xmin, ymin, zmin, xmax, ymax, zmax). */
int dsmc_no_inputs(const float* p, const float* sphere, const float* box, int B, int Z, int Y, int X, int C,
                   void* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DSMC_C_H_ */
