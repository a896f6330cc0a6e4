/*
 * dsmc_cpu.h - CPU oracle of the DSMC step (TEST INFRASTRUCTURE).
 *
 * A sequential C++ restatement of the reference's GPU algorithm
 * (kernels/dsmc_kernel.cu, kernels/mtwister.cu, kernels/tri_intersect.cu,
 * kernels/vectors.cu, sequenced as in include/gpu_dsmc.inl:221-274) and of
 * the physics-mode extensions, used only as the checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg. The product
 * (cudadsmc_amd) never links or calls it.
 */
#ifndef DSMC_CPU_H_
#define DSMC_CPU_H_

#include <stdint.h>

#include "dsmc/dsmc_c.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sim oracle_sim;

/* Generic Mersenne Twister in block form (Matsumoto & Nishimura 1998):
 * n words, middle m, one separation bit (upper mask ~lower), tempering
 * shifts u,s,t,l and masks b,c; Knuth's 1812433253 initialisation. */
void oracle_mt_generic(uint32_t n, uint32_t m, uint32_t a, uint32_t lower_mask, uint32_t u, uint32_t s,
                       uint32_t b, uint32_t t, uint32_t c, uint32_t l, uint32_t seed, uint32_t n_out,
                       uint32_t* out);
/* MT(19,9) streams of the DSMC engine: out[d * n_streams + s]. */
void oracle_mt_streams(const uint32_t* a, const uint32_t* b, const uint32_t* c, const uint32_t* seeds,
                       uint32_t n_streams, uint32_t n_draws, uint32_t* out);
/* GetCellID3D on the grid of the settings. */
void oracle_cell_ids(const dsmc_settings* s, const float* x, const float* y, const float* z, uint32_t n,
                     uint32_t* keys);
/* Stable sort by key (0xFFFFFFFF dropped): perm and cell offsets. */
void oracle_stable_sort(const uint32_t* keys, uint32_t n, uint32_t n_cells, uint32_t* perm,
                        uint32_t* cell_start);
/* Moller-Trumbore of the reference (tri_intersect.cu:14-56). */
int oracle_intersect(const float o[3], const float d[3], const float v0[3], const float v1[3],
                     const float v2[3], float out_tuv[3]);
/* detmath probes for the libm accuracy tests: f = 0 sin, 1 cos, 2 log, 3 exp2, 4 pow(x, y). */
void oracle_detmath(int f, const float* x, const float* y, uint32_t n, float* out);

/* Simulation. tile_cells/tile_threads/n_streams must equal the engine's
 * layout (dsmc_get_layout); stream s uses parameters (a[s], b[s], c[s]). */
oracle_sim* oracle_create(const dsmc_settings* s, uint32_t tile_cells, uint32_t tile_threads,
                          uint32_t n_streams, const uint32_t* a, const uint32_t* b, const uint32_t* c);
void oracle_destroy(oracle_sim* o);
/* Collision grid as the engine built it (bounds = min[3], max[3], cell[3];
 * flags, ranges, entries), the mesh and the per-cell fluid fractions
 * (NULL: all 1). */
int oracle_set_geometry(oracle_sim* o, const float* verts, uint32_t nv, const uint32_t* tris,
                        uint32_t nt, const float* normals, const float cg_bounds[9], const uint32_t cg_dim[3], const uint32_t* flags, const uint32_t* start,
                        const uint32_t* end, const uint32_t* entries, uint32_t n_entries,
                        const float* fluid_frac);
/* Derived scalars the engine computed at init (fnum, (sigma c_r)max start,
 * VHS constant and exponent, freestream injection per face cell). */
int oracle_set_scalars(oracle_sim* o, float fnum, float vhs_k, float vhs_exp, const float inj_lambda[6]);
/* RNG windows: 20 words per stream, stream-major (as dsmc_get_rng_state). */
int oracle_set_rng(oracle_sim* o, const uint32_t* words, uint32_t n_words);
int oracle_get_rng(oracle_sim* o, uint32_t* words, uint32_t n_words);
int oracle_set_ntc(oracle_sim* o, const float* ntc, uint32_t n_cells);
int oracle_get_ntc(oracle_sim* o, float* ntc, uint32_t n_cells);
/* The locality order in which the step visits the tiles; returns the tile
 * count, or -1 when n is too small. */
int oracle_tile_order(oracle_sim* o, uint32_t* out, uint32_t n);
/* Positions of that order visited by workgroup vw, in order; returns their
 * number, or -1. */
int oracle_tile_walk(oracle_sim* o, uint32_t vw, uint32_t* out, uint32_t n);
/* State in storage order: pos/vel as 3 blocks of n. */
int oracle_set_state(oracle_sim* o, const float* pos, const float* vel, uint32_t n);
uint32_t oracle_n_stored(oracle_sim* o);
int oracle_get_state(oracle_sim* o, float* pos, float* vel, uint32_t* keys, uint32_t cap);
/* Particle loading (InitSystemState) with the stream velocity. */
int oracle_load(oracle_sim* o, float ux, float uy, float uz);
/* One step; sample != 0 accumulates the per-cell sums. */
int oracle_step(oracle_sim* o, int sample);
/* PHYSICS NTC of one cell (n particles, cell-local arrays, fluid volume
 * vol, nt = {(sigma c_r)max, remainder}) at step `stepno`; acc3 receives
 * {candidates, collisions, sum of c_r of the collisions}. */
int oracle_ntc_cell(oracle_sim* o, uint32_t cell, uint32_t stepno, float* vx, float* vy, float* vz, uint32_t n,
                    float vol, float* nt, double* acc3);
/* The keyed permutation of the PHYSICS pairing: pi(x) on [0, n). */
uint32_t oracle_pair_perm(uint32_t n, const uint32_t keys[4], uint32_t x);
int oracle_get_perm(oracle_sim* o, uint32_t* out, uint32_t n);
int oracle_get_cell_start(oracle_sim* o, uint32_t* out, uint32_t n);
int oracle_get_hits(oracle_sim* o, uint32_t* out, uint32_t n);
/* Per-cell sums {N, sx, sy, sz, s2} as doubles (FP32 sums widened in REFERENCE mode). */
int oracle_get_samples(oracle_sim* o, double* out, uint32_t n_cells);
/* Per-cell REFERENCE counters {sel, coll, sum c_r, 0}. */
int oracle_get_cstat(oracle_sim* o, float* out, uint32_t n_cells);
/* Totals {sel, coll, sum c_r, Fx, Fy, Fz, heat} of the PHYSICS accumulators (without fnum). */
int oracle_get_acc(oracle_sim* o, double* out, uint32_t n);
/* Number of threads used by oracle_step (OpenMP). */
int oracle_threads(void);
/* OpenMP threads of the following calls (timing a one-thread baseline). */
void oracle_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif /* DSMC_CPU_H_ */
