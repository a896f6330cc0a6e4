// Device-side data layout of one DSMC instance and the per-thread RNG.
//
// HBM layout (all SoA, 4-byte lanes, 256-byte aligned allocations):
//   particles  x,y,z,vx,vy,vz : float[cap], double-buffered (parity = step & 1)
//   keys                      : uint32[cap] per buffer (cell of the particle,
//                               0xFFFFFFFF = removed)
//   count                     : uint32[ncells]      next-step counting
//   cursor                    : uint32[ncells]      scatter: next free sorted position of each cell
//   local_cnt                 : uint32[ncells]      particles moved into the cell from its own tile
//   tile_p0                   : uint32[ntiles + 1]  first stored slot of each tile's particles
//   cell_start                : uint32[ncells + 1]  offsets of the current order
//   perm / perm_tmp           : uint32[cap]         sorted position -> storage index
//   ntc                       : float2[ncells]      {(sigma c_r)max, remainder}
//   samples                   : double[5][ncells]   {N, sum vx, sum vy, sum vz, sum |v|^2}
//   RNG                       : uint32[20][n_streams] (19-word window + spare)
// A step is scan(count) -> scatter -> cells(tile kernel) [-> inject].
#pragma once
#include <stdint.h>

#include "detmath.h"

namespace dsmc_dev {

constexpr uint32_t kInvalidKey = 0xFFFFFFFFu;
constexpr uint32_t kDeferredKey = 0xFFFFFFFEu;  // PHYSICS: moved by the wall pass
// A key the wall pass stored carries this bit, so that the scatter does not
// take the particle for one that stayed in its tile (cells < 2^31).
constexpr uint32_t kWallKeyBit = 0x80000000u;
__host__ __device__ inline uint32_t key_cell(uint32_t k) { return k >= kDeferredKey ? kInvalidKey : (k & ~kWallKeyBit); }
constexpr uint32_t kTileCells = 256;    // most cells per tile (one lane per cell)
constexpr uint32_t kTileThreads = 256;  // threads of the tile kernel
constexpr uint32_t kLdsCap = 2048;      // particles staged in LDS per tile
constexpr uint32_t kMaxCap = 1u << 30;   // particles per instance: 4-byte arrays stay below 4 GiB
constexpr uint32_t kMtWords = 20;       // 19 window words + 1 spare (zero)
constexpr uint32_t kRefMtStreams = 65536;  // the reference's MT_RNG_COUNT (MersenneTwister.h:31)
constexpr int kMtN = 19, kMtM = 9;
constexpr uint32_t kMtUpper = 0xFFFFFFFEu, kMtLower = 0x1u;
constexpr int kShift0 = 12, kShiftB = 7, kShiftC = 15, kShift1 = 18;

// Physical constants (reference values where the reference has them).
constexpr float kBoltzRef = 1.3806e-23f;          // DSMC_BOLZ, include/dsmc.h:33
constexpr double kBoltz = 1.380649e-23;           // CODATA, physics mode
constexpr float kPiRef = 3.14159265358979f;       // DSMC_PI, include/dsmc.h:27
constexpr double kBounceDelta = 0.001;            // BOUNCE_DELTA, dsmc_kernel.cu:53
constexpr float kGeomEps = 0.0001f;               // EPS, dsmc_kernel.cu:1251
constexpr float kModelWidth = 0.0001f;            // MODEL_WIDTH, dsmc_kernel.cu:1253
constexpr int kCollisionLimit = 10;               // COLLISION_LIMIT, dsmc_kernel.cu:1250
constexpr double kTriEpsilon = 0.000001;          // EPSILON, tri_intersect.cu:1
constexpr double kTriRelEpsilon = 1.0e-9;         // PHYSICS: |det| vs (|e1|^2 + |e2|^2)
constexpr int kPhysMaxBounces = 8;
// Smallest fluid fraction used for a cell's NTC volume: the fraction is
// sampled at 4^3 points, so a cell reported empty may still hold a sliver of
// fluid that particles reach; its volume is then taken as one sample's share.
constexpr float kMinFluidFrac = 1.0f / 64.0f;
// Candidate pairs allowed per cell and step. A larger count only arises from
// a corrupted state (non-finite (sigma c_r)max); the count is clamped and
// status bit kStatusPairs is raised instead of spinning for ~2^32 pairs.
constexpr uint32_t kMaxPairsPerCell = 1u << 20;
// Fixed-point unit of the summed relative speeds of the collision counters:
// each collision adds round(c_r * 2^16).
constexpr double kCrScale = 65536.0;
// Bits of the sticky status word ctr[3].
constexpr uint32_t kStatusCapacity = 1u, kStatusPairs = 2u;

// Collision-grid cell flags.
constexpr uint32_t kCgEmpty = 0, kCgSurface = 1, kCgFilled = 2;

enum Mode : uint32_t { kModeReference = 0, kModePhysics = 1 };
enum Face : uint32_t { kFaceSpecular = 0, kFacePeriodic = 1, kFaceFreestream = 2 };

struct InstDesc {
  // ---- simulation grid
  float gmin[3], gmax[3], cell[3];
  float inv_cell[3];   // 1 / cell
  uint32_t dim[3];
  uint32_t ncells;
  uint32_t tile_cells;    // cells per tile: a power of two in [32, kTileCells]
  uint32_t ntiles;
  uint32_t tile_dyn;      // large tiles: the tile kernel takes tiles from atomic queues (else a fixed
                          // sequence) and counts the particles that stay in their tile for the scatter
  float cell_volume;
  // ---- gas and NTC
  float mass;
  float sigma_hs;   // REFERENCE: (float)(pi d^2) evaluated in double
  float vhs_k;      // PHYSICS: sigma*c_r = vhs_k * c_r^vhs_exp
  float vhs_exp;    // 2 - 2 omega (1 for hard spheres)
  float avg_conc;   // REFERENCE: global particles per cell
  float fnum;
  float dt;
  float vm_wall;    // sqrt(2 k T_wall / m); <= 0 specular walls (PHYSICS)
  float wall_temp;
  uint32_t mode;
  uint32_t sample_on;     // this instance samples in launches flagged for sampling
  uint32_t periodic;      // REFERENCE box: 'x','y','z','n'
  uint32_t face_bc[6];    // PHYSICS box faces
  uint32_t has_geom;
  // ---- freestream (PHYSICS injection)
  float n_inf, T_inf, U[3], vm_inf;  // vm_inf = sqrt(2 k T_inf / m)
  float inj_lambda[6];               // expected injections per face cell per step
  uint32_t face_off[7];              // prefix of face cells per face (only freestream faces)
  // ---- collision grid
  float cg_min[3], cg_max[3], cg_cell[3];
  uint32_t cg_dim[3];
  uint32_t cg_ncells;
  const uint32_t* cg_flag;
  const uint32_t* cg_start;  // per cg cell: first entry (0xFFFFFFFF: none)
  const uint32_t* cg_end;
  const uint32_t* cg_tri;    // entries: triangle index
  const float4* cg_rec;      // PHYSICS entries: {v0 + index bits in w, e1 = v1 - v0, e2 = v2 - v0, normal}
  const uint2* cg_se;        // PHYSICS per cg cell: {first, end} of its entries (equal: not a surface cell)
  const float4* tv0;         // per triangle vertices and unit face normal (w unused)
  const float4* tv1;
  const float4* tv2;
  const float4* tn;
  // ---- particles [2] = double buffer
  float* px[2];
  float* py[2];
  float* pz[2];
  float* vx[2];
  float* vy[2];
  float* vz[2];
  uint32_t* key[2];
  uint32_t cap;
  // ---- cells
  uint32_t* count;
  uint32_t* cell_start;
  uint32_t* perm;
  uint32_t* cursor;     // per cell: next free position of the scatter for particles from other tiles
                        // (the scan sets it to cell_start + local_cnt)
  uint32_t* local_cnt;  // per cell: particles the tile kernel moved into it from the cell's own tile
  uint32_t* tile_p0;    // [ntiles + 1]: slot of the first particle the tile kernel stored for each tile;
                        // [ntiles] = end of the tile-ordered slots (all zero: no tile layout yet)
  uint32_t* perm_tmp;   // overflow tiles: sorted copy of the segment
  uint8_t* owner_tmp;   // overflow tiles: cell of each position within the tile
  float2* ntc;
  float4* cstat;        // REFERENCE per-cell {selections, collisions, sum c_r, 0}
  double* samp;         // [5][ncells]
  float* fluid_frac;    // PHYSICS per-cell fluid volume fraction
  const uint8_t* near_cell;  // PHYSICS with a body: near-body code per cell (255: far; else the distance
                             // to the body's collision-grid cells in units of near_unit)
  float near_unit;      // smallest cell size / 64
  // ---- scalars and per-workgroup accumulators
  uint32_t* ctr;        // [0] stored, [1] sorted, [2] injected particles, [3] status bits, [4] step number,
                        // [5] deferred particles of the step, [8..15] the tile kernel's tile queues
  const uint32_t* tile_order;  // tiles in locality order [ntiles] (null: identity)
  unsigned long long* defer_bits;  // PHYSICS with a body: bit s set = slot s deferred [cap / 64 + 2]
  uint32_t* defer_list; // PHYSICS with a body: deferred slots of this step, compacted [cap]
  uint32_t n_wall_wg;   // PHYSICS with a body: workgroups of the wall pass
  long long* wall_part;  // PHYSICS with a body: [n_wall_wg][4] each wall-pass workgroup's fixed-point sums
                        // of the step (folded into wall_acc by one workgroup after the pass)
  double wall_scale[2]; // fixed-point scale of a velocity change (2^36 / v_ref) and of a change of |v|^2
                        // (2^33 / v_ref^2)
  double wall_unit[2];  // value of one fixed-point unit: m v_ref / 2^36 (momentum), m v_ref^2 / 2^34 (energy)
  double* wall_acc;     // PHYSICS with a body: wall {Fx, Fy, Fz, heat} sums over the steps [4]
  uint32_t wall_key[2]; // Philox key of the wall re-emission draws
  long long* wg_acc;    // [n_vwg][kAccSlots] collision counters per workgroup (integers, see CollAcc)
  uint32_t* hits;       // per sorted particle wall hits (debug, may be null)
  // ---- RNG
  uint32_t* taus;       // REFERENCE with mt = 0: Hybrid-Taus state per cell, 4 words (else null)
  uint32_t* mt_state;   // [20][n_streams]
  const uint32_t* mt_a;
  const uint32_t* mt_b;
  const uint32_t* mt_c;
  uint32_t n_streams;
  uint32_t n_vwg;       // virtual workgroups of the tile kernel = n_streams / 256
  // ---- injection face cells
  uint32_t n_face_cells;
  uint32_t* face_count; // [n_face_cells + 1] counts then offsets
};

// Global-memory view of a pointer held in a descriptor. Pointers loaded from
// memory are generic to the compiler, so accesses through them compile to
// FLAT instructions, which count against the LDS wait counter as well and
// serialise LDS work behind outstanding global loads; through these views
// they compile to global_* instructions.
template <class T>
using gptr = __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ gptr<T> gmem(T* p) {
  return (gptr<T>)p;
}

// Element i of a global array addressed by a 32-bit byte offset from its
// base (every particle and cell array is below 4 GiB: the capacity is at
// most kMaxCap, the cells at most 16M). With a workgroup-uniform base the access compiles to the
// scalar-base form (global_load vdst, voffset, sbase) instead of a 64-bit
// address per lane, which costs two VALU instructions per access.
template <class T>
__device__ __forceinline__ gptr<T> gelem(gptr<T> base, uint32_t i) {
  return (gptr<T>)((gptr<const char>)base + (uint32_t)(i * (uint32_t)sizeof(T)));
}

// HIP's vector structs cannot be copied out of a global-space reference: go
// through the compiler's native vector types.
typedef float dsmc_f2 __attribute__((ext_vector_type(2)));
typedef float dsmc_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t dsmc_u4 __attribute__((ext_vector_type(4)));
typedef uint32_t dsmc_u2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float4 gload(gptr<const float4> p) {
  const dsmc_f4 v = *(gptr<const dsmc_f4>)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 gload(gptr<const float2> p) {
  const dsmc_f2 v = *(gptr<const dsmc_f2>)p;
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ uint4 gload(gptr<const uint4> p) {
  const dsmc_u4 v = *(gptr<const dsmc_u4>)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 gload(gptr<const uint2> p) {
  const dsmc_u2 v = *(gptr<const dsmc_u2>)p;
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ void gstore(gptr<float4> p, float4 a) {
  *(gptr<dsmc_f4>)p = dsmc_f4{a.x, a.y, a.z, a.w};
}
__device__ __forceinline__ void gstore(gptr<float2> p, float2 a) { *(gptr<dsmc_f2>)p = dsmc_f2{a.x, a.y}; }
__device__ __forceinline__ void gstore(gptr<uint4> p, uint4 a) {
  *(gptr<dsmc_u4>)p = dsmc_u4{a.x, a.y, a.z, a.w};
}

__device__ __forceinline__ uint32_t gatomic_add(gptr<uint32_t> p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits for every vector memory operation of the wave (s_waitcnt vmcnt(0);
// expcnt and lgkmcnt not waited for). The compiler counts loads issued under
// a branch as possibly not issued, so its own wait for the first use of such
// a load also waits for every store issued after it; one wait here, where
// the loads of a batch are all needed anyway, leaves no load pending for the
// stores that follow.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Walk of workgroup vw (of a grid of G) over the positions of the tile
// order. The hardware deals workgroups out to the 8 XCDs round-robin (XCD =
// vw mod 8). With G a multiple of 8, the order is cut into patches of
// kTilePatch positions dealt out to the XCDs round-robin as well, and the
// G/8 workgroups of an XCD take the positions of its patches in turn: the
// tiles of one patch (neighbours in space) run together under one L2, and
// every XCD gets patches from all over the domain (the same share of
// expensive near-body tiles). Otherwise workgroup vw takes positions vw,
// vw + G, ... Together with the stream of each lane (vw * 256 + lane) the
// walk fixes the RNG assignment. The static walk is tile_queue_pos below
// with e = vw / 8 + k * G / 8 and q = vw mod 8 (G a multiple of 8), or
// e = vw + k * G otherwise; positions >= ntiles end the walk.
constexpr uint32_t kTilePatch = 128;

// Tile queues of the tile kernel: workgroups take tiles dynamically, one
// atomic increment per tile. With a workgroup count that is a multiple of 8
// there is one queue per XCD (workgroup w uses queue w mod 8); element e of
// queue q is position ((e / kTilePatch) * 8 + q) * kTilePatch + e mod
// kTilePatch of the locality order, i.e. the queue runs through that XCD's
// patches in order, so neighbouring tiles run together under one L2 and
// every XCD gets the same share of the domain. Else one queue over all
// positions. No result depends on which workgroup takes which tile.
__host__ __device__ inline uint32_t tile_queue_pos(uint32_t e, uint32_t q, bool per_xcd) {
  return per_xcd ? ((e / kTilePatch) * 8u + q) * kTilePatch + e % kTilePatch : e;
}
constexpr uint32_t kTileQueueCtr = 8;  // ctr[8 + q]: next element of queue q

// Per virtual-workgroup accumulator slots.
enum AccSlot {
  kAccSel = 0, kAccColl, kAccCr, kAccFx, kAccFy, kAccFz, kAccHeat, kAccHits, kAccSlots
};

// ------------------------------------------------------------------- MT(19,9)
// Sliding-window form of the DCMT generator: the state is the last 19
// untempered words (x_k .. x_{k+18}); a draw computes
//   x_{k+19} = x_{k+9} ^ twist(x_k, x_{k+1})
// shifts the window by one and returns the tempered x_{k+19}. This is the
// recurrence of the reference's incremental form (GetRandomIntegerFast,
// kernels/mtwister.cu:70-100) and of Matsumoto-Nishimura's block form
// (genrand_mt); keeping every index static lets the window live in VGPRs.
struct Mt {
  uint32_t st[kMtN];
  uint32_t a, b, c;
};

DM_HD uint32_t mt_twist(uint32_t lo, uint32_t hi, uint32_t a) {
  uint32_t x = (lo & kMtUpper) | (hi & kMtLower);
  return (x >> 1) ^ ((x & 1u) ? a : 0u);
}

DM_HD uint32_t mt_next(Mt& m) {
  uint32_t x = m.st[kMtM] ^ mt_twist(m.st[0], m.st[1], m.a);
#pragma unroll
  for (int k = 0; k < kMtN - 1; ++k) m.st[k] = m.st[k + 1];
  m.st[kMtN - 1] = x;
  x ^= (x >> kShift0);
  x ^= (x << kShiftB) & m.b;
  x ^= (x << kShiftC) & m.c;
  x ^= (x >> kShift1);
  return x;
}

// GetRandomFloatFast (kernels/mtwister.cu:102-105): ((float)u + 1) / 2^32, in (0, 1].
DM_HD float mt_float(Mt& m) { return ((float)mt_next(m) + 1.0f) / 4294967296.0f; }

// InitTwisterFast (kernels/mtwister.cu:107-116): Knuth's 1812433253 seeding.
DM_HD void mt_seed(Mt& m, uint32_t seed) {
  m.st[0] = seed;
#pragma unroll
  for (int i = 1; i < kMtN; ++i) m.st[i] = 1812433253u * (m.st[i - 1] ^ (m.st[i - 1] >> 30)) + (uint32_t)i;
}

// Generic draws for the NTC loop.
DM_HD uint32_t rng_u32(Mt& m) { return mt_next(m); }
DM_HD float rng_f(Mt& m) { return mt_float(m); }

// ----------------------------------------------------------- Hybrid Taus
// The reference's collision RNG when the Mersenne Twister is off
// (HybridTausRng / HybridTausRngInt, kernels/mtwister.cu:118-153): three
// Tausworthe components with the taus88 parameters of L'Ecuyer (1996),
// (k, q, s) = (31, 13, 12), (29, 2, 4), (28, 3, 17), combined by XOR with the
// LCG x <- 1664525 x + 1013904223 (mod 2^32). One state per cell.
struct Taus {
  uint32_t x, y, z, w;
};

DM_HD uint32_t taus_component(uint32_t& z, int s1, int s2, int s3, uint32_t m) {
  uint32_t b = ((z << s1) ^ z) >> s2;
  z = ((z & m) << s3) ^ b;
  return z;
}

DM_HD uint32_t rng_u32(Taus& t) {
  uint32_t a = taus_component(t.x, 13, 19, 12, 0xFFFFFFFEu);
  uint32_t b = taus_component(t.y, 2, 25, 4, 0xFFFFFFF8u);
  uint32_t c = taus_component(t.z, 3, 11, 17, 0xFFFFFFF0u);
  t.w = t.w * 1664525u + 1013904223u;
  return a ^ b ^ c ^ t.w;
}

// 2.3283064365387e-10 * u evaluated in double, rounded to float.
DM_HD float rng_f(Taus& t) { return (float)(2.3283064365387e-10 * (double)rng_u32(t)); }

}  // namespace dsmc_dev
