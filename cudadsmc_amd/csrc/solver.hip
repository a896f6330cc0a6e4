// Solver object and C ABI (include/dsmc/dsmc_c.h).
//
// Host orchestration of one DSMC instance on one GPU: owns the device
// buffers, the RNG streams, the collision grid and the step sequence. It
// replaces GPUDSMCSolver (include/gpu_dsmc.h, include/gpu_dsmc.inl) and the
// DSMCSolver base (kernels/dsmc_base.cu); the reference's module globals
// (dsmc_kernel.cu:117-137) become per-handle members.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "dsmc/dsmc_c.h"
#include "host/host_internal.h"
#include "kernels/launch.h"

using namespace dsmc_dev;

namespace dsmc_host {

static thread_local std::string g_err;
int set_global_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
const char* global_error() { return g_err.c_str(); }

}  // namespace dsmc_host

using namespace dsmc_host;

// A second stream (default priority) on which the freestream injection runs
// beside the wall pass: both only add to the cell counts, injection writes
// the slots after the sorted ones and the wall pass deferred slots below
// them, so they may overlap; the main stream waits for the injection before
// the next step. Under graph capture the side stream joins the capture
// through the fork event.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  hipError_t create() {
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&join, hipEventDisableTiming);
    return e;
  }
  ~SideStream() {
    if (s) (void)hipStreamSynchronize(s);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (s) (void)hipStreamDestroy(s);
  }
};

struct dsmc_solver {
  std::string err;
  dsmc_settings s{};
  bool inited = false, have_state = false;
  hipStream_t stream = nullptr;
  SideStream side;
  InstDesc h{};
  InstDesc* d_desc = nullptr;
  std::vector<void*> bufs;      // lifetime of the handle
  std::vector<void*> geom_bufs; // replaced by set_geometry
  Mesh mesh;
  CGrid cg;
  bool have_geom = false;
  std::vector<float> frac;
  std::vector<uint8_t> nearc;  // PHYSICS near-body cell flags
  uint32_t par = 0;
  uint64_t steps = 0;
  uint32_t samples = 0;
  // simulation_state_t::simTime: FP32 sum of dt over the steps, as
  // advanceSimTime accumulates it (include/draw_helper.h:71)
  float sim_time = 0.0f;
  void advance_time(uint64_t n) {
    for (uint64_t k = 0; k < n; ++k) sim_time += h.dt;
  }
  // Expected particles per cell; tiles expecting at least 512 particles
  // are taken from the tile queues, smaller ones in a fixed sequence: their
  // phases are too short to hide the queue atomic, and it costs more than
  // the balance gains. The same flag has the scatter rank the particles
  // that stayed in their tile in LDS (with small tiles, the per-pass work
  // costs more than the global atomics it saves). No result depends on
  // either choice.
  double ppc_exp = 0.0;
  void set_tile_schedule() { h.tile_dyn = (double)h.tile_cells * ppc_exp >= 512.0 ? 1u : 0u; }
  bool sampling = false;
  uint64_t sample_steps0 = 0;
  double acc_snap[kAccSlots] = {0};
  uint32_t* scan_scratch = nullptr;
  uint32_t scatter_grid = 0;
  uint32_t ppc = 0;
  float ntc_init = 0.0f;
  bool timing = false;
  double t_ms[6] = {0, 0, 0, 0, 0, 0};  // summed ms: scan, scatter, tile, inject, whole step, walls
  uint64_t timed_steps = 0;
  std::vector<hipEvent_t> ev_pool;   // 5 events per timed step, resolved lazily
  size_t ev_used = 0;
  uint32_t debug_flags = 0;
  std::vector<uint32_t> rng_a, rng_b, rng_c;
  std::map<uint64_t, hipGraphExec_t> graphs;  // (parity, flags, flags) -> two captured steps
  bool use_graphs = true;

  int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
  int hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return DSMC_OK;
    return fail(DSMC_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
  }
  template <class T>
  int alloc(T** p, size_t count, bool geometry = false) {
    *p = nullptr;
    if (count == 0) count = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) return hip(e, "hipMalloc");
    (geometry ? geom_bufs : bufs).push_back(q);
    *p = (T*)q;
    return DSMC_OK;
  }
  void free_geom() {
    for (void* p : geom_bufs) (void)hipFree(p);
    geom_bufs.clear();
  }
  ~dsmc_solver() {
    if (stream) (void)hipStreamSynchronize(stream);
    free_geom();
    for (void* p : bufs) (void)hipFree(p);
    for (auto& e : ev_pool) (void)hipEventDestroy(e);
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    if (stream) (void)hipStreamDestroy(stream);
  }
  int upload_desc() { return hip(hipMemcpyAsync(d_desc, &h, sizeof(h), hipMemcpyHostToDevice, stream), "desc upload"); }
  uint32_t get_ctr(int i) {
    uint32_t v = 0;
    (void)hipMemcpyAsync(&v, h.ctr + i, 4, hipMemcpyDeviceToHost, stream);
    (void)hipStreamSynchronize(stream);
    return v;
  }
  bool physics() const { return s.mode == DSMC_MODE_PHYSICS; }
  // Sticky device status word (ctr[3]), read at synchronising calls.
  int check_status() {
    uint32_t st = get_ctr(3);
    if (st & kStatusCapacity)
      return fail(DSMC_ERR_CAPACITY, "particle capacity %u exceeded: injected particles were dropped", h.cap);
    if (st & kStatusPairs)
      return fail(DSMC_ERR_NUMERIC, "more than %u NTC candidate pairs in one cell (non-finite collision state)",
                  kMaxPairsPerCell);
    return DSMC_OK;
  }
};

#define CHECK_RC(expr)          \
  do {                          \
    int rc__ = (expr);          \
    if (rc__ != DSMC_OK) return rc__; \
  } while (0)
#define HIPC(expr) CHECK_RC(h->hip((expr), #expr))

namespace {

const double kPi = 3.14159265358979323846;

// Gamma function for the VHS constant (Lanczos approximation, g = 7).
double gamma_fn(double x) {
  static const double c[9] = {0.99999999999980993, 676.5203681218851, -1259.1392167224028,
                              771.32342877765313, -176.61502916214059, 12.507343278686905,
                              -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7};
  if (x < 0.5) return kPi / (std::sin(kPi * x) * gamma_fn(1.0 - x));
  x -= 1.0;
  double a = c[0];
  double t = x + 7.5;
  for (int i = 1; i < 9; ++i) a += c[i] / (x + i);
  return std::sqrt(2.0 * kPi) * std::pow(t, x + 0.5) * std::exp(-t) * a;
}

// Bird's freestream number flux through a face per unit area (1994, eq. 4.22):
// n * c_m / (2 sqrt(pi)) * [exp(-s^2) + sqrt(pi) s (1 + erf(s))], s = inward
// speed ratio.
double inflow_flux(double n, double vm, double un) {
  double s = un / vm;
  return n * vm / (2.0 * std::sqrt(kPi)) * (std::exp(-s * s) + std::sqrt(kPi) * s * (1.0 + std::erf(s)));
}

uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1u) / m * m; }

int step_once(dsmc_solver* h);
uint32_t step_flags(const dsmc_solver* h, uint64_t stepno, bool* sample);
int run_graph_pair(dsmc_solver* h, uint32_t f1, uint32_t f2);

}  // namespace

extern "C" {

const char* dsmc_version(void) { return "dsmc-mi355x 0.1 (gfx950)"; }
const char* dsmc_global_last_error(void) { return global_error(); }

int dsmc_create(int solver_type, dsmc_solver** out) {
  if (!out) return set_global_error(DSMC_ERR_INVALID, "null output handle");
  *out = nullptr;
  if (solver_type == DSMC_SOLVER_CPU)
    return set_global_error(DSMC_ERR_UNSUPPORTED, "no CPU solver (the reference's CreateSolver(SolverCPU) returns none)");
  if (solver_type == DSMC_SOLVER_GPU_VBO)
    return set_global_error(DSMC_ERR_UNSUPPORTED, "the OpenGL VBO solver is not part of this engine");
  if (solver_type != DSMC_SOLVER_GPU) return set_global_error(DSMC_ERR_INVALID, "unknown solver type");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return set_global_error(DSMC_ERR_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
  auto* h = new dsmc_solver();
  e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = h->side.create();
  if (e != hipSuccess) {
    delete h;
    return set_global_error(DSMC_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *out = h;
  return DSMC_OK;
}

void dsmc_destroy(dsmc_solver* h) { delete h; }

const char* dsmc_last_error(const dsmc_solver* h) { return h ? h->err.c_str() : global_error(); }

static int physics_fnum(dsmc_solver* h);
static int setup_injection(dsmc_solver* h);
static int sum_acc(dsmc_solver* h, double out[kAccSlots]);

int dsmc_init(dsmc_solver* h, const dsmc_settings* sp) {
  if (!h || !sp) return set_global_error(DSMC_ERR_INVALID, "null argument");
  if (h->inited) return h->fail(DSMC_ERR_STATE, "solver already initialised");
  dsmc_settings s = *sp;
  for (int a = 0; a < 3; ++a)
    if (s.grid_dim[a] <= 0 || !(s.bmax[a] > s.bmin[a]))
      return h->fail(DSMC_ERR_INVALID, "grid dimensions and bounds must be positive");
  if (!(s.dt > 0.0f)) return h->fail(DSMC_ERR_INVALID, "dt must be positive");
  if (!(s.molmass > 0.0f) || !(s.diameter > 0.0f)) return h->fail(DSMC_ERR_INVALID, "bad gas properties");
  for (int a = 0; a < 3; ++a) s.cell_size[a] = (s.bmax[a] - s.bmin[a]) / (float)s.grid_dim[a];
  const uint64_t C64 = (uint64_t)s.grid_dim[0] * s.grid_dim[1] * s.grid_dim[2];
  if (C64 > 16000000ull) return h->fail(DSMC_ERR_CAPACITY, "too many cells (%llu)", (unsigned long long)C64);
  const uint32_t C = (uint32_t)C64;
  const bool phys = s.mode == DSMC_MODE_PHYSICS;
  InstDesc& I = h->h;
  std::memset(&I, 0, sizeof(I));
  for (int a = 0; a < 3; ++a) {
    I.gmin[a] = s.bmin[a];
    I.gmax[a] = s.bmax[a];
    I.cell[a] = s.cell_size[a];
    I.inv_cell[a] = 1.0f / s.cell_size[a];
    I.dim[a] = (uint32_t)s.grid_dim[a];
  }
  I.ncells = C;
  I.sample_on = 1u;
  I.tile_cells = kTileCells;
  I.ntiles = (C + kTileCells - 1) / kTileCells;
  I.cell_volume = s.cell_size[0] * s.cell_size[1] * s.cell_size[2];
  I.mass = s.molmass;
  I.dt = s.dt;
  I.mode = phys ? kModePhysics : kModeReference;
  I.periodic = (uint32_t)(unsigned char)s.periodic_condition;
  for (int f = 0; f < 6; ++f) I.face_bc[f] = (uint32_t)s.face_bc[f];
  I.wall_temp = s.wall_temp;
  I.wall_key[0] = s.seed;
  I.wall_key[1] = 0x6A09E667u;
  I.vm_wall = s.wall_temp > 0.0f ? (float)std::sqrt(2.0 * kBoltz * s.wall_temp / s.molmass) : 0.0f;
  // InitSimulationProperties (dsmc_kernel.cu:309-348), FP32 as written there
  I.sigma_hs = kPiRef * s.diameter * s.diameter;
  {
    float d = s.diameter;
    h->ntc_init = d * d * kPiRef * sqrtf((float)(s.temp / 300.0)) * 300.0f;
  }
  if (phys) {
    // VHS: sigma c_r = pi d^2 (2 k Tref / (m_r c_r^2))^(w - 1/2) / Gamma(5/2 - w) * c_r
    double w = s.omega, mr = 0.5 * s.molmass;
    double k = kPi * (double)s.diameter * s.diameter * std::pow(2.0 * kBoltz * s.Tref / mr, w - 0.5) /
               gamma_fn(2.5 - w);
    I.vhs_k = (float)k;
    I.vhs_exp = (float)(2.0 - 2.0 * w);
    if (w == 0.5) {
      I.vhs_k = (float)(kPi * (double)s.diameter * s.diameter);
      I.vhs_exp = 1.0f;
    }
    I.n_inf = s.density;
    I.T_inf = s.temp;
    for (int a = 0; a < 3; ++a) I.U[a] = s.stream_velocity[a];
    I.vm_inf = (float)std::sqrt(2.0 * kBoltz * s.temp / s.molmass);
    // initial (sigma c_r)max from a generous relative speed
    double umag = std::sqrt((double)I.U[0] * I.U[0] + (double)I.U[1] * I.U[1] + (double)I.U[2] * I.U[2]);
    double cr0 = 4.0 * I.vm_inf + umag;
    h->ntc_init = (float)(I.vhs_exp == 1.0f ? I.vhs_k * cr0 : I.vhs_k * std::pow(cr0, (double)I.vhs_exp));
    I.fnum = s.fnum;
  } else {
    if (s.part_per_cell == 0 || s.num_particles == 0) {
      int rc = dsmc_settings_derive_particles(&s);
      if (rc) return h->fail(rc, "%s", global_error());
    }
    h->ppc = s.part_per_cell;
    I.avg_conc = (float)s.part_per_cell;
    I.fnum = s.fnum;
  }
  // capacity and RNG streams
  uint32_t cap;
  if (phys) {
    double vol = (double)(s.bmax[0] - s.bmin[0]) * (s.bmax[1] - s.bmin[1]) * (s.bmax[2] - s.bmin[2]);
    double n_exp = s.target_particles ? (double)s.target_particles : (double)s.density * vol / s.fnum;
    cap = s.capacity ? s.capacity : (uint32_t)std::min(4.0e9, 1.6 * n_exp + 65536.0);
  } else {
    cap = s.capacity ? std::max(s.capacity, s.num_particles) : std::max<uint32_t>(s.num_particles, 1u);
  }
  // the kernels address particle arrays by 32-bit byte offsets
  if (cap > kMaxCap) return h->fail(DSMC_ERR_CAPACITY, "capacity %u above %u particles", cap, kMaxCap);
  // Cells per tile: keep the expected particles of a tile near half the LDS
  // capacity so that few tiles take the global-memory path; on small grids,
  // halve the tiles (down to 64 cells) until there are 4096 of them, about
  // four per resident workgroup (c1, 64x32x32, 100k particles: 0.123 ms per
  // step with 64-cell tiles, 0.132 with 256, 0.149 with 32; c2, 128x64x64,
  // 1M particles: 0.155 ms with 128-cell tiles, 0.171 with 256, 0.182 with
  // 64), and to 32 cells where such a tile still expects 512 particles (the
  // reference's 32x32x32 benchmark case with 17 per cell: 0.135 ms per step
  // with 32-cell tiles, 0.143 with 64).
  {
    double ppc_exp = phys ? (s.target_particles ? (double)s.target_particles / C : (double)s.density *
                             I.cell_volume / (double)s.fnum)
                          : (double)h->ppc;
    uint32_t tc = kTileCells;
    while (tc > 32u && tc * ppc_exp > 0.5 * kLdsCap) tc >>= 1;
    while (tc > 64u && (C + tc - 1) / tc < 4096u) tc >>= 1;
    if (tc == 64u && (C + tc - 1) / tc < 4096u && 32.0 * ppc_exp >= 512.0) tc = 32u;
    I.tile_cells = tc;
    I.ntiles = (C + tc - 1) / tc;
    h->ppc_exp = ppc_exp;
    h->set_tile_schedule();
  }
  // RNG streams = 256 x the tile kernel's workgroups. By default one tile
  // per workgroup up to 1024 workgroups (small grids fill the GPU), else
  // about eight tiles per workgroup, at most 4096 workgroups: enough
  // workgroups to balance tiles of uneven cost (walls, dense cells) over the
  // CUs while the per-step stream state traffic (80 B per stream, read and
  // written) stays small next to the particle traffic.
  uint32_t ns = s.n_streams;
  if (!ns) {
    uint32_t wg = std::max(std::min(I.ntiles, 1024u), (I.ntiles + 7u) / 8u);
    ns = std::min(wg, 4096u) * kTileThreads;
  }
  ns = round_up(ns, kTileThreads);
  if (!phys) {
    // REFERENCE mode keeps the reference's stream layout: loading thread
    // X*Y*i + y*X + x (dsmc_kernel.cu:633) and, with -mt, cell c in the
    // collisions (:1030-1031) index the stream table directly. The reference
    // allocates MT_RNG_COUNT = 65536 streams (MersenneTwister.h:31) and
    // overruns them beyond that, so larger cases have no reference result.
    const uint64_t load_threads = (uint64_t)s.grid_dim[0] * s.grid_dim[1] * h->ppc;
    if (load_threads > kRefMtStreams || (s.mt && C > kRefMtStreams))
      return h->fail(DSMC_ERR_INVALID,
                     "REFERENCE mode: %llu loading threads and %u cells exceed the reference's %u MT streams",
                     (unsigned long long)load_threads, C, kRefMtStreams);
    ns = std::max(ns, round_up(std::max(C, (uint32_t)load_threads), kTileThreads));
  }
  I.cap = cap;
  I.n_streams = ns;
  I.n_vwg = ns / kTileThreads;
  h->s = s;

  // ---- device buffers
  for (int b = 0; b < 2; ++b) {
    CHECK_RC(h->alloc(&I.px[b], cap));
    CHECK_RC(h->alloc(&I.py[b], cap));
    CHECK_RC(h->alloc(&I.pz[b], cap));
    CHECK_RC(h->alloc(&I.vx[b], cap));
    CHECK_RC(h->alloc(&I.vy[b], cap));
    CHECK_RC(h->alloc(&I.vz[b], cap));
    CHECK_RC(h->alloc(&I.key[b], cap));
  }
  CHECK_RC(h->alloc(&I.perm, cap));
  CHECK_RC(h->alloc(&I.cursor, C));
  CHECK_RC(h->alloc(&I.local_cnt, C));
  CHECK_RC(h->alloc(&I.perm_tmp, cap));
  CHECK_RC(h->alloc(&I.owner_tmp, cap));
  CHECK_RC(h->alloc(&I.hits, cap));
  CHECK_RC(h->alloc(&I.count, C));
  CHECK_RC(h->alloc(&I.cell_start, (size_t)C + 1));
  CHECK_RC(h->alloc(&I.ntc, C));
  CHECK_RC(h->alloc(&I.cstat, C));
  CHECK_RC(h->alloc(&I.samp, (size_t)5 * C));
  CHECK_RC(h->alloc(&I.ctr, 16));
  CHECK_RC(h->alloc(&I.wg_acc, (size_t)I.n_vwg * kAccSlots));
  CHECK_RC(h->alloc(&I.mt_state, (size_t)kMtWords * ns));
  uint32_t *ma, *mb, *mc, *seeds;
  CHECK_RC(h->alloc(&ma, ns));
  CHECK_RC(h->alloc(&mb, ns));
  CHECK_RC(h->alloc(&mc, ns));
  CHECK_RC(h->alloc(&seeds, ns));
  I.mt_a = ma;
  I.mt_b = mb;
  I.mt_c = mc;
  CHECK_RC(h->alloc(&h->scan_scratch, dsmc_launch::kScanMaxTiles + 1));
  CHECK_RC(h->alloc(&h->d_desc, 1));
  hipStream_t st = h->stream;
  HIPC(hipMemsetAsync(I.count, 0, (size_t)C * 4, st));
  HIPC(hipMemsetAsync(I.cell_start, 0, ((size_t)C + 1) * 4, st));
  HIPC(hipMemsetAsync(I.cstat, 0, (size_t)C * sizeof(float4), st));
  HIPC(hipMemsetAsync(I.samp, 0, (size_t)5 * C * 8, st));
  HIPC(hipMemsetAsync(I.ctr, 0, 64, st));
  HIPC(hipMemsetAsync(I.wg_acc, 0, (size_t)I.n_vwg * kAccSlots * 8, st));
  HIPC(hipMemsetAsync(I.hits, 0, (size_t)cap * 4, st));
  // ---- RNG parameters and seeds
  std::vector<uint32_t> ra, rb, rc;
  if (s.rng_param_file[0]) {
    std::string e;
    if (!load_mt_rows(s.rng_param_file, ra, rb, rc, e)) return h->fail(DSMC_ERR_IO, "%s", e.c_str());
  } else {
    uint32_t rows = 0;
    const uint32_t* t = builtin_mt_table(&rows);
    ra.assign(t, t + rows);
    rb.assign(rows, 0x9d2c5680u);
    rc.assign(rows, 0xefc60000u);
  }
  std::vector<uint32_t> ha(ns), hb(ns), hc(ns), hs(ns);
  for (uint32_t i = 0; i < ns; ++i) {
    ha[i] = ra[i % ra.size()];
    hb[i] = rb[i % rb.size()];
    hc[i] = rc[i % rc.size()];
  }
  if (!phys && s.seed == 0) {
    // InitGPUTwisters seed chain (mtwister.cu:41-64): seed_i = lcg^(i+1)(999)
    uint64_t seed = 999;
    for (uint32_t i = 0; i < ns; ++i) {
      seed = (seed * 279470273ull) % 4294967291ull;
      hs[i] = (uint32_t)seed;
    }
  } else {
    for (uint32_t i = 0; i < ns; ++i) hs[i] = s.seed + i + 1u;
  }
  h->rng_a = ha;
  h->rng_b = hb;
  h->rng_c = hc;
  HIPC(hipMemcpyAsync(ma, ha.data(), (size_t)ns * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(mb, hb.data(), (size_t)ns * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(mc, hc.data(), (size_t)ns * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(seeds, hs.data(), (size_t)ns * 4, hipMemcpyHostToDevice, st));
  dsmc_launch::mt_seed(I.mt_state, seeds, ns, st);
  HIPC(hipGetLastError());
  // ---- NTC state: InitMaximums (dsmc_kernel.cu:1104-1120). The reference
  // draws the remainder from a zero MT state, which yields 1/2^32.
  std::vector<float2> nt(C);
  for (uint32_t c = 0; c < C; ++c) nt[c] = make_float2(h->ntc_init, phys ? 0.5f : 2.3283064365386963e-10f);
  HIPC(hipMemcpyAsync(I.ntc, nt.data(), (size_t)C * sizeof(float2), hipMemcpyHostToDevice, st));
  if (!phys && !s.mt) {
    // Hybrid-Taus collision streams, one per cell (ComputeBirdCollisionsNoMTDev).
    // CreateRandomSeedsUint4 (kernels/dsmc_kernel.cu:1846-1886) reads
    // /dev/urandom; here each word comes from splitmix64 of (seed, cell, k),
    // reduced the same way (x mod (2^32 - 129) + 128), so runs repeat.
    // dsmc_set_taus_state installs any other state.
    CHECK_RC(h->alloc(&I.taus, (size_t)4 * C));
    std::vector<uint32_t> ts((size_t)4 * C);
    for (uint32_t c = 0; c < C; ++c)
      for (uint32_t k = 0; k < 4; ++k) {
        uint64_t z = ((uint64_t)s.seed << 40) + (uint64_t)c * 4u + k + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        ts[(size_t)4 * c + k] = (uint32_t)((z & 0xFFFFFFFFull) % 4294967167ull) + 128u;
      }
    HIPC(hipMemcpyAsync(I.taus, ts.data(), ts.size() * 4, hipMemcpyHostToDevice, st));
  }
  HIPC(hipStreamSynchronize(st));
  // grid-stride scatter: 256 positions per workgroup and pass
  h->scatter_grid = std::min<uint32_t>((cap + 255) / 256, 4096u);
  if (phys) {
    // F_n and the inflow per face for the whole box; dsmc_set_geometry
    // redoes both for the fluid volume
    CHECK_RC(physics_fnum(h));
    CHECK_RC(setup_injection(h));
  }
  CHECK_RC(h->upload_desc());
  HIPC(hipStreamSynchronize(st));
  h->inited = true;
  return DSMC_OK;
}

int dsmc_init_from_config(dsmc_solver* h, const char* path) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  dsmc_settings s;
  dsmc_settings_default(&s);
  int rc = dsmc_settings_read_file(&s, path);
  if (rc) return h->fail(rc, "%s", global_error());
  if (s.mode != DSMC_MODE_PHYSICS) {
    rc = dsmc_settings_derive_particles(&s);
    if (rc) return h->fail(rc, "%s", global_error());
  }
  return dsmc_init(h, &s);
}

int dsmc_set_geometry(dsmc_solver* h, const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                      const float* normals, uint32_t gx, uint32_t gy, uint32_t gz) {
  return dsmc_set_geometry_moved(h, verts, nv, tris, nt, normals, gx, gy, gz, nullptr);
}

int dsmc_set_geometry_moved(dsmc_solver* h, const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                            const float* normals, uint32_t gx, uint32_t gy, uint32_t gz, const float move[3]) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->inited) return h->fail(DSMC_ERR_STATE, "set_geometry before init");
  if (h->have_state) return h->fail(DSMC_ERR_STATE, "set_geometry after the particles were loaded");
  if (!verts || !tris || nv == 0 || nt == 0 || gx == 0 || gy == 0 || gz == 0)
    return h->fail(DSMC_ERR_INVALID, "empty mesh or grid");
  for (uint64_t i = 0; i < 3ull * nt; ++i)
    if (tris[i] >= nv) return h->fail(DSMC_ERR_INVALID, "triangle index out of range");
  Mesh& m = h->mesh;
  m.v.assign(verts, verts + 3ull * nv);
  m.t.assign(tris, tris + 3ull * nt);
  if (normals) m.n.assign(normals, normals + 3ull * nt);
  else m.compute_normals();
  if (h->physics()) build_cgrid_physics(m, gx, gy, gz, h->cg);
  else build_cgrid_reference(m, gx, gy, gz, h->cg);
  CGrid& g = h->cg;
  if (move) {
    // Model::MoveModel after CreateGrid (src/model.cpp:709-731): vertices and
    // grid bounds shift, the cell size and the triangle lists stay
    for (size_t i = 0; i < m.v.size(); ++i) m.v[i] += move[i % 3];
    for (int k = 0; k < 3; ++k) {
      g.mn[k] += move[k];
      g.mx[k] += move[k];
    }
  }
  h->free_geom();
  InstDesc& I = h->h;
  uint32_t* flag;
  uint32_t *cs, *ce, *ct;
  float4 *t0, *t1, *t2, *tn;
  CHECK_RC(h->alloc(&flag, g.ncells(), true));
  CHECK_RC(h->alloc(&cs, g.ncells(), true));
  CHECK_RC(h->alloc(&ce, g.ncells(), true));
  CHECK_RC(h->alloc(&ct, std::max<size_t>(g.tri.size(), 1), true));
  CHECK_RC(h->alloc(&t0, nt, true));
  CHECK_RC(h->alloc(&t1, nt, true));
  CHECK_RC(h->alloc(&t2, nt, true));
  CHECK_RC(h->alloc(&tn, nt, true));
  std::vector<float4> h0(nt), h1(nt), h2(nt), hn(nt);
  for (uint32_t i = 0; i < nt; ++i) {
    const float* a = &m.v[3 * m.t[3 * i]];
    const float* b = &m.v[3 * m.t[3 * i + 1]];
    const float* c = &m.v[3 * m.t[3 * i + 2]];
    h0[i] = make_float4(a[0], a[1], a[2], 0.0f);
    h1[i] = make_float4(b[0], b[1], b[2], 0.0f);
    h2[i] = make_float4(c[0], c[1], c[2], 0.0f);
    hn[i] = make_float4(m.n[3 * i], m.n[3 * i + 1], m.n[3 * i + 2], 0.0f);
  }
  hipStream_t st = h->stream;
  HIPC(hipMemcpyAsync(flag, g.flag.data(), g.flag.size() * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(cs, g.start.data(), g.start.size() * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(ce, g.end.data(), g.end.size() * 4, hipMemcpyHostToDevice, st));
  if (!g.tri.empty()) HIPC(hipMemcpyAsync(ct, g.tri.data(), g.tri.size() * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(t0, h0.data(), nt * sizeof(float4), hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(t1, h1.data(), nt * sizeof(float4), hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(t2, h2.data(), nt * sizeof(float4), hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(tn, hn.data(), nt * sizeof(float4), hipMemcpyHostToDevice, st));
  I.cg_rec = nullptr;
  I.cg_se = nullptr;
  if (h->physics()) {
    const size_t ne = std::max<size_t>(g.tri.size(), 1);
    float4* rec;
    CHECK_RC(h->alloc(&rec, 4 * ne, true));
    std::vector<float4> hr(4 * ne, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (size_t i = 0; i < g.tri.size(); ++i) {
      const uint32_t t = g.tri[i];
      const float4 a = h0[t], b = h1[t], c = h2[t];
      uint32_t bits = t;
      float tf;
      std::memcpy(&tf, &bits, 4);
      hr[4 * i] = make_float4(a.x, a.y, a.z, tf);
      hr[4 * i + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
      hr[4 * i + 2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
      hr[4 * i + 3] = hn[t];
    }
    HIPC(hipMemcpyAsync(rec, hr.data(), hr.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    // entry ranges of the surface cells, one 8-byte load per cell visited
    uint2* se;
    CHECK_RC(h->alloc(&se, g.ncells(), true));
    std::vector<uint2> hse(g.ncells(), make_uint2(0u, 0u));
    for (size_t c = 0; c < g.ncells(); ++c)
      if (g.flag[c] == kCgSurface) hse[c] = make_uint2(g.start[c], g.end[c]);
    HIPC(hipMemcpyAsync(se, hse.data(), hse.size() * sizeof(uint2), hipMemcpyHostToDevice, st));
    HIPC(hipStreamSynchronize(st));
    I.cg_rec = rec;
    I.cg_se = se;
  }
  for (int a = 0; a < 3; ++a) {
    I.cg_min[a] = g.mn[a];
    I.cg_max[a] = g.mx[a];
    I.cg_cell[a] = g.cell[a];
    I.cg_dim[a] = g.dim[a];
  }
  I.cg_ncells = g.ncells();
  I.cg_flag = flag;
  I.cg_start = cs;
  I.cg_end = ce;
  I.cg_tri = ct;
  I.tv0 = t0;
  I.tv1 = t1;
  I.tv2 = t2;
  I.tn = tn;
  I.has_geom = 1u;
  h->have_geom = true;
  if (h->physics()) {
    fluid_fractions(m, g, I.gmin, I.cell, I.dim, h->frac);
    float* df;
    CHECK_RC(h->alloc(&df, I.ncells, true));
    HIPC(hipMemcpyAsync(df, h->frac.data(), (size_t)I.ncells * 4, hipMemcpyHostToDevice, st));
    I.fluid_frac = df;
    near_cells(g, I.gmin, I.cell, I.dim, h->nearc);
    I.near_unit = std::min(I.cell[0], std::min(I.cell[1], I.cell[2])) / 64.0f;
    uint8_t* dn;
    CHECK_RC(h->alloc(&dn, I.ncells, true));
    HIPC(hipMemcpyAsync(dn, h->nearc.data(), (size_t)I.ncells, hipMemcpyHostToDevice, st));
    I.near_cell = dn;
    // F_n and the inflow per face depend on the fluid volume
    CHECK_RC(physics_fnum(h));
    CHECK_RC(setup_injection(h));
  }
  CHECK_RC(h->upload_desc());
  HIPC(hipStreamSynchronize(st));
  return DSMC_OK;
}

// Tiles in locality order: sorted by (z / B, y / B, z mod B, y mod B, x) of
// their first cell, so that consecutive positions of the order cover B x B
// patches of cell rows. B is the largest power of two for which a patch
// holds at most 128 tiles, about what one XCD keeps resident (32 CUs x 4
// workgroups): the tiles a particle can come from in one step then mostly
// run together on that XCD.
static std::vector<uint32_t> tile_locality_order(uint32_t X, uint32_t Y, uint32_t Z, uint32_t tc, uint32_t ntiles) {
  (void)Z;
  uint32_t B = 1;
  while ((uint64_t)(2 * B) * (2 * B) * X <= 128ull * tc && 2 * B <= std::max(Y, Z)) B *= 2;
  std::vector<std::pair<uint64_t, uint32_t>> k(ntiles);
  for (uint32_t t = 0; t < ntiles; ++t) {
    const uint64_t c = (uint64_t)t * tc;
    const uint64_t x = c % X, y = (c / X) % Y, z = c / ((uint64_t)X * Y);
    const uint64_t key = ((((z / B) * ((Y + B - 1) / B) + y / B) * B + z % B) * B + y % B) * X + x;
    k[t] = {key, t};
  }
  std::sort(k.begin(), k.end());
  std::vector<uint32_t> order(ntiles);
  for (uint32_t t = 0; t < ntiles; ++t) order[t] = k[t].second;
  return order;
}

namespace {
uint32_t wall_grid(const InstDesc& I);
}  // namespace

static int finish_loading(dsmc_solver* h, uint32_t n, uint32_t* n_out) {
  InstDesc& I = h->h;
  hipStream_t st = h->stream;
  if (!I.tile_order) {
    // the tile count is fixed from here on
    std::vector<uint32_t> order = tile_locality_order(I.dim[0], I.dim[1], I.dim[2], I.tile_cells, I.ntiles);
    uint32_t* d = nullptr;
    CHECK_RC(h->alloc(&d, I.ntiles));
    HIPC(hipMemcpyAsync(d, order.data(), (size_t)I.ntiles * 4, hipMemcpyHostToDevice, st));
    HIPC(hipStreamSynchronize(st));
    I.tile_order = d;
    CHECK_RC(h->alloc(&I.tile_p0, (size_t)I.ntiles + 1));
    CHECK_RC(h->upload_desc());
  }
  if (h->physics() && I.has_geom && !I.defer_bits) {
    // deferred-slot bitmask and the wall pass's force sums; the pass's grid
    // is fixed here so that its per-workgroup sums are deterministic
    const size_t words = (size_t)I.cap / 64 + 2;
    CHECK_RC(h->alloc(&I.defer_bits, words));
    HIPC(hipMemsetAsync(I.defer_bits, 0, words * 8, st));
    CHECK_RC(h->alloc(&I.defer_list, I.cap));
    I.n_wall_wg = wall_grid(I);
    CHECK_RC(h->alloc(&I.wall_part, (size_t)I.n_wall_wg * 4));
    HIPC(hipMemsetAsync(I.wall_part, 0, (size_t)I.n_wall_wg * 4 * sizeof(long long), st));
    CHECK_RC(h->alloc(&I.wall_acc, 4));
    HIPC(hipMemsetAsync(I.wall_acc, 0, 4 * sizeof(double), st));
    // fixed-point units of the wall sums, from the largest of the speeds
    // that set the scale of the flow
    const double umag = std::sqrt((double)I.U[0] * I.U[0] + (double)I.U[1] * I.U[1] + (double)I.U[2] * I.U[2]);
    const double vref = std::max({umag, (double)I.vm_inf, (double)I.vm_wall, 1.0});
    // (a term is at most ~2^40 (momentum) or 2^43 (energy) for speeds up to
    // 20 v_ref, so 2^20 wall hits per step stay below 2^63)
    I.wall_scale[0] = std::ldexp(1.0, 36) / vref;
    I.wall_scale[1] = std::ldexp(1.0, 33) / (vref * vref);
    I.wall_unit[0] = (double)I.mass * vref / std::ldexp(1.0, 36);
    I.wall_unit[1] = (double)I.mass * vref * vref / std::ldexp(1.0, 34);
    CHECK_RC(h->upload_desc());
  }
  uint32_t ctr[2] = {n, n};
  HIPC(hipMemcpyAsync(I.ctr, ctr, 8, hipMemcpyHostToDevice, st));
  HIPC(hipMemsetAsync(I.count, 0, (size_t)I.ncells * 4, st));
  // no tile layout yet: the first scatter places every particle through the cursors
  HIPC(hipMemsetAsync(I.local_cnt, 0, (size_t)I.ncells * 4, st));
  HIPC(hipMemsetAsync(I.tile_p0, 0, ((size_t)I.ntiles + 1) * 4, st));
  dsmc_launch::keys_count(h->d_desc, h->par, n, st);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(st));
  h->have_state = true;
  if (n_out) *n_out = n;
  return DSMC_OK;
}

// Freestream injection set-up: face cells of the open faces and their
// expected inflow per step.
static int setup_injection(dsmc_solver* h) {
  InstDesc& I = h->h;
  uint32_t total = 0;
  I.face_off[0] = 0;
  for (int f = 0; f < 6; ++f) {
    uint32_t axis = f / 2, ta = axis == 0 ? 1 : 0, tb = axis == 2 ? 1 : 2;
    uint32_t n = (I.face_bc[f] == kFaceFreestream) ? I.dim[ta] * I.dim[tb] : 0u;
    if (n) {
      double sign = (f & 1) ? -1.0 : 1.0;
      double flux = inflow_flux(I.n_inf, I.vm_inf, sign * I.U[axis]);
      double area = (double)I.cell[ta] * I.cell[tb];
      I.inj_lambda[f] = (float)(flux * area * I.dt / I.fnum);
    } else {
      I.inj_lambda[f] = 0.0f;
    }
    total += n;
    I.face_off[f + 1] = total;
  }
  I.n_face_cells = total;
  if (total && !I.face_count) CHECK_RC(h->alloc(&I.face_count, (size_t)total + 1));
  return DSMC_OK;
}

static int physics_fnum(dsmc_solver* h) {
  InstDesc& I = h->h;
  double vfl = 0.0;
  for (uint32_t c = 0; c < I.ncells; ++c) vfl += h->frac.empty() ? 1.0 : h->frac[c];
  vfl *= I.cell_volume;
  if (h->s.target_particles) I.fnum = (float)((double)I.n_inf * vfl / h->s.target_particles);
  if (!(I.fnum > 0.0f)) return h->fail(DSMC_ERR_INVALID, "fnum must be positive");
  h->s.fnum = I.fnum;
  return DSMC_OK;
}

int dsmc_init_state(dsmc_solver* h, float ux, float uy, float uz, uint32_t* n_out) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->inited) return h->fail(DSMC_ERR_STATE, "init_state before init");
  if (h->have_state) return h->fail(DSMC_ERR_STATE, "state already loaded");
  InstDesc& I = h->h;
  hipStream_t st = h->stream;
  if (!h->physics()) {
    // GenerateInitialVelocities / GenInitialVelsUniform + move with dt = 0
    // (include/gpu_dsmc.inl:69-92)
    if (h->s.geometry && !h->have_geom) I.has_geom = 0u;
    const uint32_t ppc = h->ppc;
    const uint32_t n = ppc * I.ncells;
    if (n > I.cap) return h->fail(DSMC_ERR_CAPACITY, "capacity %u < %u particles", I.cap, n);
    CHECK_RC(h->upload_desc());
    dsmc_launch::load_reference(h->d_desc, h->par, ppc, I.dim[0] * I.dim[1] * ppc, ux, uy, uz, h->s.temp, st);
    dsmc_launch::move_zero(h->d_desc, h->par, n, st);
    HIPC(hipGetLastError());
    return finish_loading(h, n, n_out);
  }
  I.U[0] = ux;
  I.U[1] = uy;
  I.U[2] = uz;
  CHECK_RC(physics_fnum(h));
  CHECK_RC(setup_injection(h));
  CHECK_RC(h->upload_desc());
  dsmc_launch::load_count(h->d_desc, I.n_streams, nullptr, st);
  if (dsmc_launch::scan(I.count, I.ncells, I.cell_start, nullptr, true, h->scan_scratch, I.ctr + 1, st))
    return h->fail(DSMC_ERR_CAPACITY, "too many cells to scan");
  uint32_t n = h->get_ctr(1);
  if (n > I.cap) return h->fail(DSMC_ERR_CAPACITY, "capacity %u < %u particles", I.cap, n);
  dsmc_launch::load_generate(h->d_desc, h->par, I.n_streams, nullptr, nullptr, nullptr, nullptr, st);
  HIPC(hipGetLastError());
  return finish_loading(h, n, n_out);
}

int dsmc_init_state_from_fields(dsmc_solver* h, const float* fn, const float* fu, const float* fv, const float* fw,
                                const float* fT, uint32_t n_cells, uint32_t* n_out) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->inited) return h->fail(DSMC_ERR_STATE, "init_state before init");
  if (h->have_state) return h->fail(DSMC_ERR_STATE, "state already loaded");
  if (!h->physics()) return h->fail(DSMC_ERR_UNSUPPORTED, "field loading needs PHYSICS mode");
  InstDesc& I = h->h;
  if (n_cells != I.ncells || !fn || !fu || !fv || !fw || !fT) return h->fail(DSMC_ERR_INVALID, "field size mismatch");
  for (uint32_t c = 0; c < n_cells; ++c)
    if (!(fn[c] >= 0.0f) || !(fT[c] > 0.0f) || !std::isfinite(fu[c]) || !std::isfinite(fv[c]) || !std::isfinite(fw[c]))
      return h->fail(DSMC_ERR_INVALID, "non-physical field value in cell %u", c);
  hipStream_t st = h->stream;
  CHECK_RC(physics_fnum(h));
  CHECK_RC(setup_injection(h));
  CHECK_RC(h->upload_desc());
  float* d[5];
  const float* src[5] = {fn, fu, fv, fw, fT};
  for (int k = 0; k < 5; ++k) {
    CHECK_RC(h->alloc(&d[k], n_cells));
    HIPC(hipMemcpyAsync(d[k], src[k], (size_t)n_cells * 4, hipMemcpyHostToDevice, st));
  }
  dsmc_launch::load_count(h->d_desc, I.n_streams, d[0], st);
  if (dsmc_launch::scan(I.count, I.ncells, I.cell_start, nullptr, true, h->scan_scratch, I.ctr + 1, st))
    return h->fail(DSMC_ERR_CAPACITY, "too many cells to scan");
  uint32_t n = h->get_ctr(1);
  if (n > I.cap) return h->fail(DSMC_ERR_CAPACITY, "capacity %u < %u particles", I.cap, n);
  dsmc_launch::load_generate(h->d_desc, h->par, I.n_streams, d[1], d[2], d[3], d[4], st);
  HIPC(hipGetLastError());
  return finish_loading(h, n, n_out);
}

int dsmc_step(dsmc_solver* h, float dt) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->have_state) return h->fail(DSMC_ERR_STATE, "step before init_state");
  if (dt != h->h.dt) {
    if (!(dt > 0.0f)) return h->fail(DSMC_ERR_INVALID, "dt must be positive");
    h->h.dt = dt;
    if (h->physics() && h->h.n_face_cells) {
      // inflow per step scales with dt
      for (int f = 0; f < 6; ++f) h->h.inj_lambda[f] = 0.0f;
      h->h.n_face_cells = 0;
      CHECK_RC(setup_injection(h));
    }
    CHECK_RC(h->upload_desc());
  }
  return step_once(h);
}

int dsmc_run(dsmc_solver* h, uint32_t steps, float dt) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->have_state) return h->fail(DSMC_ERR_STATE, "run before init_state");
  if (steps == 0) return DSMC_OK;
  CHECK_RC(dsmc_step(h, dt));  // applies a dt change, then eager
  uint32_t done = 1;
  // pairs of steps replayed from captured graphs (timing needs per-step events)
  while (h->use_graphs && !h->timing && steps - done >= 2) {
    bool s1, s2;
    uint32_t f1 = step_flags(h, h->steps + 1, &s1);
    uint32_t f2 = step_flags(h, h->steps + 2, &s2);
    CHECK_RC(run_graph_pair(h, f1, f2));
    h->steps += 2;
    h->advance_time(2);
    h->samples += (s1 ? 1u : 0u) + (s2 ? 1u : 0u);
    done += 2;
  }
  for (; done < steps; ++done) CHECK_RC(step_once(h));
  return DSMC_OK;
}

int dsmc_set_sampling(dsmc_solver* h, int enabled, int reset) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->inited) return h->fail(DSMC_ERR_STATE, "not initialised");
  h->sampling = enabled != 0;
  if (reset) {
    InstDesc& I = h->h;
    HIPC(hipMemsetAsync(I.samp, 0, (size_t)5 * I.ncells * 8, h->stream));
    h->samples = 0;
    h->sample_steps0 = h->steps;
    CHECK_RC(sum_acc(h, h->acc_snap));
  }
  return DSMC_OK;
}

int dsmc_sample(dsmc_solver* h) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->have_state) return h->fail(DSMC_ERR_STATE, "sample before init_state");
  dsmc_launch::sample_cells(h->d_desc, h->par, h->h.mode, h->h.ncells, h->stream);
  HIPC(hipGetLastError());
  h->samples += 1;
  return DSMC_OK;
}

int dsmc_get_cell_stats(dsmc_solver* h, dsmc_cell_stats* out, uint32_t n) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  if (!h->inited) return h->fail(DSMC_ERR_STATE, "not initialised");
  InstDesc& I = h->h;
  if (n < I.ncells) return h->fail(DSMC_ERR_INVALID, "need %u cells", I.ncells);
  const uint32_t C = I.ncells;
  std::vector<double> buf((size_t)5 * C);
  HIPC(hipMemcpyAsync(buf.data(), I.samp, buf.size() * 8, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  const float* f = reinterpret_cast<const float*>(buf.data());
  for (uint32_t c = 0; c < C; ++c) {
    if (h->physics()) {
      out[c].num_in_sample = buf[c];
      out[c].xyz_sum[0] = buf[C + c];
      out[c].xyz_sum[1] = buf[2 * (size_t)C + c];
      out[c].xyz_sum[2] = buf[3 * (size_t)C + c];
      out[c].vel_mag = buf[4 * (size_t)C + c];
    } else {
      out[c].num_in_sample = f[c];
      out[c].xyz_sum[0] = f[C + c];
      out[c].xyz_sum[1] = f[2 * (size_t)C + c];
      out[c].xyz_sum[2] = f[3 * (size_t)C + c];
      out[c].vel_mag = f[4 * (size_t)C + c];
    }
  }
  return DSMC_OK;
}

int dsmc_get_fields(dsmc_solver* h, float* n, float* u, float* v, float* w, float* T, uint32_t nc) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  InstDesc& I = h->h;
  if (nc < I.ncells) return h->fail(DSMC_ERR_INVALID, "need %u cells", I.ncells);
  std::vector<dsmc_cell_stats> cs(I.ncells);
  CHECK_RC(dsmc_get_cell_stats(h, cs.data(), I.ncells));
  const double S = h->samples > 0 ? (double)h->samples : 1.0;
  const double kb = h->physics() ? kBoltz : (double)kBoltzRef;
  for (uint32_t c = 0; c < I.ncells; ++c) {
    double N = cs[c].num_in_sample;
    double fr = (h->physics() && !h->frac.empty()) ? h->frac[c] : 1.0;
    double vol = (double)I.cell_volume * fr;
    double dens = vol > 0.0 ? (double)I.fnum * N / (vol * S) : 0.0;
    double ux = 0, uy = 0, uz = 0, tt = 0;
    if (N > 0) {
      ux = cs[c].xyz_sum[0] / N;
      uy = cs[c].xyz_sum[1] / N;
      uz = cs[c].xyz_sum[2] / N;
      tt = (double)I.mass * (cs[c].vel_mag / N - (ux * ux + uy * uy + uz * uz)) / (3.0 * kb);
    }
    if (n) n[c] = (float)dens;
    if (u) u[c] = (float)ux;
    if (v) v[c] = (float)uy;
    if (w) w[c] = (float)uz;
    if (T) T[c] = (float)tt;
  }
  return DSMC_OK;
}

static int sum_acc(dsmc_solver* h, double out[kAccSlots]) {
  InstDesc& I = h->h;
  std::vector<long long> acc((size_t)I.n_vwg * kAccSlots);
  HIPC(hipMemcpyAsync(acc.data(), I.wg_acc, acc.size() * 8, hipMemcpyDeviceToHost, h->stream));
  double tacc[4] = {0.0, 0.0, 0.0, 0.0};
  if (I.wall_acc) HIPC(hipMemcpyAsync(tacc, I.wall_acc, sizeof(tacc), hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  // the workgroups' integer counters summed exactly (128-bit), converted once
  for (int k = 0; k < kAccSlots; ++k) {
    __int128 t = 0;
    for (uint32_t w = 0; w < I.n_vwg; ++w) t += acc[(size_t)w * kAccSlots + k];
    out[k] = (double)t;
  }
  out[kAccCr] /= kCrScale;
  // wall forces and heat of the wall pass
  for (int k = 0; k < 4; ++k) out[kAccFx + k] += tacc[k];
  return DSMC_OK;
}

int dsmc_get_collision_stats(dsmc_solver* h, double out[4]) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  if (!h->inited) return h->fail(DSMC_ERR_STATE, "not initialised");
  InstDesc& I = h->h;
  if (!h->physics()) {
    // GetStatistics (dsmc_kernel.cu:840-854): sum of the per-cell counters
    std::vector<float4> cs(I.ncells);
    HIPC(hipMemcpyAsync(cs.data(), I.cstat, cs.size() * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    double a = 0, b = 0, c = 0;
    for (auto& x : cs) {
      a += x.x;
      b += x.y;
      c += x.z;
    }
    out[0] = a;
    out[1] = b;
    out[2] = c;
    out[3] = 0.0;
    return DSMC_OK;
  }
  double acc[kAccSlots];
  CHECK_RC(sum_acc(h, acc));
  out[0] = acc[kAccSel];
  out[1] = acc[kAccColl];
  out[2] = acc[kAccCr];
  out[3] = 0.0;
  return DSMC_OK;
}

int dsmc_get_drag(dsmc_solver* h, double out[4]) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  if (!h->inited) return h->fail(DSMC_ERR_STATE, "not initialised");
  for (int k = 0; k < 4; ++k) out[k] = 0.0;
  if (!h->physics()) return DSMC_OK;
  double acc[kAccSlots];
  CHECK_RC(sum_acc(h, acc));
  double tw = (double)(h->steps - h->sample_steps0) * h->h.dt;
  if (tw <= 0.0) return DSMC_OK;
  double f = (double)h->h.fnum / tw;
  out[0] = (acc[kAccFx] - h->acc_snap[kAccFx]) * f;
  out[1] = (acc[kAccFy] - h->acc_snap[kAccFy]) * f;
  out[2] = (acc[kAccFz] - h->acc_snap[kAccFz]) * f;
  out[3] = (acc[kAccHeat] - h->acc_snap[kAccHeat]) * f;
  return DSMC_OK;
}

namespace {

// ProcessAndOutputStatistics (kernels/dsmc_base.cu:104-163) from the current
// per-cell sums and the given simulation-state counters.
int format_statistics(dsmc_solver* h, const dsmc_sim_state& st, std::string& text) {
  double coll[4];
  CHECK_RC(dsmc_get_collision_stats(h, coll));
  InstDesc& I = h->h;
  std::vector<dsmc_cell_stats> cs(I.ncells);
  CHECK_RC(dsmc_get_cell_stats(h, cs.data(), I.ncells));
  std::ostringstream out;
  // GetStatistics hands over a float4 (dsmc_kernel.cu:840-854)
  const float sel = (float)coll[0], col = (float)coll[1], cr = (float)coll[2];
  out << "Selections: " << sel << "\nCollisions: " << col << "\nMean separation: " << cr / col
      << "\nCollision ratio: " << col / sel << "\n";
  out << "Simulation time: " << st.sim_time << " Running time: " << st.real_time << "\n";
  out << "Cell N  ||  SampNum  ||  Dens  ||          Velocity          ||  Temperature\n";
  double tot = 0, denn = 0;
  const float cell_volume = h->s.cell_size[0] * h->s.cell_size[1] * h->s.cell_size[2];
  const float kb = kBoltzRef;
  for (uint32_t i = 0; i < I.ncells; ++i) {
    const dsmc_cell_stats& c = cs[i];
    float a = h->s.fnum / (cell_volume * (float)st.samples_count);
    const float n = (float)c.num_in_sample;
    denn = a * n;
    float vx = (float)c.xyz_sum[0] / n, vy = (float)c.xyz_sum[1] / n, vz = (float)c.xyz_sum[2] / n;
    float uu = vx * vx + vy * vy + vz * vz;
    float tt = h->s.molmass * ((float)c.vel_mag / n - uu) / (3.0f * kb);
    tot += tt;
    out << i << "  ||  " << (uint32_t)n << "  ||  " << denn << "  ||  " << vx << "," << vy << "," << vz << "  ||  "
        << tt << "\n";
  }
  tot /= I.ncells;
  const float gv = (h->s.bmax[0] - h->s.bmin[0]) * (h->s.bmax[1] - h->s.bmin[1]) * (h->s.bmax[2] - h->s.bmin[2]);
  // the reference names these dens2 / diam2 but forms n*d twice; the product
  // n^2 d^2 is the hard-sphere collision-rate factor
  const double dens2 = (double)h->s.density * h->s.diameter, diam2 = (double)h->s.diameter * h->s.density;
  const double molmas = h->s.molmass;
  const float Tref = h->s.Tref;
  float tColl = (float)(2.0f * st.sim_time * dens2 * gv * kPiRef * diam2 * std::pow(tot / Tref, 0.25) *
                        sqrtf((float)(kb * Tref / (kPiRef * molmas))) / h->s.fnum);
  out << "\n\nExpected collisions: " << tColl << " Ratio: " << tColl / col << "\n";
  // the reference prints the last cell's density under "Average density"
  // (dsmc_base.cu:162); kept for output compatibility
  out << "Average temperature: " << tot << " Reference temperature is: " << h->s.temp
      << "\nAverage density: " << denn << " Reference density: " << h->s.density << "\n\n" << std::endl;
  text = out.str();
  return DSMC_OK;
}

dsmc_sim_state engine_state(const dsmc_solver* h) {
  dsmc_sim_state st{};
  st.sim_time = h->sim_time;
  st.real_time = 0.0f;  // the benchmark path never advances realTime (kernels/dsmc.cu:270 is interactive)
  st.step_count = (uint32_t)h->steps;
  st.samples_count = h->samples;
  return st;
}

}  // namespace

int dsmc_get_sim_state(dsmc_solver* h, dsmc_sim_state* out) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  *out = engine_state(h);
  return DSMC_OK;
}

int dsmc_sample_statistics(dsmc_solver* h, float coll_out[4], float* stats, uint32_t n_cells) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->have_state) return h->fail(DSMC_ERR_STATE, "no state");
  InstDesc& I = h->h;
  if (stats && n_cells < I.ncells) return h->fail(DSMC_ERR_INVALID, "need %u cells", I.ncells);
  // SampleStatistics (include/gpu_dsmc.inl:277-282): GetStatistics, then one
  // more SampleCellStatistics that the samples counter does not include
  double coll[4];
  CHECK_RC(dsmc_get_collision_stats(h, coll));
  if (coll_out)
    for (int k = 0; k < 4; ++k) coll_out[k] = (float)coll[k];
  dsmc_launch::sample_cells(h->d_desc, h->par, h->h.mode, h->h.ncells, h->stream);
  HIPC(hipGetLastError());
  if (stats) {
    std::vector<dsmc_cell_stats> cs(I.ncells);
    CHECK_RC(dsmc_get_cell_stats(h, cs.data(), I.ncells));
    for (uint32_t c = 0; c < I.ncells; ++c) {
      float* r = stats + (size_t)8 * c;  // statistics_t: n, xyzSum[3], velMag, pad (8 floats)
      r[0] = (float)cs[c].num_in_sample;
      r[1] = (float)cs[c].xyz_sum[0];
      r[2] = (float)cs[c].xyz_sum[1];
      r[3] = (float)cs[c].xyz_sum[2];
      r[4] = (float)cs[c].vel_mag;
      r[5] = r[6] = r[7] = 0.0f;
    }
  }
  return DSMC_OK;
}

int dsmc_format_statistics(dsmc_solver* h, const dsmc_sim_state* st, char* buf, uint64_t cap, uint64_t* len) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->have_state) return h->fail(DSMC_ERR_STATE, "no state");
  std::string text;
  CHECK_RC(format_statistics(h, st ? *st : engine_state(h), text));
  if (len) *len = text.size();
  if (!buf) return DSMC_OK;
  if (cap < text.size() + 1) return h->fail(DSMC_ERR_CAPACITY, "statistics need %zu bytes", text.size() + 1);
  std::memcpy(buf, text.c_str(), text.size() + 1);
  return DSMC_OK;
}

int dsmc_dump_statistics(dsmc_solver* h, const char* path) {
  if (!h || !path) return set_global_error(DSMC_ERR_INVALID, "null argument");
  if (!h->have_state) return h->fail(DSMC_ERR_STATE, "no state");
  // DumpStatistics = SampleStatistics + ProcessAndOutputStatistics
  // (kernels/dsmc_base.cu:165-174)
  CHECK_RC(dsmc_sample_statistics(h, nullptr, nullptr, 0));
  std::string text;
  CHECK_RC(format_statistics(h, engine_state(h), text));
  std::ofstream out(path);
  if (!out) return h->fail(DSMC_ERR_IO, "cannot write %s", path);
  out << text;
  return out ? DSMC_OK : h->fail(DSMC_ERR_IO, "cannot write %s", path);
}

int dsmc_get_counts(dsmc_solver* h, uint64_t* steps, uint32_t* n_particles, uint32_t* samples) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (steps) *steps = h->steps;
  if (samples) *samples = h->samples;
  if (n_particles) {
    *n_particles = 0;
    if (h->have_state) {
      // live particles = stored minus removed; the next scan's total equals it
      InstDesc& I = h->h;
      uint32_t stored = h->get_ctr(0);
      std::vector<uint32_t> keys(stored);
      if (stored)
        HIPC(hipMemcpyAsync(keys.data(), I.key[h->par], (size_t)stored * 4, hipMemcpyDeviceToHost, h->stream));
      HIPC(hipStreamSynchronize(h->stream));
      uint32_t live = 0;
      for (uint32_t k : keys) live += k != kInvalidKey;
      *n_particles = live;
    }
  }
  return DSMC_OK;
}

int dsmc_synchronize(dsmc_solver* h) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  HIPC(hipStreamSynchronize(h->stream));
  return h->inited ? h->check_status() : DSMC_OK;
}

int dsmc_validate(dsmc_solver* h, uint32_t* n_bad) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->have_state) return h->fail(DSMC_ERR_STATE, "no state");
  uint32_t* bad = h->h.ctr + 8;
  HIPC(hipMemsetAsync(bad, 0, 4, h->stream));
  uint32_t n = h->get_ctr(0);
  dsmc_launch::validate(h->d_desc, h->par, n, bad, h->stream);
  HIPC(hipGetLastError());
  uint32_t b = h->get_ctr(8);
  if (n_bad) *n_bad = b;
  CHECK_RC(h->check_status());
  return b ? h->fail(DSMC_ERR_NUMERIC, "%u particles with non-finite state or outside the box", b) : DSMC_OK;
}

int dsmc_get_particles(dsmc_solver* h, float* pos, float* vel, uint32_t* keys, uint32_t cap, uint32_t* n_out) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->have_state) return h->fail(DSMC_ERR_STATE, "no state");
  InstDesc& I = h->h;
  uint32_t n = h->get_ctr(0);
  if (n_out) *n_out = n;
  if (cap < n) return pos || vel || keys ? h->fail(DSMC_ERR_CAPACITY, "need room for %u particles", n) : DSMC_OK;
  hipStream_t st = h->stream;
  uint32_t p = h->par;
  if (pos) {
    HIPC(hipMemcpyAsync(pos, I.px[p], (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(pos + cap, I.py[p], (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(pos + 2 * (size_t)cap, I.pz[p], (size_t)n * 4, hipMemcpyDeviceToHost, st));
  }
  if (vel) {
    HIPC(hipMemcpyAsync(vel, I.vx[p], (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(vel + cap, I.vy[p], (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(vel + 2 * (size_t)cap, I.vz[p], (size_t)n * 4, hipMemcpyDeviceToHost, st));
  }
  if (keys) HIPC(hipMemcpyAsync(keys, I.key[p], (size_t)n * 4, hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  if (keys)
    for (uint32_t i = 0; i < n; ++i) keys[i] = key_cell(keys[i]);  // without the wall pass's mark
  return DSMC_OK;
}

int dsmc_set_particles(dsmc_solver* h, const float* pos, const float* vel, uint32_t n) {
  if (!h || !pos || !vel) return set_global_error(DSMC_ERR_INVALID, "null argument");
  if (!h->inited) return h->fail(DSMC_ERR_STATE, "not initialised");
  InstDesc& I = h->h;
  if (n > I.cap) return h->fail(DSMC_ERR_CAPACITY, "capacity %u < %u", I.cap, n);
  if (!I.has_geom && h->physics() == false && h->s.geometry && h->have_geom == false) I.has_geom = 0u;
  if (h->physics() && I.n_face_cells == 0 && !h->have_state) {
    CHECK_RC(physics_fnum(h));
    CHECK_RC(setup_injection(h));
  }
  CHECK_RC(h->upload_desc());
  hipStream_t st = h->stream;
  uint32_t p = h->par;
  HIPC(hipMemcpyAsync(I.px[p], pos, (size_t)n * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(I.py[p], pos + n, (size_t)n * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(I.pz[p], pos + 2 * (size_t)n, (size_t)n * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(I.vx[p], vel, (size_t)n * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(I.vy[p], vel + n, (size_t)n * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(I.vz[p], vel + 2 * (size_t)n, (size_t)n * 4, hipMemcpyHostToDevice, st));
  h->have_state = false;
  return finish_loading(h, n, nullptr);
}

int dsmc_get_cell_start(dsmc_solver* h, uint32_t* out, uint32_t n) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc& I = h->h;
  if (n < I.ncells + 1) return h->fail(DSMC_ERR_INVALID, "need %u entries", I.ncells + 1);
  HIPC(hipMemcpyAsync(out, I.cell_start, ((size_t)I.ncells + 1) * 4, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return DSMC_OK;
}

int dsmc_get_perm(dsmc_solver* h, uint32_t* out, uint32_t n) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc& I = h->h;
  uint32_t m = std::min(n, I.cap);
  HIPC(hipMemcpyAsync(out, I.perm, (size_t)m * 4, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return DSMC_OK;
}

int dsmc_get_hit_counts(dsmc_solver* h, uint32_t* out, uint32_t n) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc& I = h->h;
  uint32_t m = std::min(n, I.cap);
  HIPC(hipMemcpyAsync(out, I.hits, (size_t)m * 4, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return DSMC_OK;
}

int dsmc_get_ntc_state(dsmc_solver* h, float* out, uint32_t n) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc& I = h->h;
  if (n < I.ncells) return h->fail(DSMC_ERR_INVALID, "need %u cells", I.ncells);
  HIPC(hipMemcpyAsync(out, I.ntc, (size_t)I.ncells * 8, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return DSMC_OK;
}

int dsmc_get_rng_state(dsmc_solver* h, uint32_t* out, uint32_t n) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc& I = h->h;
  size_t need = (size_t)kMtWords * I.n_streams;
  if (n < need) return h->fail(DSMC_ERR_INVALID, "need %zu words", need);
  std::vector<uint32_t> tmp(need);
  HIPC(hipMemcpyAsync(tmp.data(), I.mt_state, need * 4, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  for (uint32_t s = 0; s < I.n_streams; ++s)
    for (uint32_t w = 0; w < kMtWords; ++w) out[(size_t)s * kMtWords + w] = tmp[(size_t)w * I.n_streams + s];
  return DSMC_OK;
}

int dsmc_set_rng_state(dsmc_solver* h, const uint32_t* in, uint32_t n) {
  if (!h || !in) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc& I = h->h;
  size_t need = (size_t)kMtWords * I.n_streams;
  if (n < need) return h->fail(DSMC_ERR_INVALID, "need %zu words", need);
  std::vector<uint32_t> tmp(need);
  for (uint32_t s = 0; s < I.n_streams; ++s)
    for (uint32_t w = 0; w < kMtWords; ++w) tmp[(size_t)w * I.n_streams + s] = in[(size_t)s * kMtWords + w];
  HIPC(hipMemcpyAsync(I.mt_state, tmp.data(), need * 4, hipMemcpyHostToDevice, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return DSMC_OK;
}

int dsmc_set_tile_cells(dsmc_solver* h, uint32_t tile_cells) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->inited) return h->fail(DSMC_ERR_STATE, "not initialised");
  if (h->have_state || h->h.tile_order)
    return h->fail(DSMC_ERR_STATE, "tile size must be set before the particles are loaded");
  if (tile_cells < 32u || tile_cells > kTileCells || (tile_cells & (tile_cells - 1u)))
    return h->fail(DSMC_ERR_INVALID, "tile_cells must be a power of two in [32, %u]", kTileCells);
  h->h.tile_cells = tile_cells;
  h->h.ntiles = (h->h.ncells + tile_cells - 1) / tile_cells;
  h->set_tile_schedule();
  return h->upload_desc();
}

int dsmc_get_taus_state(dsmc_solver* h, uint32_t* out, uint32_t n) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc& I = h->h;
  if (!I.taus) return h->fail(DSMC_ERR_STATE, "no Hybrid-Taus state (REFERENCE mode with mt = 0 only)");
  if (n < 4u * I.ncells) return h->fail(DSMC_ERR_INVALID, "need %u words", 4u * I.ncells);
  HIPC(hipMemcpyAsync(out, I.taus, (size_t)16 * I.ncells, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return DSMC_OK;
}

int dsmc_set_taus_state(dsmc_solver* h, const uint32_t* in, uint32_t n) {
  if (!h || !in) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc& I = h->h;
  if (!I.taus) return h->fail(DSMC_ERR_STATE, "no Hybrid-Taus state (REFERENCE mode with mt = 0 only)");
  if (n < 4u * I.ncells) return h->fail(DSMC_ERR_INVALID, "need %u words", 4u * I.ncells);
  HIPC(hipMemcpyAsync(I.taus, in, (size_t)16 * I.ncells, hipMemcpyHostToDevice, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return DSMC_OK;
}

int dsmc_get_near_cells(dsmc_solver* h, uint8_t* out, uint32_t n) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc& I = h->h;
  if (n < I.ncells) return h->fail(DSMC_ERR_INVALID, "need %u cells", I.ncells);
  for (uint32_t c = 0; c < I.ncells; ++c) out[c] = h->nearc.empty() ? 0 : h->nearc[c];
  return DSMC_OK;
}

int dsmc_get_fluid_fraction(dsmc_solver* h, float* out, uint32_t n) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc& I = h->h;
  if (n < I.ncells) return h->fail(DSMC_ERR_INVALID, "need %u cells", I.ncells);
  for (uint32_t c = 0; c < I.ncells; ++c) out[c] = h->frac.empty() ? 1.0f : h->frac[c];
  return DSMC_OK;
}

int dsmc_get_collision_grid(dsmc_solver* h, uint32_t* flags, uint32_t n) {
  if (!h || !flags) return set_global_error(DSMC_ERR_INVALID, "null argument");
  if (!h->have_geom) return h->fail(DSMC_ERR_STATE, "no geometry");
  if (n < h->cg.ncells()) return h->fail(DSMC_ERR_INVALID, "need %u cells", h->cg.ncells());
  std::copy(h->cg.flag.begin(), h->cg.flag.end(), flags);
  return DSMC_OK;
}

// Resolve the recorded per-step events into the kernel-class sums.
static int fold_timings(dsmc_solver* h) {
  if (h->ev_used == 0) return DSMC_OK;
  HIPC(hipEventSynchronize(h->ev_pool[h->ev_used - 1]));
  // events per step: start, after scan, after scatter, after tile, after
  // walls, after injection
  static const int slot_of[5] = {0, 1, 2, 5, 3};
  for (size_t g = 0; g + 6 <= h->ev_used; g += 6) {
    float t;
    for (int k = 0; k < 5; ++k) {
      HIPC(hipEventElapsedTime(&t, h->ev_pool[g + k], h->ev_pool[g + k + 1]));
      h->t_ms[slot_of[k]] += t;
    }
    HIPC(hipEventElapsedTime(&t, h->ev_pool[g], h->ev_pool[g + 5]));
    h->t_ms[4] += t;
    h->timed_steps += 1;
  }
  h->ev_used = 0;
  return DSMC_OK;
}

int dsmc_get_timings(dsmc_solver* h, float* out_ms, uint32_t n) {
  if (!h || !out_ms) return set_global_error(DSMC_ERR_INVALID, "null argument");
  CHECK_RC(fold_timings(h));
  double d = h->timed_steps ? (double)h->timed_steps : 1.0;
  for (uint32_t k = 0; k < n && k < 5; ++k) out_ms[k] = (float)(h->t_ms[k] / d);
  if (n > 5) out_ms[5] = (float)h->timed_steps;
  if (n > 6) out_ms[6] = (float)(h->t_ms[5] / d);
  return DSMC_OK;
}

int dsmc_set_timing(dsmc_solver* h, int enabled) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  CHECK_RC(fold_timings(h));
  h->timing = enabled != 0;
  for (auto& t : h->t_ms) t = 0.0;
  h->timed_steps = 0;
  return DSMC_OK;
}

int dsmc_set_scatter_grid(dsmc_solver* h, uint32_t blocks) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->inited) return h->fail(DSMC_ERR_STATE, "not initialised");
  if (blocks == 0 || blocks > 65535) return h->fail(DSMC_ERR_INVALID, "scatter grid must be in [1, 65535]");
  h->scatter_grid = blocks;
  for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.second);
  h->graphs.clear();
  return DSMC_OK;
}

int dsmc_set_graphs(dsmc_solver* h, int enabled) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  h->use_graphs = enabled != 0;
  return DSMC_OK;
}

int dsmc_set_debug(dsmc_solver* h, uint32_t flags) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  h->debug_flags = flags;
  return DSMC_OK;
}

int dsmc_get_layout(dsmc_solver* h, uint32_t* tile_cells, uint32_t* threads, uint32_t* n_streams, uint32_t* n_vwg) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (tile_cells) *tile_cells = h->h.tile_cells;
  if (threads) *threads = kTileThreads;
  if (n_streams) *n_streams = h->h.n_streams;
  if (n_vwg) *n_vwg = h->h.n_vwg;
  return DSMC_OK;
}

int dsmc_get_rng_params(dsmc_solver* h, uint32_t* a, uint32_t* b, uint32_t* c, uint32_t n) {
  if (!h || !a || !b || !c) return set_global_error(DSMC_ERR_INVALID, "null argument");
  if (!h->inited) return h->fail(DSMC_ERR_STATE, "not initialised");
  if (n < h->rng_a.size()) return h->fail(DSMC_ERR_INVALID, "need %zu streams", h->rng_a.size());
  std::copy(h->rng_a.begin(), h->rng_a.end(), a);
  std::copy(h->rng_b.begin(), h->rng_b.end(), b);
  std::copy(h->rng_c.begin(), h->rng_c.end(), c);
  return DSMC_OK;
}

int dsmc_get_derived(dsmc_solver* h, float* out, uint32_t n) {
  if (!h || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc& I = h->h;
  float v[12] = {I.fnum, I.vhs_k, I.vhs_exp, h->ntc_init, I.cell_volume, I.sigma_hs,
                 I.inj_lambda[0], I.inj_lambda[1], I.inj_lambda[2], I.inj_lambda[3], I.inj_lambda[4], I.inj_lambda[5]};
  for (uint32_t k = 0; k < n && k < 12; ++k) out[k] = v[k];
  return DSMC_OK;
}

int dsmc_get_collision_grid_full(dsmc_solver* h, float bounds[9], uint32_t dims[3], uint32_t* flags, uint32_t* start,
                                 uint32_t* end, uint32_t* entries, uint32_t cap_cells, uint32_t cap_entries,
                                 uint32_t* n_entries) {
  if (!h) return set_global_error(DSMC_ERR_INVALID, "null handle");
  if (!h->have_geom) return h->fail(DSMC_ERR_STATE, "no geometry");
  const CGrid& g = h->cg;
  if (bounds)
    for (int k = 0; k < 3; ++k) {
      bounds[k] = g.mn[k];
      bounds[3 + k] = g.mx[k];
      bounds[6 + k] = g.cell[k];
    }
  if (dims)
    for (int k = 0; k < 3; ++k) dims[k] = g.dim[k];
  if (n_entries) *n_entries = (uint32_t)g.tri.size();
  if ((flags || start || end) && cap_cells < g.ncells()) return h->fail(DSMC_ERR_CAPACITY, "cell buffers too small");
  if (entries && cap_entries < g.tri.size()) return h->fail(DSMC_ERR_CAPACITY, "entry buffer too small");
  if (flags) std::copy(g.flag.begin(), g.flag.end(), flags);
  if (start) std::copy(g.start.begin(), g.start.end(), start);
  if (end) std::copy(g.end.begin(), g.end.end(), end);
  if (entries) std::copy(g.tri.begin(), g.tri.end(), entries);
  return DSMC_OK;
}

// ------------------------------------------------------------- operators
int dsmc_op_mt_generate(const uint32_t* a, const uint32_t* b, const uint32_t* c, const uint32_t* seeds,
                        uint32_t n_streams, uint32_t n_draws, uint32_t* out, void* stream) {
  if (!a || !b || !c || !seeds || !out) return set_global_error(DSMC_ERR_INVALID, "null argument");
  dsmc_launch::mt_generate(a, b, c, seeds, n_streams, n_draws, out, (hipStream_t)stream);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_global_error(DSMC_ERR_DEVICE, hipGetErrorString(e));
  return DSMC_OK;
}

int dsmc_op_cell_ids(const dsmc_settings* s, const float* x, const float* y, const float* z, uint32_t n,
                     uint32_t* keys, void* stream) {
  if (!s || !x || !y || !z || !keys) return set_global_error(DSMC_ERR_INVALID, "null argument");
  InstDesc I;
  std::memset(&I, 0, sizeof(I));
  for (int a = 0; a < 3; ++a) {
    if (s->grid_dim[a] <= 0) return set_global_error(DSMC_ERR_INVALID, "bad grid");
    I.gmin[a] = s->bmin[a];
    I.gmax[a] = s->bmax[a];
    I.cell[a] = (s->bmax[a] - s->bmin[a]) / (float)s->grid_dim[a];
    I.inv_cell[a] = 1.0f / I.cell[a];
    I.dim[a] = (uint32_t)s->grid_dim[a];
  }
  dsmc_launch::cell_ids(I, x, y, z, n, keys, (hipStream_t)stream);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_global_error(DSMC_ERR_DEVICE, hipGetErrorString(e));
  return DSMC_OK;
}

uint64_t dsmc_op_sort_work_words(uint32_t n, uint32_t n_cells) { return dsmc_launch::op_sort_work_words(n, n_cells); }

int dsmc_op_sort(const uint32_t* keys, uint32_t n, uint32_t n_cells, uint32_t* perm, uint32_t* cell_start,
                 uint32_t* work, void* stream) {
  if ((n && (!keys || !perm)) || !cell_start || !work) return set_global_error(DSMC_ERR_INVALID, "null argument");
  if (dsmc_launch::op_sort(keys, n, n_cells, perm, cell_start, work, (hipStream_t)stream))
    return set_global_error(DSMC_ERR_CAPACITY, "too many keys or cells for the scan");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_global_error(DSMC_ERR_DEVICE, hipGetErrorString(e));
  return DSMC_OK;
}

// ---------------------------------------------------- sort without allocation
// dsmc_op_sort with its workspace in a file-scope device array instead of a
// caller buffer, for callers that cannot allocate (code under HIP graph
// capture, serving loops). Limits: n and n_cells up to 2^20; the array
// (op_sort_work_words(2^20, 2^20) words, 17.3 MB) is held on every GPU the
// library runs on until the process ends. Every call on a GPU uses that
// GPU's copy: calls that overlap on one GPU race.
namespace {
constexpr uint32_t kStaticSortMax = 1u << 20;
constexpr uint64_t kStaticSortBlocks = (kStaticSortMax + 4095u) / 4096u;
constexpr uint64_t kStaticSortWords =
    kStaticSortMax + 2ull * (256ull * kStaticSortBlocks + 1ull) + (dsmc_launch::kScanMaxTiles + 1ull) +
    3ull * kStaticSortMax;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) uint32_t scratch[kStaticSortWords];

namespace {
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (host code must not pass the symbol itself to a kernel).
int scratch_on_current_device(uint32_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_global_error(DSMC_ERR_DEVICE, hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return set_global_error(DSMC_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return set_global_error(DSMC_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed");
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return DSMC_OK;
}
}  // namespace

int dsmc_op_sort_static(const uint32_t* keys, uint32_t n, uint32_t n_cells, uint32_t* perm, uint32_t* cell_start,
                        void* stream) {
  if ((n && (!keys || !perm)) || !cell_start) return set_global_error(DSMC_ERR_INVALID, "null argument");
  if (n > kStaticSortMax || n_cells > kStaticSortMax)
    return set_global_error(DSMC_ERR_CAPACITY, "dsmc_op_sort_static takes at most " + std::to_string(kStaticSortMax) +
                                                   " keys and cells (got " + std::to_string(n) + " keys, " +
                                                   std::to_string(n_cells) + " cells)");
  if (dsmc_launch::op_sort_work_words(n, n_cells) > kStaticSortWords)
    return set_global_error(DSMC_ERR_CAPACITY, "workspace exceeds the static array");
  uint32_t* work = nullptr;
  int rc = scratch_on_current_device(&work);
  if (rc != DSMC_OK) return rc;
  if (dsmc_launch::op_sort(keys, n, n_cells, perm, cell_start, work, (hipStream_t)stream))
    return set_global_error(DSMC_ERR_CAPACITY, "too many keys or cells for the scan");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_global_error(DSMC_ERR_DEVICE, hipGetErrorString(e));
  return DSMC_OK;
}

}  // extern "C"

namespace {

// Flags of step number `stepno` (1-based): sampling cadence and debug outputs.
uint32_t step_flags(const dsmc_solver* h, uint64_t stepno, bool* sample) {
  uint32_t flags = h->debug_flags;
  bool smp;
  if (h->physics()) smp = h->sampling;
  else {
    uint32_t every = h->s.sample_every ? h->s.sample_every : 100u;
    smp = (stepno % every) == 0;
  }
  if (smp) flags |= dsmc_launch::kFlagSample;
  if (sample) *sample = smp;
  return flags;
}

// Enqueue the kernels of one step reading buffer `par` (no host sync, no
// memsets: safe to capture in a HIP graph).
// Workgroups of the wall pass (both launches): one wave per chunk of 64
// bitmask words for small instances, at most 1024 workgroups (about what
// the GPU keeps resident at the pass's occupancy); grid-stride beyond.
uint32_t wall_grid(const InstDesc& I) {
  const uint64_t chunks = ((uint64_t)I.cap / 64 + 2 + 63) / 64;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(1024, (chunks + 3) / 4));
}

int enqueue_step(dsmc_solver* h, uint32_t par, uint32_t flags, hipEvent_t* ev) {
  InstDesc& I = h->h;
  hipStream_t st = h->stream;
  if (ev) HIPC(hipEventRecord(ev[0], st));
  dsmc_launch::step_scan(h->d_desc, 1, 0, I.ncells, h->scan_scratch, st);
  if (ev) HIPC(hipEventRecord(ev[1], st));
  dsmc_launch::scatter(h->d_desc, par, h->scatter_grid, st);
  if (ev) HIPC(hipEventRecord(ev[2], st));
  dsmc_launch::tile(I.mode, h->d_desc, par, flags, I.n_vwg, st);
  if (ev) HIPC(hipEventRecord(ev[3], st));
  // injection beside the wall pass (a phase timer then sees only the part
  // of the injection that outlasts the wall pass)
  const bool overlap = I.defer_bits && I.n_face_cells;
  auto inject = [&](hipStream_t s) {
    dsmc_launch::inject_count(h->d_desc, std::min(I.n_streams, I.n_face_cells), s);
    dsmc_launch::step_scan(h->d_desc, 1, 1, I.n_face_cells, h->scan_scratch, s);
    dsmc_launch::inject_generate(h->d_desc, par ^ 1u, std::min(I.n_streams, I.n_face_cells), s);
  };
  if (overlap) {
    HIPC(hipEventRecord(h->side.fork, st));
    HIPC(hipStreamWaitEvent(h->side.s, h->side.fork, 0));
    inject(h->side.s);
    HIPC(hipEventRecord(h->side.join, h->side.s));
  }
  if (I.defer_bits) dsmc_launch::walls(h->d_desc, par ^ 1u, flags, I.n_wall_wg, st);
  if (ev) HIPC(hipEventRecord(ev[4], st));
  if (overlap) HIPC(hipStreamWaitEvent(st, h->side.join, 0));
  else if (I.n_face_cells) inject(st);
  if (ev) HIPC(hipEventRecord(ev[5], st));
  return DSMC_OK;
}

int step_once(dsmc_solver* h) {
  const uint64_t stepno = h->steps + 1;
  bool sample = false;
  uint32_t flags = step_flags(h, stepno, &sample);
  hipEvent_t* ev = nullptr;
  if (h->timing) {
    if (h->ev_used + 6 > 6 * 4096) CHECK_RC(fold_timings(h));
    while (h->ev_pool.size() < h->ev_used + 6) {
      hipEvent_t e;
      HIPC(hipEventCreate(&e));
      h->ev_pool.push_back(e);
    }
    ev = &h->ev_pool[h->ev_used];
    h->ev_used += 6;
  }
  CHECK_RC(enqueue_step(h, h->par, flags, ev));
  HIPC(hipGetLastError());
  h->par ^= 1u;
  h->advance_time(stepno - h->steps);
  h->steps = stepno;
  if (sample) h->samples += 1;
  return DSMC_OK;
}

// Two consecutive steps (parity par, then par^1) captured as one graph.
int run_graph_pair(dsmc_solver* h, uint32_t f1, uint32_t f2) {
  uint64_t key = ((uint64_t)h->par << 63) | ((uint64_t)f1 << 32) | f2;
  auto it = h->graphs.find(key);
  hipGraphExec_t exec;
  if (it == h->graphs.end()) {
    hipGraph_t g;
    HIPC(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc1 = enqueue_step(h, h->par, f1, nullptr);
    int rc2 = rc1 == DSMC_OK ? enqueue_step(h, h->par ^ 1u, f2, nullptr) : rc1;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc2 != DSMC_OK) return rc2;
    HIPC(e);
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPC(e);
    h->graphs[key] = exec;
  } else {
    exec = it->second;
  }
  HIPC(hipGraphLaunch(exec, h->stream));
  return DSMC_OK;
}

}  // namespace

// ------------------------------------------------------------------ batches
// Several instances advanced by the same kernel launches (grid.y = instance):
// the batched-ensemble step. Every instance keeps its own buffers, RNG
// streams, dt and descriptor; the batch holds the descriptor array.
struct dsmc_batch {
  std::vector<dsmc_solver*> hs;
  std::string err;
  hipStream_t stream = nullptr;
  SideStream side;
  hipEvent_t ev = nullptr;
  InstDesc* d_descs = nullptr;
  uint32_t* scratch = nullptr;
  uint32_t n = 0, mode = 0, max_cells = 0, max_faces = 0, max_vwg = 0, max_inj_threads = 0, scatter_grid = 0;
  uint32_t wall_grid = 0;
  bool use_graphs = true;
  std::map<uint64_t, hipGraphExec_t> graphs;
  ~dsmc_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    if (d_descs) (void)hipFree(d_descs);
    if (scratch) (void)hipFree(scratch);
    if (ev) (void)hipEventDestroy(ev);
    if (stream) (void)hipStreamDestroy(stream);
  }
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

namespace {

int batch_hip(dsmc_batch* b, hipError_t e, const char* what) {
  if (e == hipSuccess) return DSMC_OK;
  return b->fail(DSMC_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define BHIP(expr) CHECK_RC(batch_hip(b, (expr), #expr))

int enqueue_batch_step(dsmc_batch* b, uint32_t par, uint32_t flags) {
  hipStream_t st = b->stream;
  dsmc_launch::step_scan(b->d_descs, b->n, 0, b->max_cells, b->scratch, st);
  dsmc_launch::scatter(b->d_descs, par, b->scatter_grid, st, b->n);
  dsmc_launch::tile(b->mode, b->d_descs, par, flags, b->max_vwg, st, b->n);
  // injection beside the wall pass, as for one solver
  const bool overlap = b->wall_grid && b->max_faces;
  auto inject = [&](hipStream_t s) {
    dsmc_launch::inject_count(b->d_descs, b->max_inj_threads, s, b->n);
    dsmc_launch::step_scan(b->d_descs, b->n, 1, b->max_faces, b->scratch, s);
    dsmc_launch::inject_generate(b->d_descs, par ^ 1u, b->max_inj_threads, s, b->n);
  };
  hipError_t e = hipSuccess;
  if (overlap) {
    e = hipEventRecord(b->side.fork, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->side.s, b->side.fork, 0);
    if (e == hipSuccess) inject(b->side.s);
    if (e == hipSuccess) e = hipEventRecord(b->side.join, b->side.s);
  }
  if (b->wall_grid) dsmc_launch::walls(b->d_descs, par ^ 1u, flags, b->wall_grid, st, b->n);
  if (overlap && e == hipSuccess) e = hipStreamWaitEvent(st, b->side.join, 0);
  else if (!overlap && b->max_faces) inject(st);
  if (e != hipSuccess) return b->fail(DSMC_ERR_DEVICE, std::string("batch side stream: ") + hipGetErrorString(e));
  return DSMC_OK;
}

}  // namespace

extern "C" {

int dsmc_batch_create(dsmc_solver* const* hs, uint32_t n, dsmc_batch** out) {
  if (!hs || !out || n == 0) return set_global_error(DSMC_ERR_INVALID, "empty batch");
  *out = nullptr;
  for (uint32_t i = 0; i < n; ++i) {
    if (!hs[i] || !hs[i]->have_state) return set_global_error(DSMC_ERR_STATE, "batch member without loaded state");
    if (hs[i]->h.mode != hs[0]->h.mode) return set_global_error(DSMC_ERR_INVALID, "batch members differ in mode");
    if (hs[i]->par != hs[0]->par || hs[i]->steps != hs[0]->steps)
      return set_global_error(DSMC_ERR_STATE, "batch members must be at the same step");
    if (!hs[i]->physics() && hs[i]->s.sample_every != hs[0]->s.sample_every)
      return set_global_error(DSMC_ERR_INVALID, "REFERENCE batch members must share sample_every");
    for (uint32_t j = 0; j < i; ++j)
      if (hs[j] == hs[i]) return set_global_error(DSMC_ERR_INVALID, "a solver appears twice in the batch");
  }
  auto* b = new dsmc_batch();
  b->hs.assign(hs, hs + n);
  b->n = n;
  b->mode = hs[0]->h.mode;
  for (auto* h : b->hs) {
    b->max_cells = std::max(b->max_cells, h->h.ncells);
    b->max_faces = std::max(b->max_faces, h->h.n_face_cells);
    b->max_vwg = std::max(b->max_vwg, h->h.n_vwg);
    b->max_inj_threads = std::max(b->max_inj_threads, std::min(h->h.n_streams, h->h.n_face_cells));
    b->scatter_grid = std::max(b->scatter_grid, h->scatter_grid);
    if (h->h.defer_bits) b->wall_grid = std::max(b->wall_grid, h->h.n_wall_wg);
  }
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = b->side.create();
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&b->d_descs, sizeof(InstDesc) * n);
  if (e == hipSuccess) e = hipMalloc(&b->scratch, sizeof(uint32_t) * n * (dsmc_launch::kScanMaxTiles + 1));
  if (e != hipSuccess) {
    delete b;
    return set_global_error(DSMC_ERR_DEVICE, std::string("batch allocation: ") + hipGetErrorString(e));
  }
  *out = b;
  return DSMC_OK;
}

void dsmc_batch_destroy(dsmc_batch* b) { delete b; }

const char* dsmc_batch_last_error(const dsmc_batch* b) { return b ? b->err.c_str() : global_error(); }

int dsmc_batch_set_graphs(dsmc_batch* b, int enabled) {
  if (!b) return set_global_error(DSMC_ERR_INVALID, "null batch");
  b->use_graphs = enabled != 0;
  return DSMC_OK;
}

int dsmc_batch_set_sampling(dsmc_batch* b, int enabled, int reset) {
  if (!b) return set_global_error(DSMC_ERR_INVALID, "null batch");
  for (auto* h : b->hs) {
    int rc = dsmc_set_sampling(h, enabled, reset);
    if (rc) return b->fail(rc, h->err);
  }
  return DSMC_OK;
}

int dsmc_batch_run(dsmc_batch* b, uint32_t steps) {
  if (!b) return set_global_error(DSMC_ERR_INVALID, "null batch");
  dsmc_solver* h0 = b->hs[0];
  for (auto* h : b->hs)
    if (h->par != h0->par || h->steps != h0->steps) return b->fail(DSMC_ERR_STATE, "batch members drifted apart");
  // order after each member's queued work, refresh the descriptors
  for (auto* h : b->hs) {
    BHIP(hipEventRecord(b->ev, h->stream));
    BHIP(hipStreamWaitEvent(b->stream, b->ev, 0));
  }
  // Sampling: in PHYSICS mode each member samples while its own sampling is
  // on (sample_on in its descriptor); launches carry the sample flag when any
  // member samples. REFERENCE members share the every-Nth-step schedule.
  std::vector<InstDesc> descs(b->n);
  bool any_sampling = false;
  for (uint32_t i = 0; i < b->n; ++i) {
    descs[i] = b->hs[i]->h;
    descs[i].sample_on = (!b->hs[i]->physics() || b->hs[i]->sampling) ? 1u : 0u;
    any_sampling = any_sampling || b->hs[i]->sampling;
  }
  BHIP(hipMemcpyAsync(b->d_descs, descs.data(), sizeof(InstDesc) * b->n, hipMemcpyHostToDevice, b->stream));
  auto flags_at = [&](uint64_t stepno, bool* smp) {
    uint32_t f = step_flags(h0, stepno, smp);
    if (h0->physics()) {
      *smp = any_sampling;
      f = any_sampling ? (f | dsmc_launch::kFlagSample) : (f & ~dsmc_launch::kFlagSample);
    }
    return f;
  };
  uint32_t done = 0;
  uint32_t par = h0->par;
  uint64_t stepno = h0->steps;
  uint32_t samples = 0;
  while (done < steps) {
    bool s1 = false, s2 = false;
    uint32_t f1 = flags_at(stepno + 1, &s1);
    if (b->use_graphs && steps - done >= 2) {
      uint32_t f2 = flags_at(stepno + 2, &s2);
      uint64_t key = ((uint64_t)par << 63) | ((uint64_t)f1 << 32) | f2;
      auto it = b->graphs.find(key);
      hipGraphExec_t exec;
      if (it == b->graphs.end()) {
        hipGraph_t g;
        BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        enqueue_batch_step(b, par, f1);
        enqueue_batch_step(b, par ^ 1u, f2);
        BHIP(hipStreamEndCapture(b->stream, &g));
        hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        BHIP(e);
        b->graphs[key] = exec;
      } else {
        exec = it->second;
      }
      BHIP(hipGraphLaunch(exec, b->stream));
      stepno += 2;
      done += 2;
      samples += (s1 ? 1u : 0u) + (s2 ? 1u : 0u);
    } else {
      CHECK_RC(enqueue_batch_step(b, par, f1));
      par ^= 1u;
      stepno += 1;
      done += 1;
      samples += s1 ? 1u : 0u;
    }
  }
  BHIP(hipGetLastError());
  for (auto* h : b->hs) {
    h->par = par;
    h->advance_time(stepno - h->steps);
    h->steps = stepno;
    if (!h->physics() || h->sampling) h->samples += samples;
  }
  // members' later calls run after the batch's work
  BHIP(hipEventRecord(b->ev, b->stream));
  for (auto* h : b->hs) BHIP(hipStreamWaitEvent(h->stream, b->ev, 0));
  return DSMC_OK;
}

int dsmc_batch_synchronize(dsmc_batch* b) {
  if (!b) return set_global_error(DSMC_ERR_INVALID, "null batch");
  BHIP(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < b->hs.size(); ++i) {
    int rc = b->hs[i]->check_status();
    if (rc != DSMC_OK) return b->fail(rc, "member " + std::to_string(i) + ": " + b->hs[i]->err);
  }
  return DSMC_OK;
}

}  // extern "C"
