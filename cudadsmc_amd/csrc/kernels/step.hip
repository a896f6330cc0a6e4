// HIP kernels of the DSMC step for gfx950 (CDNA4, wave64).
//
// Step = scan(count) -> scatter -> tile [-> inject], replacing the
// reference's CalculateCellIDs -> thrust::sort_by_key -> cudaMemset ->
// FindCellBoundaries -> ComputeBird -> IntegrateAndProcessCollisions ->
// SampleCellStatistics sequence (include/gpu_dsmc.inl:221-274).
//
//  * Keys of the next step are computed by the move itself and counted per
//    cell with wave-aggregated atomics, so the "sort" is a counting sort:
//    an exclusive scan of the counts gives the cell offsets
//    (cellStartEnd of FindCellBoundariesDev, dsmc_kernel.cu:1734-1787). The
//    counting atomics return each particle's rank in its cell, which is
//    kept, so the scatter that builds the permutation needs no atomics.
//  * The tile kernel owns a tile of consecutive cells: it restores the stable
//    order inside each cell segment (ascending storage index, i.e. the
//    permutation of a stable sort_by_key, dsmc_kernel.cu:2048), gathers the
//    particles (the REORDER gather, :1777-1786), runs the NTC loop one lane
//    per cell out of LDS (:1002-1089), moves each particle with walls and
//    box faces (:1596-1657), writes the next key and samples per cell
//    (:856-890).
//  * RNG (PHYSICS): stream s = 256 * virtual_workgroup + lane; the tile
//    kernel runs a fixed grid of n_streams/256 workgroups, each walking a
//    fixed sequence of tiles, so a cell always draws from the same stream.
//    REFERENCE mode keeps the reference's layout: cell c collides with
//    stream c and loading thread t draws from stream t.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "physics.h"

using namespace dsmc_dev;

namespace {

constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;  // 4096

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ void mt_load(const InstDesc& I, uint32_t s, Mt& m) {
  const gptr<uint32_t> st = gmem(I.mt_state);
#pragma unroll
  for (int w = 0; w < kMtN; ++w) m.st[w] = st[(uint32_t)w * I.n_streams + s];
  m.a = gmem(I.mt_a)[s];
  m.b = gmem(I.mt_b)[s];
  m.c = gmem(I.mt_c)[s];
}

__device__ __forceinline__ void mt_store(const InstDesc& I, uint32_t s, const Mt& m) {
#pragma unroll
  for (int w = 0; w < kMtN; ++w) gmem(I.mt_state)[(uint32_t)w * I.n_streams + s] = m.st[w];
}

// Wave-aggregated atomicAdd(&cnt[key], 1) for keys of consecutive lanes,
// returning each lane's count before its own increment: one atomic per run
// of equal keys, whose lanes take consecutive values. Invalid keys are
// skipped (their result is meaningless). Split in two so that a caller can
// issue the atomics of several rounds before waiting for any: count_issue
// returns the run head's old count (0 elsewhere) and the wave's run heads,
// count_finish the lane's value.
__device__ __forceinline__ uint32_t count_issue(gptr<uint32_t> cnt, uint32_t key, uint64_t& heads) {
  const uint32_t lane = lane_id();
  const uint32_t prev = __shfl_up(key, 1);
  const bool head = (lane == 0) || (key != prev);
  heads = __ballot(head);
  uint32_t base = 0;
  if (head && key != kInvalidKey) {
    const uint64_t below_incl = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const uint64_t above = heads & ~below_incl;
    const uint32_t next = above ? (uint32_t)__ffsll((unsigned long long)above) - 1u : 64u;
    base = gatomic_add(gelem(cnt, key), next - lane);
  }
  return base;
}

__device__ __forceinline__ uint32_t count_finish(uint32_t base, uint64_t heads) {
  const uint32_t lane = lane_id();
  const uint64_t below_incl = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const uint32_t hl = 63u - (uint32_t)__clzll((long long)(heads & below_incl));
  return __shfl(base, (int)hl) + (lane - hl);
}

// Wave-aggregated atomicAdd(&cnt[key], 1) for keys of consecutive lanes:
// one atomic per run of equal keys. Invalid keys are skipped.
__device__ __forceinline__ void count_keys(gptr<uint32_t> cnt, uint32_t key) {
  uint32_t lane = lane_id();
  uint32_t prev = __shfl_up(key, 1);
  bool head = (lane == 0) || (key != prev);
  uint64_t heads = __ballot(head);
  if (head && key != kInvalidKey) {
    uint64_t above = heads & ~((2ull << lane) - 1ull);
    if (lane == 63) above = 0;
    uint32_t next = above ? (uint32_t)__ffsll((unsigned long long)above) - 1u : 64u;
    gatomic_add(gelem(cnt, key), next - lane);
  }
}

}  // namespace

// ======================================================================= RNG
// InitMersenneTwistersDev (dsmc_kernel.cu:548-562): seed every stream.
__global__ void k_mt_seed(uint32_t* state, const uint32_t* seeds, uint32_t n_streams) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_streams) return;
  Mt m;
  mt_seed(m, seeds[s]);
#pragma unroll
  for (int w = 0; w < kMtN; ++w) state[(uint32_t)w * n_streams + s] = m.st[w];
  state[(uint32_t)kMtN * n_streams + s] = 0u;
}

// Stream outputs for the RNG parity tests (GenerateRandomSamplesUint4,
// mtwister.cu:207-230, with one stream per thread).
__global__ void k_mt_generate(const uint32_t* a, const uint32_t* b, const uint32_t* c, const uint32_t* seeds,
                              uint32_t n_streams, uint32_t n_draws, uint32_t* out) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_streams) return;
  Mt m;
  m.a = a[s];
  m.b = b[s];
  m.c = c[s];
  mt_seed(m, seeds[s]);
  for (uint32_t d = 0; d < n_draws; ++d) out[(size_t)d * n_streams + s] = mt_next(m);
}

// ====================================================================== keys
// CalculateCellIDsDev (dsmc_kernel.cu:1712-1724) on SoA positions.
__global__ void k_cell_ids(InstDesc I, const float* x, const float* y, const float* z, uint32_t n, uint32_t* keys) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = cell_key(I, x[i], y[i], z[i]);
}

// Keys and counts of a freshly loaded state (storage order, buffer `par`).
__global__ void k_keys_count(const InstDesc* __restrict__ Ip, uint32_t par, uint32_t n) {
  const InstDesc& I = *Ip;
  uint32_t base = blockIdx.x * blockDim.x;
  uint32_t i = base + threadIdx.x;
  uint32_t key = kInvalidKey;
  if (i < n) {
    const float x = I.px[par][i], y = I.py[par][i], z = I.pz[par][i];
    key = I.mode == kModePhysics ? cell_key_phys(I, x, y, z) : cell_key(I, x, y, z);
    I.key[par][i] = key;
  }
  count_keys(gmem(I.count), key);
}

// Plain counting of arbitrary keys (sort operator); keys >= n_cells are
// dropped.
__global__ void k_count_keys(const uint32_t* keys, uint32_t n, uint32_t n_cells, uint32_t* cnt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t key = i < n ? keys[i] : kInvalidKey;
  count_keys(gmem(cnt), key < n_cells ? key : kInvalidKey);
}

// ====================================================================== scan
// Exclusive scan of uint32 counts in three passes (tile sums, scan of the
// tile sums, tile scan). Optionally copies the offsets into a cursor array,
// clears the counts and stores the total.
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t* in, uint32_t n, uint32_t* tile_sums) {
  __shared__ uint32_t red[kScanThreads / 64];
  uint32_t base = blockIdx.x * kScanTile;
  uint32_t sum = 0;
  if (base + kScanTile <= n && (reinterpret_cast<uintptr_t>(in) & 15u) == 0u) {
    // a whole tile of a 16-byte aligned array: vector loads, consecutive lanes on consecutive vectors
    const gptr<const uint4> src = gmem<const uint4>(reinterpret_cast<const uint4*>(in + base));
#pragma unroll
    for (uint32_t k = 0; k < kScanItems / 4u; ++k) {
      const uint4 q = gload(src + k * kScanThreads + threadIdx.x);
      sum += (q.x + q.y) + (q.z + q.w);
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k) {
      uint32_t i = base + k * kScanThreads + threadIdx.x;
      if (i < n) sum += in[i];
    }
  }
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off);
  if (lane_id() == 0) red[threadIdx.x / 64] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < kScanThreads / 64; ++w) t += red[w];
    tile_sums[blockIdx.x] = t;
  }
}

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total, uint32_t* sh) {
  // sh: kScanThreads/64 + 1 words
  uint32_t lane = lane_id(), wid = threadIdx.x / 64;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (uint32_t w = 0; w < kScanThreads / 64; ++w) {
      uint32_t t = sh[w];
      sh[w] = run;
      run += t;
    }
    sh[kScanThreads / 64] = run;
  }
  __syncthreads();
  uint32_t r = sh[wid] + x - v;
  *total = sh[kScanThreads / 64];
  __syncthreads();
  return r;
}

// Scans tile_sums in place (n_tiles <= kScanTile) and writes the grand total.
__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(uint32_t* tile_sums, uint32_t n_tiles, uint32_t* total_out) {
  __shared__ uint32_t sh[kScanThreads / 64 + 1];
  uint32_t v[kScanItems];
  uint32_t local = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k) {
    uint32_t i = threadIdx.x * kScanItems + k;
    v[k] = i < n_tiles ? tile_sums[i] : 0u;
    local += v[k];
  }
  uint32_t total;
  uint32_t off = block_exclusive_scan(local, &total, sh);
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k) {
    uint32_t i = threadIdx.x * kScanItems + k;
    if (i < n_tiles) tile_sums[i] = off;
    off += v[k];
  }
  if (threadIdx.x == 0 && total_out) *total_out = total;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(uint32_t* in, uint32_t n, const uint32_t* tile_sums,
                                                            uint32_t* out, uint32_t* cursor, uint32_t clear_in,
                                                            const uint32_t* total) {
  __shared__ uint32_t sh[kScanThreads / 64 + 1];
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t local = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k) {
    uint32_t i = base + k;
    v[k] = i < n ? in[i] : 0u;
    local += v[k];
  }
  uint32_t tot;
  uint32_t off = block_exclusive_scan(local, &tot, sh) + tile_sums[blockIdx.x];
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k) {
    uint32_t i = base + k;
    if (i < n) {
      out[i] = off;
      if (cursor) cursor[i] = off;
      if (clear_in) in[i] = 0u;
    }
    off += v[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = *total;
}

// Step scans over a batch of instances (grid.y = instance). which = 0: the
// cell counts (count -> cell_start, count cleared, total in ctr[1]);
// which = 1: the injection face counts, in place (total in face_count[n]).
// scratch: kScanMaxTiles + 1 words per instance.
__device__ __forceinline__ void bscan_args(const InstDesc& I, uint32_t which, uint32_t*& in, uint32_t& n) {
  in = which ? I.face_count : I.count;
  n = which ? I.n_face_cells : I.ncells;
}

__global__ void __launch_bounds__(kScanThreads) k_bscan_reduce(const InstDesc* __restrict__ D, uint32_t which,
                                                              uint32_t* scratch) {
  const InstDesc& I = D[blockIdx.y];
  uint32_t* in;
  uint32_t n;
  bscan_args(I, which, in, n);
  if (blockIdx.x * kScanTile >= n) return;
  __shared__ uint32_t red[kScanThreads / 64];
  uint32_t base = blockIdx.x * kScanTile;
  uint32_t sum = 0;
  if (base + kScanTile <= n && (reinterpret_cast<uintptr_t>(in) & 15u) == 0u) {
    // a whole tile of a 16-byte aligned array: vector loads, consecutive lanes on consecutive vectors
    const gptr<const uint4> src = gmem<const uint4>(reinterpret_cast<const uint4*>(in + base));
#pragma unroll
    for (uint32_t k = 0; k < kScanItems / 4u; ++k) {
      const uint4 q = gload(src + k * kScanThreads + threadIdx.x);
      sum += (q.x + q.y) + (q.z + q.w);
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k) {
      uint32_t i = base + k * kScanThreads + threadIdx.x;
      if (i < n) sum += in[i];
    }
  }
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off);
  if (lane_id() == 0) red[threadIdx.x / 64] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < kScanThreads / 64; ++w) t += red[w];
    scratch[blockIdx.y * (dsmc_launch::kScanMaxTiles + 1) + blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kScanThreads) k_bscan_tiles(const InstDesc* __restrict__ D, uint32_t which,
                                                             uint32_t* scratch) {
  const InstDesc& I = D[blockIdx.y];
  uint32_t* in;
  uint32_t n;
  bscan_args(I, which, in, n);
  if (n == 0) return;
  uint32_t* ts = scratch + blockIdx.y * (dsmc_launch::kScanMaxTiles + 1);
  const uint32_t n_tiles = (n + kScanTile - 1) / kScanTile;
  __shared__ uint32_t sh[kScanThreads / 64 + 1];
  uint32_t v[kScanItems];
  uint32_t local = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k) {
    uint32_t i = threadIdx.x * kScanItems + k;
    v[k] = i < n_tiles ? ts[i] : 0u;
    local += v[k];
  }
  uint32_t total;
  uint32_t off = block_exclusive_scan(local, &total, sh);
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k) {
    uint32_t i = threadIdx.x * kScanItems + k;
    if (i < n_tiles) ts[i] = off;
    off += v[k];
  }
  if (threadIdx.x == 0) {
    ts[dsmc_launch::kScanMaxTiles] = total;
    if (which == 0) {
      I.ctr[1] = total;
      I.ctr[4] += 1u;  // step number (counter of the wall draws)
      I.ctr[5] = 0u;   // deferred particles of the step
      for (uint32_t q = 0; q < 8u; ++q) I.ctr[kTileQueueCtr + q] = 0u;  // the tile kernel's queues
    }
  }
}

__global__ void __launch_bounds__(kScanThreads) k_bscan_apply(const InstDesc* __restrict__ D, uint32_t which,
                                                             const uint32_t* scratch) {
  const InstDesc& I = D[blockIdx.y];
  uint32_t* in;
  uint32_t n;
  bscan_args(I, which, in, n);
  if (blockIdx.x * kScanTile >= n) return;
  const uint32_t* ts = scratch + blockIdx.y * (dsmc_launch::kScanMaxTiles + 1);
  uint32_t* out = which ? I.face_count : I.cell_start;
  __shared__ uint32_t sh[kScanThreads / 64 + 1];
  // each thread scans kScanItems consecutive entries; a thread whose entries
  // all lie below n moves them as 16-byte vectors when the arrays are 16-byte
  // aligned (the engine's are 256-byte aligned; base is a multiple of 16)
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  const bool whole = base + kScanItems <= n &&
                     ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out) |
                       (which ? 0u : (reinterpret_cast<uintptr_t>(I.cursor) | reinterpret_cast<uintptr_t>(I.local_cnt)))) &
                      15u) == 0u;
  uint32_t v[kScanItems];
  if (whole) {
    const gptr<const uint4> src = gmem<const uint4>(reinterpret_cast<const uint4*>(in + base));
#pragma unroll
    for (uint32_t k = 0; k < kScanItems / 4u; ++k) {
      const uint4 q = gload(src + k);
      v[4 * k] = q.x;
      v[4 * k + 1] = q.y;
      v[4 * k + 2] = q.z;
      v[4 * k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k) v[k] = base + k < n ? in[base + k] : 0u;
  }
  // the cells' counts of particles that stayed in their tile (cursor =
  // offset + count), loaded before the scan's barriers
  uint4 l[kScanItems / 4u];
#pragma unroll
  for (uint32_t k = 0; k < kScanItems / 4u; ++k)
    l[k] = (which == 0 && whole) ? gload(gmem<const uint4>(reinterpret_cast<const uint4*>(I.local_cnt + base)) + k)
                                 : make_uint4(0u, 0u, 0u, 0u);
  uint32_t local = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k) local += v[k];
  uint32_t tot;
  uint32_t off = block_exclusive_scan(local, &tot, sh) + ts[blockIdx.x];
  if (whole) {
    const gptr<uint4> dst = gmem(reinterpret_cast<uint4*>(out + base));
    const gptr<uint4> clr = gmem(reinterpret_cast<uint4*>(in + base));
#pragma unroll
    for (uint32_t k = 0; k < kScanItems / 4u; ++k) {
      const uint32_t o0 = off, o1 = o0 + v[4 * k], o2 = o1 + v[4 * k + 1], o3 = o2 + v[4 * k + 2];
      off = o3 + v[4 * k + 3];
      gstore(dst + k, make_uint4(o0, o1, o2, o3));
      if (which == 0) {
        gstore(clr + k, make_uint4(0u, 0u, 0u, 0u));
        gstore(gmem(reinterpret_cast<uint4*>(I.cursor + base)) + k,
               make_uint4(o0 + l[k].x, o1 + l[k].y, o2 + l[k].z, o3 + l[k].w));
      }
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k) {
      uint32_t i = base + k;
      if (i < n) {
        out[i] = off;
        if (which == 0) {
          in[i] = 0u;
          I.cursor[i] = off + I.local_cnt[i];
        }
      }
      off += v[k];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = ts[dsmc_launch::kScanMaxTiles];
}

// =================================================================== scatter
// Places every stored particle i in the sorted order: perm[dst] = i, dst in
// its cell's segment. The tile kernel stored the particles of each tile in
// one run of slots (tile_p0) and counted per cell how many of them stayed
// in the tile's own cells (local_cnt); the scan set cursor = cell_start +
// local_cnt. A workgroup takes up to kScatterSpan consecutive cells (whole
// tiles) and the slots of their tiles: the particles that stayed in their
// tile are ranked with LDS atomics into the first local_cnt places of their
// cells; the others (they left their tile, went through the wall pass, or
// were injected after the tile-ordered slots) take the places from the
// cursor on, with one wave-aggregated global atomic per run of equal keys
// and the atomics of kScatterRounds rounds of 256 slots in flight together
// (one atomic per particle is far slower: runs are common, e.g. a cell's
// particles after the wall pass).
// The order within a cell is the order in which the atomics landed; the
// tile kernel sorts each cell's segment by storage index.
constexpr uint32_t kScatterRounds = 4;
constexpr uint32_t kScatterSpan = 1024;  // cells per workgroup pass (a multiple of every tile size)

__global__ void __launch_bounds__(256) k_scatter(const InstDesc* __restrict__ Ip, uint32_t par) {
  const InstDesc& I = Ip[blockIdx.y];
  __shared__ uint32_t s_cnt[kScatterSpan];
  __shared__ uint32_t s_tp[kScatterSpan / 32u + 1u];
  const uint32_t tid = threadIdx.x;
  const uint32_t n = I.ctr[0];
  const gptr<const uint32_t> keys = gmem<const uint32_t>(I.key[par]);
  const gptr<const uint32_t> start = gmem<const uint32_t>(I.cell_start);
  const gptr<const uint32_t> tp0 = gmem<const uint32_t>(I.tile_p0);
  const gptr<uint32_t> perm = gmem(I.perm);
  const gptr<uint32_t> cursor = gmem(I.cursor);
  const uint32_t cap = I.cap, tile_cells = I.tile_cells;
  // end of the tile-ordered slots; instances with small tiles place every
  // particle through the cursors (their tile kernel counts none as stayed)
  const uint32_t S = I.tile_dyn ? min(tp0[I.ntiles], n) : 0u;
  // tiles per pass: up to kScatterSpan cells, fewer when the instance has
  // too few tiles to give every workgroup a pass
  const uint32_t K = max(1u, min(kScatterSpan / tile_cells, I.ntiles / gridDim.x));
  // Slots of the pass's tiles [t0, t0 + K) (runs starting at s_tp[]), in
  // rounds of 256 slots from q0 (below q1). Each lane keeps the tile of its
  // last slot in t (its slots only grow). A particle that stayed in its
  // tile takes the next LDS place of its cell (s_cnt, seeded with the cell
  // starts); any other takes its cell's cursor, with one wave-aggregated
  // atomic per run of equal keys (particles of one cell that leave together,
  // or a cell's particles after the wall pass), the atomics of
  // kScatterRounds rounds issued before any result is used. kk holds the
  // keys of slots q0...; the keys of the next kScatterRounds rounds are
  // requested before these atomics, so both latencies overlap.
  auto load_keys = [&](uint32_t q0, uint32_t q1, uint32_t (&kk)[kScatterRounds]) {
#pragma unroll
    for (uint32_t r = 0; r < kScatterRounds; ++r) {
      const uint32_t i = q0 + r * 256u + tid;
      kk[r] = i < q1 ? *gelem(keys, i) : kInvalidKey;
    }
  };
  auto place_tiled = [&](uint32_t q0, uint32_t q1, uint32_t c0, uint32_t K, uint32_t& t,
                         uint32_t (&kk)[kScatterRounds]) {
    uint32_t kc[kScatterRounds], loc[kScatterRounds], g[kScatterRounds], base[kScatterRounds], lds[kScatterRounds];
    uint64_t heads[kScatterRounds];
#pragma unroll
    for (uint32_t r = 0; r < kScatterRounds; ++r) kc[r] = kk[r];
    if (q0 + 256u * kScatterRounds < q1) load_keys(q0 + 256u * kScatterRounds, q1, kk);  // uniform
#pragma unroll
    for (uint32_t r = 0; r < kScatterRounds; ++r) {
      const uint32_t i = q0 + r * 256u + tid;
      const uint32_t key = kc[r];
      while (t + 1u < K && s_tp[t + 1u] <= i) ++t;
      const uint32_t rel = key - c0;
      const bool stayed = rel - t * tile_cells < tile_cells;  // its cell lies in the tile that holds its slot
      loc[r] = stayed ? rel : kInvalidKey;
      g[r] = stayed ? kInvalidKey : key_cell(key);
      base[r] = count_issue(cursor, g[r], heads[r]);
      if (stayed) lds[r] = atomicAdd(&s_cnt[rel], 1u);
    }
#pragma unroll
    for (uint32_t r = 0; r < kScatterRounds; ++r) {
      const uint32_t i = q0 + r * 256u + tid;
      uint32_t dst = count_finish(base[r], heads[r]);
      if (loc[r] != kInvalidKey) dst = lds[r];
      if ((loc[r] != kInvalidKey || g[r] != kInvalidKey) && dst < cap) *gelem(perm, dst) = i;
    }
  };
  // slots after the tile-ordered ones (or all of them before the first
  // step): through the cursors, one wave-aggregated atomic per run of
  // equal keys
  auto place_rest = [&](uint32_t q0, uint32_t q1) {
    uint32_t key[kScatterRounds], base[kScatterRounds];
    uint64_t heads[kScatterRounds];
#pragma unroll
    for (uint32_t r = 0; r < kScatterRounds; ++r) {
      const uint32_t i = q0 + r * 256u + tid;
      key[r] = i < q1 ? key_cell(*gelem(keys, i)) : kInvalidKey;
      base[r] = count_issue(cursor, key[r], heads[r]);
    }
#pragma unroll
    for (uint32_t r = 0; r < kScatterRounds; ++r) {
      const uint32_t i = q0 + r * 256u + tid;
      const uint32_t dst = count_finish(base[r], heads[r]);
      if (key[r] != kInvalidKey && dst < cap) *gelem(perm, dst) = i;
    }
  };
  const uint32_t passes = S ? (I.ntiles + K - 1u) / K : 0u;
  for (uint32_t g = blockIdx.x; g < passes; g += gridDim.x) {  // uniform over the workgroup
    const uint32_t t0 = g * K;
    const uint32_t c0 = t0 * I.tile_cells;
    for (uint32_t c = tid; c < K * tile_cells; c += 256u) s_cnt[c] = c0 + c < I.ncells ? start[c0 + c] : 0u;
    if (tid <= K) s_tp[tid] = t0 + tid <= I.ntiles ? min(tp0[t0 + tid], S) : S;
    __syncthreads();
    const uint32_t p0 = s_tp[0], p1 = max(s_tp[K], p0);
    uint32_t t = 0;
    uint32_t kk[kScatterRounds];
    if (p0 < p1) load_keys(p0, p1, kk);
    __builtin_amdgcn_s_waitcnt(0);  // the loop starts with no load pending, so its top waits for none
    for (uint32_t q0 = p0; q0 < p1; q0 += 256u * kScatterRounds) place_tiled(q0, p1, c0, K, t, kk);
    __syncthreads();  // s_cnt and s_tp are set again for the next pass
  }
  for (uint32_t q0 = S + blockIdx.x * 256u * kScatterRounds; q0 < n; q0 += gridDim.x * 256u * kScatterRounds)
    place_rest(q0, n);
}

// ============================================================ sort operator
// Stable LSD radix sort of (key, storage index) pairs, 8 bits per pass: the
// permutation of SortParticlesIndices' thrust::sort_by_key
// (dsmc_kernel.cu:2033-2051). Keys >= n_cells (dropped particles) sort as
// n_cells, after every valid key. A pass is a per-block digit histogram
// (hist[digit * nblocks + block]), an exclusive scan of it, and a scatter in
// which each block ranks its 4096 keys stably: rounds of 256 positions in
// order, waves in order within a round, lanes in order within a wave (peers
// of equal digit found with eight ballots).
constexpr uint32_t kRsThreads = 256;
constexpr uint32_t kRsRounds = 16;
constexpr uint32_t kRsTile = kRsThreads * kRsRounds;  // 4096 keys per block

__device__ __forceinline__ uint32_t rs_key(const uint32_t* keys, uint32_t i, uint32_t n_cells, bool raw) {
  const uint32_t k = keys[i];
  return raw ? (k < n_cells ? k : n_cells) : k;
}

__global__ void __launch_bounds__(kRsThreads) k_rs_hist(const uint32_t* keys, uint32_t n, uint32_t n_cells,
                                                        uint32_t raw, uint32_t shift, uint32_t nblocks,
                                                        uint32_t* hist) {
  __shared__ uint32_t s_h[256];
  const uint32_t tid = threadIdx.x, b = blockIdx.x;
  s_h[tid] = 0u;
  __syncthreads();
  const uint32_t base = b * kRsTile;
  for (uint32_t r = 0; r < kRsRounds; ++r) {
    const uint32_t i = base + r * kRsThreads + tid;
    if (i < n) atomicAdd(&s_h[(rs_key(keys, i, n_cells, raw != 0u) >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)tid * nblocks + b] = s_h[tid];
}

__global__ void __launch_bounds__(kRsThreads) k_rs_scatter(const uint32_t* keys, const uint32_t* vals, uint32_t n,
                                                           uint32_t n_cells, uint32_t raw, uint32_t shift,
                                                           uint32_t nblocks, const uint32_t* offs,
                                                           uint32_t* keys_out, uint32_t* vals_out) {
  __shared__ uint32_t s_run[256];              // this block's keys of each digit in earlier rounds
  __shared__ uint32_t s_wc[kRsThreads / 64][256];  // keys of each digit per wave in this round
  const uint32_t tid = threadIdx.x, b = blockIdx.x, lane = lane_id(), w = tid / 64u;
  s_run[tid] = 0u;
#pragma unroll
  for (uint32_t k = 0; k < kRsThreads / 64; ++k) s_wc[k][tid] = 0u;
  const uint32_t dbase = offs[(size_t)tid * nblocks + b];  // block's first slot of digit tid
  __shared__ uint32_t s_base[256];
  s_base[tid] = dbase;
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t base = b * kRsTile;
  for (uint32_t r = 0; r < kRsRounds; ++r) {
    const uint32_t i = base + r * kRsThreads + tid;
    const bool valid = i < n;
    const uint32_t k = valid ? rs_key(keys, i, n_cells, raw != 0u) : 0u;
    const uint32_t v = valid ? (vals ? vals[i] : i) : 0u;
    const uint32_t d = (k >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (uint32_t bit = 0; bit < 8; ++bit) {
      const uint64_t m = __ballot((d >> bit) & 1u);
      peers &= ((d >> bit) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & below);
    if (valid && rank == 0u) s_wc[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t off = s_run[d] + rank;
      for (uint32_t k2 = 0; k2 < w; ++k2) off += s_wc[k2][d];
      const uint32_t dst = s_base[d] + off;
      keys_out[dst] = k;
      vals_out[dst] = v;
    }
    __syncthreads();
    uint32_t t = 0u;
#pragma unroll
    for (uint32_t k2 = 0; k2 < kRsThreads / 64; ++k2) {
      t += s_wc[k2][tid];
      s_wc[k2][tid] = 0u;
    }
    s_run[tid] += t;
    __syncthreads();
  }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ====================================================================== tile
using dsmc_launch::kFlagSample;
using dsmc_launch::kFlagKeepPerm;
using dsmc_launch::kFlagHits;

constexpr uint32_t kMoveBatch = 6;  // rounds of 256 positions whose loads are issued together
// PHYSICS: rounds of 256 permutation entries of the next tile loaded into
// registers during this tile's move (the next tile's first phase then only
// stores them to LDS)
constexpr uint32_t kPermPrefetch = 6;

// LDS of the PHYSICS NTC rounds.
struct NtcLds {
  uint32_t* off;   // [kTileCells] each cell's first item of the round
  uint32_t* nmax;  // [kTileCells] (sigma c_r)max at the round's start (float bits)
  uint32_t* rmax;  // [kTileCells] largest sigma c_r seen so far (float bits)
  uint8_t* icell;  // [kTileThreads] cell of each of the round's first kTileThreads items
  uint32_t* red;   // [8] wave totals of the scan, then the waves' largest round counts
};

// PHYSICS NTC of one tile in rounds of disjoint pairs (common/pairing.h).
// Lane l < nc computes cell l's candidate count M; round r holds
// min(floor(n / 2), M - r floor(n / 2)) pairs of each cell, and the pairs of
// all cells of a round are dealt to every thread of the workgroup (item t of
// the round: its cell from the scanned pair counts). The pairs of a round
// share no particle, so each is tested against the round's starting
// (sigma c_r)max and collided independently; the maximum then takes the
// round's largest sigma c_r. Draws: Philox4x32-10 under key
// (seed, kNtcKey1) at counter (cell, step, round, item).
__device__ __forceinline__ void tile_ntc_rounds(const InstDesc& I, uint32_t c0, uint32_t nc, uint32_t p0,
                                                const uint32_t* s_start, float* vx, float* vy, float* vz,
                                                float2& nt, float frac, uint32_t stepno, const NtcLds& L,
                                                CollAcc& acc) {
  using namespace dsmc_pairing;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid / 64u;
  uint32_t n = 0u, M = 0u;
  if (tid < nc) {
    n = s_start[tid + 1] - s_start[tid];
    if (n >= 2u) {
      const float vol = I.cell_volume * (I.fluid_frac ? fmaxf(frac, kMinFluidFrac) : 1.0f);
      const float fpc = (float)(0.5 * (double)n * (double)(n - 1u) * (double)I.fnum * (double)nt.x *
                                    (double)I.dt / (double)vol +
                                (double)nt.y);
      M = dm_f2u_sat(fpc);
      nt.y = fpc - (float)M;
      if (M > kMaxPairsPerCell) {
        M = kMaxPairsPerCell;
        atomicOr(&I.ctr[3], kStatusPairs);
      }
      acc.sel += M;
    }
    L.nmax[tid] = __float_as_uint(nt.x);
    L.rmax[tid] = __float_as_uint(nt.x);
  }
  // rounds of the tile (the most over its cells), found with the first
  // round's scan
  uint32_t R = rounds_of(n, M);
  for (int o = 32; o > 0; o >>= 1) R = max(R, (uint32_t)__shfl_xor((int)R, o));
  uint32_t rounds = 1u;
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t m = round_pairs(n, M, r);
    // exclusive scan of the round's pair counts over the cells
    uint32_t x = m;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) L.red[wid] = x;
    if (r == 0u && lane == 0) L.red[4 + wid] = R;
    __syncthreads();
    uint32_t base = 0u, total = 0u;
#pragma unroll
    for (uint32_t w = 0; w < kTileThreads / 64; ++w) {
      const uint32_t t = L.red[w];
      base += w < wid ? t : 0u;
      total += t;
      if (r == 0u) rounds = max(rounds, L.red[4 + w]);
    }
    if (total == 0u) break;  // uniform: round 0 without pairs means a tile without candidates
    const uint32_t off = base + x - m;
    if (tid < nc) {
      L.off[tid] = off;
      for (uint32_t q = off; q < off + m && q < kTileThreads; ++q) L.icell[q] = (uint8_t)tid;
    }
    __syncthreads();
    for (uint32_t it = tid; it < total; it += kTileThreads) {
      uint32_t c;
      if (it < kTileThreads) {
        c = L.icell[it];
      } else {  // the last cell whose first item is at or before it
        uint32_t lo = 0u, hi = nc - 1u;
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1u) >> 1;
          if (L.off[mid] <= it) lo = mid;
          else hi = mid - 1u;
        }
        c = lo;
      }
      const uint32_t oc = L.off[c];
      const uint32_t mc = (c + 1u < nc ? L.off[c + 1u] : total) - oc;
      const uint32_t k = it - oc;
      const uint32_t a = s_start[c] - p0, nn = s_start[c + 1u] - s_start[c];
      const dsmc_rng::U32x4 rk = dsmc_rng::philox4x32_10({{c0 + c, stepno, r, kRoundItem}}, I.wall_key[0], kNtcKey1);
      const dsmc_rng::U32x4 d = dsmc_rng::philox4x32_10({{c0 + c, stepno, r, k}}, I.wall_key[0], kNtcKey1);
      uint32_t i, j;
      pair_members(nn, mc, k, rk, i, j);
      i += a;
      j += a;
      const float dx = vx[i] - vx[j], dy = vy[i] - vy[j], dz = vz[i] - vz[j];
      const float mag = sqrtf(dx * dx + dy * dy + dz * dz);
      const float sc =
          (I.vhs_exp == 1.0f) ? I.vhs_k * mag : (mag > 0.0f ? I.vhs_k * dm_powf(mag, I.vhs_exp) : 0.0f);
      atomicMax(&L.rmax[c], __float_as_uint(sc));
      const float max0 = __uint_as_float(L.nmax[c]);
      if (sc > dsmc_rng::px_unit(d.v[0]) * max0) {
        acc.coll += 1u;
        acc.cr += __double2ll_rn((double)mag * kCrScale);
        scatter_uv(vx, vy, vz, i, j, mag, dsmc_rng::px_unit(d.v[1]), dsmc_rng::px_unit(d.v[2]));
      }
    }
    __syncthreads();
    if (tid < nc) L.nmax[tid] = L.rmax[tid];
  }
  if (tid < nc) nt.x = __uint_as_float(L.rmax[tid]);
}

template <uint32_t MODE, bool LDS>
__device__ __forceinline__ void tile_body(const InstDesc& I, uint32_t par, uint32_t flags, uint32_t tile,
                                          uint32_t c0, uint32_t nc, uint32_t p0, uint32_t np,
                                          const uint32_t* s_start, const uint8_t* s_nearc, uint32_t& tile_n,
                                          gptr<uint32_t> queue, uint32_t qx, bool per_xcd, uint32_t e_static,
                                          uint32_t* s_pop,
                                          uint32_t* sn_start, uint8_t* sn_nearc, uint32_t (&pf)[kPermPrefetch],
                                          uint32_t* s_raw,
                                          uint32_t* s_perm, uint8_t* s_owner, float* s_vx, float* s_vy,
                                          float* s_vz, uint32_t stepno, const NtcLds& ntl, CollAcc& cacc) {
  const uint32_t tid = threadIdx.x;
  const uint32_t opar = par ^ 1u;
  const gptr<const float> ipx = gmem<const float>(I.px[par]);
  const gptr<const float> ipy = gmem<const float>(I.py[par]);
  const gptr<const float> ipz = gmem<const float>(I.pz[par]);
  const gptr<const float> ivx = gmem<const float>(I.vx[par]);
  const gptr<const float> ivy = gmem<const float>(I.vy[par]);
  const gptr<const float> ivz = gmem<const float>(I.vz[par]);
  const gptr<float> opx = gmem(I.px[opar] + p0);
  const gptr<float> opy = gmem(I.py[opar] + p0);
  const gptr<float> opz = gmem(I.pz[opar] + p0);
  const gptr<float> ovx = gmem(I.vx[opar] + p0);
  const gptr<float> ovy = gmem(I.vy[opar] + p0);
  const gptr<float> ovz = gmem(I.vz[opar] + p0);
  const gptr<uint32_t> okey = gmem(I.key[opar] + p0);
  uint32_t* raw = LDS ? s_raw : (I.perm + p0);
  uint32_t* srt = LDS ? s_perm : (I.perm_tmp + p0);
  uint8_t* own = LDS ? s_owner : (I.owner_tmp + p0);
  float* vx = LDS ? s_vx : I.vx[opar] + p0;
  float* vy = LDS ? s_vy : I.vy[opar] + p0;
  float* vz = LDS ? s_vz : I.vz[opar] + p0;

  // the workgroup's next tile: taken from its queue now (dynamic
  // schedule; the atomic's latency is covered by phases 1 and 2), or the
  // next element of its fixed sequence
  const bool dyn = I.tile_dyn != 0u;
  uint32_t popped = 0u;
  if (dyn && tid == 0) popped = gatomic_add(queue, 1u);
  // (1) raw permutation of the tile and the owning cell of each position
  if (LDS) {
    if (MODE == kModePhysics) {
#pragma unroll
      for (uint32_t k = 0; k < kPermPrefetch; ++k)
        if (tid + k * kTileThreads < np) raw[tid + k * kTileThreads] = pf[k];
      for (uint32_t q = tid + kPermPrefetch * kTileThreads; q < np; q += kTileThreads) raw[q] = *gelem(gmem(I.perm), p0 + q);
    } else {
      for (uint32_t q = tid; q < np; q += kTileThreads) raw[q] = *gelem(gmem(I.perm), p0 + q);
    }
  }
  if (tid < nc) {
    uint32_t a = s_start[tid] - p0, b = s_start[tid + 1] - p0;
    for (uint32_t q = a; q < b; ++q) own[q] = (uint8_t)tid;
  }
  __syncthreads();
  // (2) rank of each entry among its cell's entries -> stable order
  for (uint32_t q = tid; q < np; q += kTileThreads) {
    uint32_t c = own[q];
    uint32_t a = s_start[c] - p0, b = s_start[c + 1] - p0;
    uint32_t v = raw[q];
    // four independent reads per iteration: dense cells are long loops
    uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    uint32_t j = a;
    for (; j + 4u <= b; j += 4u) {
      r0 += raw[j] < v ? 1u : 0u;
      r1 += raw[j + 1u] < v ? 1u : 0u;
      r2 += raw[j + 2u] < v ? 1u : 0u;
      r3 += raw[j + 3u] < v ? 1u : 0u;
    }
    for (; j < b; ++j) r0 += raw[j] < v ? 1u : 0u;
    srt[a + r0 + r1 + r2 + r3] = v;
  }
  if (dyn && tid == 0) *s_pop = popped;
  __syncthreads();
  {
    const uint32_t jn = tile_queue_pos(dyn ? *s_pop : e_static, qx, per_xcd);
    tile_n = jn < I.ntiles ? (I.tile_order ? gmem(I.tile_order)[jn] : jn) : kInvalidKey;
  }
  // the collision lane's cell state, requested with the first velocity
  // gathers so that the latencies overlap (lanes without a gather request it
  // before the loop)
  float2 nt_pf = make_float2(0.0f, 0.0f);
  float frac_pf = 1.0f;
  float4 st_pf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  auto load_cell_state = [&]() {
    if (tid < nc && s_start[tid + 1] - s_start[tid] >= 2u) {
      nt_pf = gload(gmem<const float2>(I.ntc) + c0 + tid);
      if (MODE == kModeReference) st_pf = gload(gmem<const float4>(I.cstat) + c0 + tid);
      else if (I.fluid_frac) frac_pf = gmem(I.fluid_frac)[c0 + tid];
    }
  };
  if (tid >= np) load_cell_state();
  // (3) gather velocities in sorted order; the loads of kMoveBatch positions
  // are issued before any store so that their latencies overlap
  for (uint32_t q0 = tid; q0 < np; q0 += kMoveBatch * kTileThreads) {
    float gx[kMoveBatch], gy[kMoveBatch], gz[kMoveBatch];
#pragma unroll
    for (uint32_t k = 0; k < kMoveBatch; ++k) {
      const uint32_t q = q0 + k * kTileThreads;
      gx[k] = gy[k] = gz[k] = 0.0f;
      if (q < np) {
        const uint32_t src = srt[q];
        gx[k] = *gelem(ivx, src);
        gy[k] = *gelem(ivy, src);
        gz[k] = *gelem(ivz, src);
      }
    }
    if (q0 == tid) load_cell_state();
#pragma unroll
    for (uint32_t k = 0; k < kMoveBatch; ++k) {
      const uint32_t q = q0 + k * kTileThreads;
      if (q < np) {
        vx[q] = gx[k];
        vy[q] = gy[k];
        vz[q] = gz[k];
      }
    }
  }
  __syncthreads();
  // Every load issued so far has arrived (the velocity gather waited for
  // them); saying so here keeps the compiler from waiting for the next
  // tile's loads below at the first use of an older one in the NTC phase.
  wait_vmem();
  // the next tile's cell offsets and near flags, stored to the other LDS
  // buffer at the end of this phase: the next tile starts without a load
  uint32_t nx_start = 0u, nx_end = 0u, ncn = 0u;
  uint8_t nx_near = 255u;
  if (tile_n != kInvalidKey) {
    const uint32_t c0n = tile_n * I.tile_cells;
    ncn = min(I.tile_cells, I.ncells - c0n);
    nx_start = gmem(I.cell_start)[c0n + min(tid, ncn)];
    if (tid == 0) nx_end = gmem(I.cell_start)[c0n + ncn];
    if (MODE == kModePhysics) nx_near = (I.near_cell && tid < ncn) ? gmem(I.near_cell)[c0n + tid] : (uint8_t)255;
  }
  // (4) NTC collisions: REFERENCE one lane per cell (the reference's loop),
  // PHYSICS in rounds of disjoint pairs over the whole workgroup
  if (MODE == kModeReference) {
    if (tid < nc) {
      const uint32_t c = c0 + tid;
      const uint32_t a = s_start[tid] - p0, n = s_start[tid + 1] - s_start[tid];
      if (n >= 2u) {
        float2 nt = nt_pf;
        float4 st = st_pf;
        if (I.taus) {
          uint4 q = gload(gmem<const uint4>(reinterpret_cast<const uint4*>(I.taus)) + c);
          Taus t = {q.x, q.y, q.z, q.w};
          ntc_reference(I, vx + a, vy + a, vz + a, n, nt, st, t);
          gstore(gmem(reinterpret_cast<uint4*>(I.taus)) + c, make_uint4(t.x, t.y, t.z, t.w));
        } else {
          // ComputeBirdCollisionsDev draws from the cell's own stream
          // (ds_mtState[cellid], dsmc_kernel.cu:1030-1031)
          Mt cr;
          mt_load(I, c, cr);
          ntc_reference(I, vx + a, vy + a, vz + a, n, nt, st, cr);
          mt_store(I, c, cr);
        }
        gstore(gmem(I.cstat) + c, st);
        gstore(gmem(I.ntc) + c, nt);
      }
    }
  }
  float2 ntc_state = nt_pf;
  if (MODE == kModePhysics) {
    // frac_pf is first used here: without this marker the compiler applies
    // the NTC's clamp of it right after its load, whose wait then comes
    // before the velocity gather instead of after it
    asm volatile("" : "+v"(frac_pf));
    tile_ntc_rounds(I, c0, nc, p0, s_start, vx, vy, vz, ntc_state, frac_pf, stepno, ntl, cacc);
  }
  if (tile_n != kInvalidKey) {
    sn_start[tid] = nx_start;
    if (tid == 0) sn_start[ncn] = nx_end;
    if (MODE == kModePhysics) sn_nearc[tid] = nx_near;
  }
  if (MODE == kModePhysics) {  // after the next tile's offsets, so that their wait is not a wait for this store
    if (tid < nc && s_start[tid + 1] - s_start[tid] >= 2u) gstore(gmem(I.ntc) + c0 + tid, ntc_state);
  }
  // per-cell counts of the particles that stay in the tile (the NTC is done
  // with the offsets that share this LDS)
  ntl.off[tid] = 0u;
  if (tid == 0) gmem(I.tile_p0)[tile] = p0;
  __syncthreads();
  // the next tile's first permutation entries (its offsets are in LDS now),
  // requested with the first batch of position gathers and waited for with
  // them, so that no later wait (the next tile's first use of pf comes after
  // this tile's stores) has to wait for the stores
  auto load_pf = [&]() {
    if (MODE == kModePhysics && tile_n != kInvalidKey) {
      const uint32_t p0n = sn_start[0];
      const uint32_t npn = sn_start[ncn] - p0n;
#pragma unroll
      for (uint32_t k = 0; k < kPermPrefetch; ++k) {
        const uint32_t q = tid + k * kTileThreads;
        pf[k] = (npn <= kLdsCap && q < npn) ? *gelem(gmem(I.perm), p0n + q) : 0u;
      }
    }
  };
  // (5) move, walls, box faces, next key; rounds keep waves converged for
  // the aggregated counting
  const uint32_t rounds = (np + kTileThreads - 1u) / kTileThreads;
  if (rounds == 0u) {
    load_pf();
    wait_vmem();
  }
  for (uint32_t r0 = 0; r0 < rounds; r0 += kMoveBatch) {
    if (r0 == 0u) load_pf();
    // positions of up to kMoveBatch rounds in flight before the first store
    float bx[kMoveBatch], by[kMoveBatch], bz[kMoveBatch];
    uint32_t bsrc[kMoveBatch];
#pragma unroll
    for (uint32_t k = 0; k < kMoveBatch; ++k) {
      const uint32_t q = (r0 + k) * kTileThreads + tid;
      bsrc[k] = 0u;
      bx[k] = by[k] = bz[k] = 0.0f;
      if (q < np) {
        const uint32_t src = srt[q];
        bsrc[k] = src;
        bx[k] = *gelem(ipx, src);
        by[k] = *gelem(ipy, src);
        bz[k] = *gelem(ipz, src);
      }
    }
    wait_vmem();
#pragma unroll
    for (uint32_t k = 0; k < kMoveBatch; ++k) {
    if (r0 + k >= rounds) break;  // uniform over the workgroup
    uint32_t q = (r0 + k) * kTileThreads + tid;
    uint32_t key = kInvalidKey;
    // REFERENCE: the wall flights of the wavefront run together (every lane
    // takes part, with or without a particle)
    float rx = 0.0f, ry = 0.0f, rz = 0.0f, rux = 0.0f, ruy = 0.0f, ruz = 0.0f;
    uint32_t rhits = 0;
    if (MODE == kModeReference) {
      if (q < np) {
        rx = bx[k], ry = by[k], rz = bz[k];
        rux = vx[q], ruy = vy[q], ruz = vz[q];
      }
      rhits = move_reference_wave(I, q < np, rx, ry, rz, rux, ruy, ruz, I.dt, ntl.rmax + (tid & ~63u));
    }
    if (q < np) {
      const uint32_t src = bsrc[k];
      float x = bx[k], y = by[k], z = bz[k];
      float ux = vx[q], uy = vy[q], uz = vz[q];
      uint32_t hits = 0;
      if (MODE == kModeReference) {
        x = rx, y = ry, z = rz, ux = rux, uy = ruy, uz = ruz;
        hits = rhits;
        key = cell_key(I, x, y, z);
      } else if (phys_near(I, s_nearc[own[q]], ux, uy, uz, I.dt)) {
        key = kDeferredKey;  // stored as is; k_walls moves it and counts its key
      } else {
        key = move_free(I, x, y, z, ux, uy, uz, I.dt);
      }
      opx[q] = x;
      opy[q] = y;
      opz[q] = z;
      okey[q] = key;
      if (LDS) {
        ovx[q] = ux;
        ovy[q] = uy;
        ovz[q] = uz;
      }
      if (!LDS || (flags & kFlagSample)) {  // post-move velocities: output (global path) or sampled
        vx[q] = ux;
        vy[q] = uy;
        vz[q] = uz;
      }
      if (flags & kFlagHits) gmem(I.hits)[p0 + q] = hits;
      if (flags & kFlagKeepPerm) *gelem(gmem(I.perm), p0 + q) = src;
    }
    count_keys(gmem(I.count), key == kDeferredKey ? kInvalidKey : key);
    if (I.tile_dyn && key - c0 < nc) atomicAdd(&ntl.off[key - c0], 1u);  // stayed in the tile's cells
    if (MODE == kModePhysics) {
      // mark the wave's deferred slots in the bitmask the wall pass reads
      const uint64_t m = __ballot(key == kDeferredKey);
      if (m && lane_id() == 0) {
        const uint32_t s0 = p0 + (r0 + k) * kTileThreads + (tid & ~63u);
        const uint32_t sh = s0 & 63u;
        const gptr<unsigned long long> wb = gmem(I.defer_bits) + s0 / 64u;
        __hip_atomic_fetch_or(wb, (unsigned long long)(m << sh), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (sh)
          __hip_atomic_fetch_or(wb + 1, (unsigned long long)(m >> (64u - sh)), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    }
  }

  // (6) sampling of the post-move velocities, one lane per cell
  if ((flags & kFlagSample) && I.sample_on) {
    __syncthreads();
    if (tid < nc) {
      uint32_t c = c0 + tid;
      uint32_t a = s_start[tid] - p0, b = s_start[tid + 1] - p0;
      if (b > a) {
        if (MODE == kModeReference) {
          // SampleCellStatisticsDev (dsmc_kernel.cu:856-890), FP32 sums
          const gptr<float> S = gmem(reinterpret_cast<float*>(I.samp));
          const uint32_t C = I.ncells;
          float sn = S[c], sx = S[C + c], sy = S[2 * C + c], sz = S[3 * C + c], s2 = S[4 * C + c];
          for (uint32_t q = a; q < b; ++q) {
            float ux = vx[q], uy = vy[q], uz = vz[q];
            sn += 1.0f;
            sx += ux;
            sy += uy;
            sz += uz;
            s2 += ux * ux + uy * uy + uz * uz;
          }
          S[c] = sn; S[C + c] = sx; S[2 * C + c] = sy; S[3 * C + c] = sz; S[4 * C + c] = s2;
        } else {
          const gptr<double> S = gmem(I.samp);
          const uint32_t C = I.ncells;
          double sn = S[c], sx = S[C + c], sy = S[2 * C + c], sz = S[3 * C + c], s2 = S[4 * C + c];
          for (uint32_t q = a; q < b; ++q) {
            double ux = vx[q], uy = vy[q], uz = vz[q];
            sn += 1.0;
            sx += ux;
            sy += uy;
            sz += uz;
            s2 += ux * ux + uy * uy + uz * uz;
          }
          S[c] = sn; S[C + c] = sx; S[2 * C + c] = sy; S[3 * C + c] = sz; S[4 * C + c] = s2;
        }
      }
    }
  }
}

template <uint32_t MODE>
__global__ void __launch_bounds__(kTileThreads, 4) k_tile(const InstDesc* __restrict__ Ip, uint32_t par, uint32_t flags) {
  const InstDesc& I = Ip[blockIdx.y];
  if (blockIdx.x >= I.n_vwg) return;  // batched launches size the grid for the largest instance
  // cell offsets and near flags of the current tile and of the next one
  __shared__ uint32_t s_start[2][kTileCells + 1];
  __shared__ uint8_t s_nearc[2][kTileCells];
  __shared__ uint32_t s_raw[kLdsCap];
  __shared__ uint32_t s_perm[kLdsCap];
  __shared__ uint8_t s_owner[kLdsCap];
  __shared__ float s_vx[kLdsCap];
  __shared__ float s_vy[kLdsCap];
  __shared__ double s_red[kTileThreads / 64][kAccSlots];
  // PHYSICS NTC rounds (the item-cell table shares s_red, used only at the end)
  __shared__ uint32_t s_ntc_off[kTileCells];
  __shared__ uint32_t s_ntc_nmax[kTileCells];
  __shared__ uint32_t s_ntc_rmax[kTileCells];
  __shared__ uint32_t s_ntc_red[8];
  static_assert(sizeof(s_red) >= kTileThreads, "item-cell table must fit in s_red");
  const NtcLds ntl = {s_ntc_off, s_ntc_nmax, s_ntc_rmax, reinterpret_cast<uint8_t*>(&s_red[0][0]), s_ntc_red};
  // The raw permutation is dead once the rank sort has run (phase 2), so
  // the z velocities reuse its space: 35 KB of LDS, four workgroups per CU.
  float* const s_vz = reinterpret_cast<float*>(s_raw);

  const uint32_t tid = threadIdx.x;
  const uint32_t vw = blockIdx.x;
  const uint32_t stepno = gmem(I.ctr)[4];  // this step's number (counter of the PHYSICS draws)
  if (vw == 0 && tid == 0) {
    I.ctr[0] = I.ctr[1];
    I.tile_p0[I.ntiles] = I.ctr[1];  // the slots this launch stores in tile order
  }
  // tiles come from the XCD's queue (dynamic schedule) or, for instances
  // whose tiles are small, as the fixed sequence of elements vw / 8,
  // vw / 8 + G / 8, ... of that queue (no atomic per tile)
  __shared__ uint32_t s_pop;
  const bool per_xcd = I.n_vwg % 8u == 0u;
  const uint32_t qx = per_xcd ? vw % 8u : 0u;
  const uint32_t e_stride = per_xcd ? I.n_vwg / 8u : I.n_vwg;
  const gptr<uint32_t> queue = gmem(I.ctr) + kTileQueueCtr + qx;
  uint32_t e = per_xcd ? vw / 8u : vw;
  if (I.tile_dyn) {
    if (tid == 0) s_pop = gatomic_add(queue, 1u);
    __syncthreads();
    e = s_pop;
  }
  const uint32_t j = tile_queue_pos(e, qx, per_xcd);
  if (j >= I.ntiles) return;  // no tile: this workgroup's counters stay as they are
  CollAcc cacc = {0ull, 0ull, 0ll};
  uint32_t tile = kInvalidKey;
  {
    tile = I.tile_order ? gmem(I.tile_order)[j] : j;
    const uint32_t c0 = tile * I.tile_cells;
    const uint32_t nc = min(I.tile_cells, I.ncells - c0);
    s_start[0][tid] = gmem(I.cell_start)[c0 + min(tid, nc)];
    if (tid == 0) s_start[0][nc] = gmem(I.cell_start)[c0 + nc];
    if (MODE == kModePhysics)
      s_nearc[0][tid] = (I.near_cell && tid < nc) ? gmem(I.near_cell)[c0 + tid] : (uint8_t)255;
  }
  __syncthreads();
  uint32_t pf[kPermPrefetch];
  if (MODE == kModePhysics && tile != kInvalidKey) {
    const uint32_t nc = min(I.tile_cells, I.ncells - tile * I.tile_cells);
    const uint32_t p0 = s_start[0][0], np = s_start[0][nc] - p0;
#pragma unroll
    for (uint32_t k = 0; k < kPermPrefetch; ++k) {
      const uint32_t q = tid + k * kTileThreads;
      pf[k] = (np <= kLdsCap && q < np) ? *gelem(gmem(I.perm), p0 + q) : 0u;
    }
  }
  wait_vmem();  // the first tile needs them now; later tiles find theirs arrived
  for (uint32_t b = 0; tile != kInvalidKey; b ^= 1u) {
    uint32_t tile_n = kInvalidKey;  // set by tile_body
    e += e_stride;                  // the next element of the fixed sequence
    const uint32_t c0 = tile * I.tile_cells;
    const uint32_t nc = min(I.tile_cells, I.ncells - c0);
    const uint32_t p0 = s_start[b][0];
    const uint32_t np = s_start[b][nc] - p0;
    if (np <= kLdsCap)
      tile_body<MODE, true>(I, par, flags, tile, c0, nc, p0, np, s_start[b], s_nearc[b], tile_n, queue, qx,
                            per_xcd, e, &s_pop, s_start[b ^ 1u],
                            s_nearc[b ^ 1u], pf, s_raw, s_perm, s_owner, s_vx, s_vy, s_vz, stepno, ntl, cacc);
    else
      tile_body<MODE, false>(I, par, flags, tile, c0, nc, p0, np, s_start[b], s_nearc[b], tile_n, queue, qx,
                            per_xcd, e, &s_pop, s_start[b ^ 1u],
                             s_nearc[b ^ 1u], pf, s_raw, s_perm, s_owner, s_vx, s_vy, s_vz, stepno, ntl, cacc);
    __syncthreads();
    if (tid < nc) gmem(I.local_cnt)[c0 + tid] = s_ntc_off[tid];
    tile = tile_n;
  }

  // per-workgroup collision counters (integers: exact in any order)
  if (MODE == kModePhysics) {
    long long v[3] = {(long long)cacc.sel, (long long)cacc.coll, cacc.cr};
    long long* red = reinterpret_cast<long long*>(&s_red[0][0]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      long long x = v[k];
      for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
      if (lane_id() == 0) red[(tid / 64) * 3 + k] = x;
    }
    __syncthreads();
    if (tid < 3) {
      long long t = 0;
      for (uint32_t w = 0; w < kTileThreads / 64; ++w) t += red[w * 3 + tid];
      gmem(I.wg_acc)[vw * kAccSlots + tid] += t;
    }
  }
}

template __global__ void k_tile<kModeReference>(const InstDesc*, uint32_t, uint32_t);
template __global__ void k_tile<kModePhysics>(const InstDesc*, uint32_t, uint32_t);

// =============================================================== wall pass
// PHYSICS: the particles the tile kernel deferred (key kDeferredKey, state
// after collisions, before the move; one bit per slot in defer_bits) fly
// with walls and faces here, in two launches.
//
// k_defer_list compacts the marked slots: each wave takes chunks of 64
// bitmask words (4096 slots) and finds their set bits with one coalesced
// load; the workgroup reserves room for its waves' chunks in defer_list with
// one atomic; each word's slots are written as one contiguous run (lane l
// tests bit l) and the words cleared.
__global__ void __launch_bounds__(kTileThreads) k_defer_list(const InstDesc* __restrict__ Ip) {
  const InstDesc& I = Ip[blockIdx.y];
  if (!I.defer_bits || blockIdx.x >= I.n_wall_wg) return;
  __shared__ uint32_t s_tot[kTileThreads / 64u + 1u];
  const uint32_t lane = lane_id(), wave = threadIdx.x / 64u;
  const uint32_t nwords = (I.ctr[1] + 63u) / 64u;  // deferred slots lie below the sorted count
  const uint32_t nchunks = (nwords + 63u) / 64u;
  const gptr<unsigned long long> bits = gmem(I.defer_bits);
  const gptr<uint32_t> list = gmem(I.defer_list);
  constexpr uint32_t kWaves = kTileThreads / 64u;
  // the workgroup takes kWaves chunks per round (one per wave) and reserves
  // room for all of them with one atomic
  for (uint32_t c0 = blockIdx.x * kWaves; c0 < nchunks; c0 += I.n_wall_wg * kWaves) {
    const uint32_t ch = c0 + wave;
    const uint32_t w = ch * 64u + lane;
    const unsigned long long word = (ch < nchunks && w < nwords) ? bits[w] : 0ull;
    const uint32_t pc = (uint32_t)__popcll(word);
    uint32_t incl = pc;
    for (uint32_t off = 1; off < 64u; off <<= 1) {
      const uint32_t t = __shfl_up(incl, off);
      if (lane >= off) incl += t;
    }
    if (lane == 63u) s_tot[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t sum = 0;
      for (uint32_t q = 0; q < kWaves; ++q) {
        const uint32_t x = s_tot[q];
        s_tot[q] = sum;
        sum += x;
      }
      s_tot[kWaves] = sum ? gatomic_add(gmem(I.ctr) + 5, sum) : 0u;
    }
    __syncthreads();
    const uint32_t base = s_tot[kWaves] + s_tot[wave];
    const bool any = __shfl(incl, 63) != 0u;
    __syncthreads();  // s_tot is rewritten in the next round
    if (!any) continue;  // uniform over the wave
    // word r of the chunk (read from lane r): each lane tests its own bit,
    // so the word's set slots are written as one contiguous run
    const uint32_t ex = incl - pc;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t r = 0; r < 64u; ++r) {
      const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)word, r);
      const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(word >> 32), r);
      const unsigned long long wr = ((unsigned long long)hi << 32) | lo;
      if (wr == 0ull) continue;  // uniform over the wave
      const uint32_t off = __builtin_amdgcn_readlane(ex, r);
      if ((wr >> lane) & 1ull) list[base + off + (uint32_t)__popcll(wr & below)] = (ch * 64u + r) * 64u + lane;
    }
    if (word) bits[w] = 0ull;
  }
}

// k_walls moves the listed particles, one per thread (a grid-stride loop
// over the list); the 64 lanes of a wavefront fly their particles together,
// testing the triangle entries of all their cells as one pool. The pass holds no RNG state (Philox draws keyed by slot
// and step). Wall forces and heat are summed in 64-bit fixed point, so the
// order in which particles are listed and run does not change the sums;
// each workgroup stores its sums in its own slot of wall_part, and
// k_wall_sum adds them to wall_acc.
__global__ void __launch_bounds__(kTileThreads, 4) k_walls(const InstDesc* __restrict__ Ip, uint32_t opar,
                                                       uint32_t flags) {
  const InstDesc& I = Ip[blockIdx.y];
  if (!I.defer_bits || blockIdx.x >= I.n_wall_wg) return;
  __shared__ long long s_red[kTileThreads / 64][4];
  __shared__ unsigned long long s_hit[kTileThreads / 64][64];
  const uint32_t cnt = I.ctr[5];
  const uint32_t stepno = I.ctr[4];
  const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid / 64;
  const gptr<float> px = gmem(I.px[opar]);
  const gptr<float> py = gmem(I.py[opar]);
  const gptr<float> pz = gmem(I.pz[opar]);
  const gptr<float> vx = gmem(I.vx[opar]);
  const gptr<float> vy = gmem(I.vy[opar]);
  const gptr<float> vz = gmem(I.vz[opar]);
  const gptr<uint32_t> key = gmem(I.key[opar]);
  const gptr<const uint32_t> list = gmem<const uint32_t>(I.defer_list);
  WallAcc acc = {0, 0, 0, 0};
  const WaveHitLds hl = {s_hit[wid]};
  // uniform trip count over the workgroup: the wall flights run the
  // wavefront's lanes together, and the counting is wave-aggregated
  // (neighbouring list entries mostly land in the same cell)
  for (uint32_t b = blockIdx.x * kTileThreads; b < cnt; b += I.n_wall_wg * kTileThreads) {
    const uint32_t i = b + tid;
    const bool act = i < cnt;
    uint32_t slot = 0;
    float x = 0.0f, y = 0.0f, z = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
    if (act) {
      slot = list[i];
      x = px[slot], y = py[slot], z = pz[slot];
      ux = vx[slot], uy = vy[slot], uz = vz[slot];
    }
    uint32_t hits;
    const uint32_t k = move_walls(I, act, x, y, z, ux, uy, uz, I.dt, slot, stepno, hits, acc, hl);
    if (act) {
      px[slot] = x;
      py[slot] = y;
      pz[slot] = z;
      vx[slot] = ux;
      vy[slot] = uy;
      vz[slot] = uz;
      key[slot] = k == kInvalidKey ? k : (k | kWallKeyBit);
      if (flags & kFlagHits) gmem(I.hits)[slot] = hits;
    }
    count_keys(gmem(I.count), k);
  }
  long long v[4] = {acc.fx, acc.fy, acc.fz, acc.heat};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    long long x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if (lane == 0) s_red[wid][k] = x;
  }
  __syncthreads();
  if (tid < 4) {  // the workgroup's own slot: plain stores, no contention
    long long t = 0;
    for (uint32_t q = 0; q < kTileThreads / 64; ++q) t += s_red[q][tid];
    gmem(I.wall_part)[blockIdx.x * 4u + tid] = t;
  }
}

// One workgroup per instance folds the wall pass's per-workgroup sums into
// the running totals: exact 64-bit integer sums (any order gives the same
// value), converted once per step.
__global__ void __launch_bounds__(kTileThreads) k_wall_sum(const InstDesc* __restrict__ Ip) {
  const InstDesc& I = Ip[blockIdx.y];
  if (!I.defer_bits) return;
  __shared__ long long s_red[kTileThreads / 64][4];
  const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid / 64;
  const gptr<const long long> part = gmem<const long long>(I.wall_part);
  long long v[4] = {0, 0, 0, 0};
  for (uint32_t w = tid; w < I.n_wall_wg; w += kTileThreads)
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] += part[w * 4u + k];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    long long x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if (lane == 0) s_red[wid][k] = x;
  }
  __syncthreads();
  if (tid < 4) {
    long long t = 0;
    for (uint32_t q = 0; q < kTileThreads / 64; ++q) t += s_red[q][tid];
    gmem(I.wall_acc)[tid] += (double)t * I.wall_unit[tid < 3 ? 0 : 1];
  }
}

// =================================================================== loading
// GenerateInitialVelocitiesDevBlocks / GenerateUniformVelocitiesDevBlocks
// (dsmc_kernel.cu:574-680): one thread per (x, y, i) column, i < ppc, walking
// z; stream X*Y*i + y*X + x; particle index = cell*ppc + i. Cells with a
// flagged collision-grid corner get parked particles (pos = vel = 0).
__device__ __forceinline__ bool geometry_in_cell(const InstDesc& I, float cx, float cy, float cz, float sx, float sy,
                                                 float sz) {
  // IsGeometryInCell (dsmc_kernel.cu:479-546), corners in the reference's order
  float px = cx, py = cy, pz = cz;
  uint32_t c;
#define DSMC_CORNER()                                                  \
  if ((c = cg_cell(I, px, py, pz)) != kInvalidKey && I.cg_flag[c] != 0u) \
    return true;
  DSMC_CORNER();
  px = px + sx;
  DSMC_CORNER();
  py = py + sy;
  DSMC_CORNER();
  pz = pz + sz;
  DSMC_CORNER();
  px = cx; py = cy; pz = cz;
  py = py + sy;
  DSMC_CORNER();
  pz = pz + sz;
  DSMC_CORNER();
  py = cy;
  DSMC_CORNER();
  px = px + sx;
  DSMC_CORNER();
#undef DSMC_CORNER
  return false;
}

__global__ void k_load_reference(const InstDesc* __restrict__ Ip, uint32_t par, uint32_t ppc, float sux, float suy,
                                 float suz, float temp) {
  const InstDesc& I = *Ip;
  const uint32_t X = I.dim[0], Y = I.dim[1], Z = I.dim[2];
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= X * Y * ppc) return;
  const uint32_t i = t / (X * Y);
  const uint32_t xy = t - i * X * Y;
  const uint32_t cy = xy / X, cx = xy - cy * X;
  const uint32_t stream = t;  // = X*Y*i + y*X + x
  Mt rng;
  mt_load(I, stream, rng);
  float cellx = I.gmin[0] + (float)cx * I.cell[0];
  float celly = I.gmin[1] + (float)cy * I.cell[1];
  const float twopi = 2.0f * kPiRef;
  for (uint32_t z = 0; z < Z; ++z) {
    uint32_t idx = (X * Y * z + cy * X + cx) * ppc + i;
    float cellz = I.gmin[2] + (float)z * I.cell[2];
    bool geom = I.has_geom && geometry_in_cell(I, cellx, celly, cellz, I.cell[0], I.cell[1], I.cell[2]);
    if (!geom) {
      // GetRandomVelocity / SampleDensityFunction (dsmc_kernel.cu:394-415)
      float vv = sqrtf(2.0f * kBoltzRef * temp / I.mass * (-dm_logf(mt_float(rng))));
      float vx = vv * dm_sinf(mt_float(rng) * twopi) + sux;
      float vy = vv * dm_sinf(mt_float(rng) * twopi) + suy;
      float vz = vv * dm_sinf(mt_float(rng) * twopi) + suz;
      // GetRandomPosition (dsmc_kernel.cu:417-427)
      float px = cellx + mt_float(rng) * I.cell[0];
      float py = celly + mt_float(rng) * I.cell[1];
      float pz = cellz + mt_float(rng) * I.cell[2];
      I.px[par][idx] = px; I.py[par][idx] = py; I.pz[par][idx] = pz;
      I.vx[par][idx] = vx; I.vy[par][idx] = vy; I.vz[par][idx] = vz;
    } else {
      I.px[par][idx] = 0.0f; I.py[par][idx] = 0.0f; I.pz[par][idx] = 0.0f;
      I.vx[par][idx] = 0.0f; I.vy[par][idx] = 0.0f; I.vz[par][idx] = 0.0f;
    }
  }
  mt_store(I, stream, rng);
}

// Move with dt = 0 after loading (include/gpu_dsmc.inl:79): box functor and
// walls on the loaded state, in place.
__global__ void k_move_zero(const InstDesc* __restrict__ Ip, uint32_t par, uint32_t n) {
  const InstDesc& I = *Ip;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x = I.px[par][i], y = I.py[par][i], z = I.pz[par][i];
  float ux = I.vx[par][i], uy = I.vy[par][i], uz = I.vz[par][i];
  move_reference(I, x, y, z, ux, uy, uz, 0.0f);
  I.px[par][i] = x; I.py[par][i] = y; I.pz[par][i] = z;
  I.vx[par][i] = ux; I.vy[par][i] = uy; I.vz[par][i] = uz;
}

// PHYSICS loading, pass 1: particles per cell by stochastic rounding of
// n * V_fluid / Fn (per-cell density when fields are given), plus a uniform
// random NTC remainder. Stream = cell mod n_streams, cells in increasing order.
__global__ void k_load_count(const InstDesc* __restrict__ Ip, const float* __restrict__ fn) {
  const InstDesc& I = *Ip;
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= I.n_streams) return;
  Mt rng;
  mt_load(I, s, rng);
  for (uint32_t c = s; c < I.ncells; c += I.n_streams) {
    float dens = fn ? fn[c] : I.n_inf;
    float frac = I.fluid_frac ? I.fluid_frac[c] : 1.0f;
    float lam = (float)((double)dens * (double)I.cell_volume * (double)frac / (double)I.fnum);
    float r = mt_float(rng);
    uint32_t k = frac > 0.0f ? dm_f2u_sat(lam + r) : 0u;
    I.count[c] = k;
    float2 nt = I.ntc[c];
    nt.y = mt_float(rng);
    I.ntc[c] = nt;
  }
  mt_store(I, s, rng);
}

// PHYSICS loading, pass 2: positions uniform in the cell (rejecting points in
// the body), velocities Maxwellian (Box-Muller) about the stream velocity.
__global__ void k_load_generate(const InstDesc* __restrict__ Ip, uint32_t par, const float* __restrict__ fu,
                                const float* __restrict__ fv, const float* __restrict__ fw,
                                const float* __restrict__ fT) {
  const InstDesc& I = *Ip;
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= I.n_streams) return;
  Mt rng;
  mt_load(I, s, rng);
  const float twopi = 6.28318530717958647692f;
  const uint32_t X = I.dim[0], Y = I.dim[1];
  for (uint32_t c = s; c < I.ncells; c += I.n_streams) {
    uint32_t off = I.cell_start[c], k = I.cell_start[c + 1] - off;
    if (k == 0) continue;
    uint32_t cz = c / (X * Y), rem = c - cz * X * Y, cy = rem / X, cx = rem - cy * X;
    float bx = I.gmin[0] + (float)cx * I.cell[0];
    float by = I.gmin[1] + (float)cy * I.cell[1];
    float bz = I.gmin[2] + (float)cz * I.cell[2];
    float ux = fu ? fu[c] : I.U[0], uy = fv ? fv[c] : I.U[1], uz = fw ? fw[c] : I.U[2];
    float T = fT ? fT[c] : I.T_inf;
    float sd = sqrtf((float)(kBoltz * (double)T / (double)I.mass));
    bool cut = I.fluid_frac && I.fluid_frac[c] < 1.0f;
    for (uint32_t j = 0; j < k; ++j) {
      float px, py, pz;
      int tries = 0;
      do {
        px = bx + mt_float(rng) * I.cell[0];
        py = by + mt_float(rng) * I.cell[1];
        pz = bz + mt_float(rng) * I.cell[2];
        ++tries;
      } while (cut && tries < 64 && inside_body(I, px, py, pz));
      float r1 = mt_float(rng), r2 = mt_float(rng), r3 = mt_float(rng), r4 = mt_float(rng);
      float a1 = sqrtf(-2.0f * dm_logf(r1)), a2 = sqrtf(-2.0f * dm_logf(r3));
      float s1, c1, s2, c2;
      dm_sincosf(twopi * r2, &s1, &c1);
      dm_sincosf(twopi * r4, &s2, &c2);
      uint32_t idx = off + j;
      I.px[par][idx] = px; I.py[par][idx] = py; I.pz[par][idx] = pz;
      I.vx[par][idx] = ux + sd * a1 * c1;
      I.vy[par][idx] = uy + sd * a1 * s1;
      I.vz[par][idx] = uz + sd * a2 * c2;
      I.key[par][idx] = c;
    }
  }
  mt_store(I, s, rng);
}

// ================================================================= injection
// Freestream inflow through open faces (PHYSICS): face cell g of face f
// injects floor(lambda_f + R) particles with the flux-weighted half-range
// Maxwellian normal velocity (acceptance-rejection on the speed-ratio
// density z exp(-(z - s)^2), Bird 1994 sec. 12.1) and Gaussian tangential
// components, placed on the face and advanced a random fraction of dt.
__device__ __forceinline__ uint32_t face_of(const InstDesc& I, uint32_t g) {
  uint32_t f = 0;
  while (f < 5 && g >= I.face_off[f + 1]) ++f;
  return f;
}

// Injection draws Philox4x32-10 under key (seed, kInjectKey1), so no stream
// state moves and every particle is independent: the count of face cell g
// draws counter (g, step, 0, 0); particle j of g draws (g, step, 1 + j, b)
// for its b-th rejection block (two tries per block) and (g, step, 1 + j,
// 2^31 + k) for its other variates.
constexpr uint32_t kInjectKey1 = 0xBB67AE85u;

__global__ void k_inject_count(const InstDesc* __restrict__ Ip) {
  const InstDesc& I = Ip[blockIdx.y];
  const uint32_t stepno = I.ctr[4];
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < I.n_face_cells; g += gridDim.x * blockDim.x) {
    const dsmc_rng::U32x4 c = {{g, stepno, 0u, 0u}};
    const dsmc_rng::U32x4 r = dsmc_rng::philox4x32_10(c, I.wall_key[0], kInjectKey1);
    I.face_count[g] = dm_f2u_sat(I.inj_lambda[face_of(I, g)] + dsmc_rng::px_unit(r.v[0]));
  }
}

// One thread per injected particle t (grid-stride): its face cell is the
// last g with face_count[g] <= t (face_count holds the exclusive offsets),
// it is particle t - face_count[g] of that cell, and it goes to slot
// sorted count + t.
__global__ void k_inject_generate(const InstDesc* __restrict__ Ip, uint32_t opar) {
  const InstDesc& I = Ip[blockIdx.y];
  const uint32_t F = I.n_face_cells;
  if (F == 0) return;
  const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t base = I.ctr[1];
  const uint32_t tot = I.face_count[F];
  if (t0 == 0) {
    uint32_t end = base + tot;
    if (end > I.cap || end < base) {  // capacity exceeded: drop the excess, flag it
      end = I.cap;
      atomicOr(&I.ctr[3], kStatusCapacity);
    }
    I.ctr[0] = end;
    I.ctr[2] += end - base;
  }
  const uint32_t stepno = I.ctr[4];
  const uint32_t k0 = I.wall_key[0];
  const gptr<const uint32_t> fc = gmem<const uint32_t>(I.face_count);
  const float twopi = 6.28318530717958647692f;
  const float vm = I.vm_inf;
  const float sd = vm * 0.70710678118654752f;  // sqrt(kT/m)
  for (uint32_t t = t0; t < tot; t += gridDim.x * blockDim.x) {
    const uint32_t idx = base + t;
    if (idx >= I.cap || idx < base) break;  // later t only go further
    uint32_t lo = 0u, hi = F;  // fc[lo] <= t < fc[hi]
    while (hi - lo > 1u) {
      const uint32_t mid = lo + (hi - lo) / 2u;
      if (fc[mid] <= t) lo = mid;
      else hi = mid;
    }
    const uint32_t g = lo, j = t - fc[g];
    const uint32_t f = face_of(I, g);
    const uint32_t local = g - I.face_off[f];
    const uint32_t axis = f >> 1;           // normal axis
    const uint32_t ta = axis == 0 ? 1u : 0u;  // first tangential axis
    const uint32_t tb = axis == 2 ? 1u : 2u;  // second tangential axis
    const uint32_t na = I.dim[ta];
    const uint32_t ia = local % na, ib = local / na;
    const float sign = (f & 1u) ? -1.0f : 1.0f;  // inward normal direction
    const float un = sign * I.U[axis];           // inward stream component
    const float sr = un / vm;                    // speed ratio of the inflow
    float zlo = sr - 5.0f;
    zlo = zlo > 0.0f ? zlo : 0.0f;
    float zhi = sr + 5.0f;
    zhi = zhi > 1.0e-3f ? zhi : 1.0e-3f;
    float zs = 0.5f * (sr + sqrtf(sr * sr + 2.0f));  // mode of z exp(-(z-s)^2)
    zs = zs < zlo ? zlo : (zs > zhi ? zhi : zs);
    const float gmax = zs * dm_expf(-(zs - sr) * (zs - sr));
    float z = zlo;
    bool ok = false;
    for (uint32_t b = 0; b < 500u && !ok; ++b) {
      const dsmc_rng::U32x4 c = {{g, stepno, 1u + j, b}};
      const dsmc_rng::U32x4 r = dsmc_rng::philox4x32_10(c, k0, kInjectKey1);
      for (int h = 0; h < 2 && !ok; ++h) {
        z = zlo + (zhi - zlo) * dsmc_rng::px_unit(r.v[2 * h]);
        const float gz = z * dm_expf(-(z - sr) * (z - sr));
        ok = gz > gmax * dsmc_rng::px_unit(r.v[2 * h + 1]);
      }
    }
    const dsmc_rng::U32x4 ca = {{g, stepno, 1u + j, 0x80000000u}};
    const dsmc_rng::U32x4 ra = dsmc_rng::philox4x32_10(ca, k0, kInjectKey1);
    const dsmc_rng::U32x4 cb = {{g, stepno, 1u + j, 0x80000001u}};
    const dsmc_rng::U32x4 rb = dsmc_rng::philox4x32_10(cb, k0, kInjectKey1);
    const float vn = z * vm;
    const float a1 = sqrtf(-2.0f * dm_logf(dsmc_rng::px_unit(ra.v[0])));
    float s1, c1;
    dm_sincosf(twopi * dsmc_rng::px_unit(ra.v[1]), &s1, &c1);
    float p[3], v[3];
    v[axis] = sign * vn;
    v[ta] = I.U[ta] + sd * a1 * c1;
    v[tb] = I.U[tb] + sd * a1 * s1;
    p[axis] = (f & 1u) ? I.gmax[axis] : I.gmin[axis];
    p[ta] = I.gmin[ta] + ((float)ia + dsmc_rng::px_unit(ra.v[2])) * I.cell[ta];
    p[tb] = I.gmin[tb] + ((float)ib + dsmc_rng::px_unit(ra.v[3])) * I.cell[tb];
    const float fr = dsmc_rng::px_unit(rb.v[0]) * I.dt;
    float x = p[0] + v[0] * fr, y = p[1] + v[1] * fr, zz = p[2] + v[2] * fr;
    x = x < I.gmin[0] ? I.gmin[0] : (x > I.gmax[0] ? I.gmax[0] : x);
    y = y < I.gmin[1] ? I.gmin[1] : (y > I.gmax[1] ? I.gmax[1] : y);
    zz = zz < I.gmin[2] ? I.gmin[2] : (zz > I.gmax[2] ? I.gmax[2] : zz);
    I.px[opar][idx] = x; I.py[opar][idx] = y; I.pz[opar][idx] = zz;
    I.vx[opar][idx] = v[0]; I.vy[opar][idx] = v[1]; I.vz[opar][idx] = v[2];
    const uint32_t key = cell_key_phys(I, x, y, zz);
    I.key[opar][idx] = key;
    gatomic_add(gmem(I.count) + key, 1u);
  }
}

// ================================================================== utility
// Count particles with non-finite state or outside the box.
__global__ void k_validate(const InstDesc* __restrict__ Ip, uint32_t par, uint32_t n, uint32_t* bad) {
  const InstDesc& I = *Ip;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (I.key[par][i] == kInvalidKey) return;
  float x = I.px[par][i], y = I.py[par][i], z = I.pz[par][i];
  float a = I.vx[par][i], b = I.vy[par][i], c = I.vz[par][i];
  bool ok = (x - x == 0.0f) && (y - y == 0.0f) && (z - z == 0.0f) && (a - a == 0.0f) && (b - b == 0.0f) &&
            (c - c == 0.0f);
  if (I.mode == kModePhysics)
    ok = ok && x >= I.gmin[0] && x <= I.gmax[0] && y >= I.gmin[1] && y <= I.gmax[1] && z >= I.gmin[2] &&
         z <= I.gmax[2];
  if (!ok) atomicAdd(bad, 1u);
}

// One extra sample of the current state (SampleCellStatistics outside the
// step, dsmc_kernel.cu:892-910): cells of the last ordering, current velocities.
__global__ void k_sample_cells(const InstDesc* __restrict__ Ip, uint32_t par, uint32_t mode) {
  const InstDesc& I = *Ip;
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= I.ncells) return;
  uint32_t a = I.cell_start[c], b = I.cell_start[c + 1];
  if (b <= a) return;
  const uint32_t C = I.ncells;
  const float* vx = I.vx[par];
  const float* vy = I.vy[par];
  const float* vz = I.vz[par];
  if (mode == kModeReference) {
    float* S = reinterpret_cast<float*>(I.samp);
    float sn = S[c], sx = S[C + c], sy = S[2 * C + c], sz = S[3 * C + c], s2 = S[4 * C + c];
    for (uint32_t q = a; q < b; ++q) {
      float ux = vx[q], uy = vy[q], uz = vz[q];
      sn += 1.0f; sx += ux; sy += uy; sz += uz;
      s2 += ux * ux + uy * uy + uz * uz;
    }
    S[c] = sn; S[C + c] = sx; S[2 * C + c] = sy; S[3 * C + c] = sz; S[4 * C + c] = s2;
  } else {
    double* S = I.samp;
    double sn = S[c], sx = S[C + c], sy = S[2 * C + c], sz = S[3 * C + c], s2 = S[4 * C + c];
    for (uint32_t q = a; q < b; ++q) {
      double ux = vx[q], uy = vy[q], uz = vz[q];
      sn += 1.0; sx += ux; sy += uy; sz += uz;
      s2 += ux * ux + uy * uy + uz * uz;
    }
    S[c] = sn; S[C + c] = sx; S[2 * C + c] = sy; S[3 * C + c] = sz; S[4 * C + c] = s2;
  }
}

// ================================================================ launchers
namespace dsmc_launch {

static inline uint32_t blocks_for(uint32_t n, uint32_t t) { return (n + t - 1u) / t; }

void mt_seed(uint32_t* state, const uint32_t* seeds, uint32_t n_streams, hipStream_t st) {
  if (!n_streams) return;
  hipLaunchKernelGGL(k_mt_seed, dim3(blocks_for(n_streams, 256)), dim3(256), 0, st, state, seeds, n_streams);
}

void mt_generate(const uint32_t* a, const uint32_t* b, const uint32_t* c, const uint32_t* seeds, uint32_t n_streams,
                 uint32_t n_draws, uint32_t* out, hipStream_t st) {
  if (!n_streams) return;
  hipLaunchKernelGGL(k_mt_generate, dim3(blocks_for(n_streams, 256)), dim3(256), 0, st, a, b, c, seeds, n_streams,
                     n_draws, out);
}

void cell_ids(const InstDesc& I, const float* x, const float* y, const float* z, uint32_t n, uint32_t* keys,
              hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_cell_ids, dim3(blocks_for(n, 256)), dim3(256), 0, st, I, x, y, z, n, keys);
}

void keys_count(const InstDesc* d, uint32_t par, uint32_t n, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_keys_count, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, par, n);
}

int scan(uint32_t* in, uint32_t n, uint32_t* out, uint32_t* cursor, bool clear_in, uint32_t* scratch,
         uint32_t* total_out, hipStream_t st) {
  uint32_t tiles = blocks_for(n > 0 ? n : 1u, kScanTile);
  if (tiles > kScanMaxTiles) return -1;
  uint32_t* total = total_out ? total_out : scratch + kScanMaxTiles;
  hipLaunchKernelGGL(k_scan_reduce, dim3(tiles), dim3(kScanThreads), 0, st, in, n, scratch);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, st, scratch, tiles, total);
  hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(kScanThreads), 0, st, in, n, scratch, out, cursor,
                     clear_in ? 1u : 0u, total);
  return 0;
}

void scatter(const InstDesc* d, uint32_t par, uint32_t grid, hipStream_t st, uint32_t n_inst) {
  hipLaunchKernelGGL(k_scatter, dim3(grid, n_inst), dim3(256), 0, st, d, par);
}

void step_scan(const InstDesc* d, uint32_t n_inst, uint32_t which, uint32_t max_n, uint32_t* scratch,
               hipStream_t st) {
  uint32_t tiles = blocks_for(max_n > 0 ? max_n : 1u, kScanTile);
  hipLaunchKernelGGL(k_bscan_reduce, dim3(tiles, n_inst), dim3(kScanThreads), 0, st, d, which, scratch);
  hipLaunchKernelGGL(k_bscan_tiles, dim3(1, n_inst), dim3(kScanThreads), 0, st, d, which, scratch);
  hipLaunchKernelGGL(k_bscan_apply, dim3(tiles, n_inst), dim3(kScanThreads), 0, st, d, which, scratch);
}

uint64_t op_sort_work_words(uint32_t n, uint32_t n_cells) {
  const uint64_t nb = blocks_for(n > 0 ? n : 1u, kRsTile);
  return (uint64_t)n_cells + 2ull * (256ull * nb + 1ull) + (kScanMaxTiles + 1ull) + 3ull * n;
}

int op_sort(const uint32_t* keys, uint32_t n, uint32_t n_cells, uint32_t* perm, uint32_t* cell_start,
            uint32_t* work, hipStream_t st) {
  const uint32_t nb = blocks_for(n > 0 ? n : 1u, kRsTile);
  if ((uint64_t)n_cells > (uint64_t)kScanMaxTiles * kScanTile || 256ull * nb > (uint64_t)kScanMaxTiles * kScanTile)
    return -1;
  uint32_t* cnt = work;
  uint32_t* hist = cnt + n_cells;
  uint32_t* offs = hist + (256u * nb + 1u);
  uint32_t* scr = offs + (256u * nb + 1u);
  uint32_t* ka = scr + (kScanMaxTiles + 1u);
  uint32_t* kb = ka + n;
  uint32_t* vt = kb + n;
  // cell offsets: counts of the valid keys and their exclusive scan
  if (n_cells) hipLaunchKernelGGL(k_fill_u32, dim3(std::min<uint32_t>(blocks_for(n_cells, 256), 4096u)), dim3(256),
                                  0, st, cnt, n_cells, 0u);
  if (n) hipLaunchKernelGGL(k_count_keys, dim3(blocks_for(n, 256)), dim3(256), 0, st, keys, n, n_cells, cnt);
  if (scan(cnt, n_cells, cell_start, nullptr, true, scr, nullptr, st)) return -1;
  if (!n) return 0;
  // radix passes over the bits of n_cells (the largest mapped key)
  uint32_t bits = 1;
  while (bits < 32 && (n_cells >> bits) != 0u) ++bits;
  const uint32_t passes = (bits + 7u) / 8u;
  const uint32_t* kin = keys;
  const uint32_t* vin = nullptr;
  for (uint32_t p = 0; p < passes; ++p) {
    const uint32_t raw = p == 0 ? 1u : 0u;
    uint32_t* kout = (p & 1u) ? kb : ka;
    uint32_t* vout = ((passes - 1u - p) & 1u) ? vt : perm;  // the last pass writes perm
    hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kRsThreads), 0, st, kin, n, n_cells, raw, 8u * p, nb, hist);
    if (scan(hist, 256u * nb, offs, nullptr, false, scr, nullptr, st)) return -1;
    hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kRsThreads), 0, st, kin, vin, n, n_cells, raw, 8u * p, nb,
                       offs, kout, vout);
    kin = kout;
    vin = vout;
  }
  return 0;
}

void walls(const InstDesc* d, uint32_t opar, uint32_t flags, uint32_t grid, hipStream_t st, uint32_t n_inst) {
  hipLaunchKernelGGL(k_defer_list, dim3(grid, n_inst), dim3(kTileThreads), 0, st, d);
  hipLaunchKernelGGL(k_walls, dim3(grid, n_inst), dim3(kTileThreads), 0, st, d, opar, flags);
  hipLaunchKernelGGL(k_wall_sum, dim3(1, n_inst), dim3(kTileThreads), 0, st, d);
}

void tile(uint32_t mode, const InstDesc* d, uint32_t par, uint32_t flags, uint32_t n_vwg, hipStream_t st,
          uint32_t n_inst) {
  if (mode == kModeReference)
    hipLaunchKernelGGL(k_tile<kModeReference>, dim3(n_vwg, n_inst), dim3(kTileThreads), 0, st, d, par, flags);
  else
    hipLaunchKernelGGL(k_tile<kModePhysics>, dim3(n_vwg, n_inst), dim3(kTileThreads), 0, st, d, par, flags);
}

void load_reference(const InstDesc* d, uint32_t par, uint32_t ppc, uint32_t n_threads, float sux, float suy,
                    float suz, float temp, hipStream_t st) {
  if (!n_threads) return;
  hipLaunchKernelGGL(k_load_reference, dim3(blocks_for(n_threads, 128)), dim3(128), 0, st, d, par, ppc, sux, suy,
                     suz, temp);
}

void move_zero(const InstDesc* d, uint32_t par, uint32_t n, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_move_zero, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, par, n);
}

void load_count(const InstDesc* d, uint32_t n_streams, const float* fn, hipStream_t st) {
  hipLaunchKernelGGL(k_load_count, dim3(blocks_for(n_streams, 256)), dim3(256), 0, st, d, fn);
}

void load_generate(const InstDesc* d, uint32_t par, uint32_t n_streams, const float* fu, const float* fv,
                   const float* fw, const float* fT, hipStream_t st) {
  hipLaunchKernelGGL(k_load_generate, dim3(blocks_for(n_streams, 256)), dim3(256), 0, st, d, par, fu, fv, fw, fT);
}

void inject_count(const InstDesc* d, uint32_t n_threads, hipStream_t st, uint32_t n_inst) {
  if (!n_threads) return;
  hipLaunchKernelGGL(k_inject_count, dim3(blocks_for(n_threads, 256), n_inst), dim3(256), 0, st, d);
}

void inject_generate(const InstDesc* d, uint32_t opar, uint32_t n_threads, hipStream_t st, uint32_t n_inst) {
  if (!n_threads) return;
  hipLaunchKernelGGL(k_inject_generate, dim3(blocks_for(n_threads, 256), n_inst), dim3(256), 0, st, d, opar);
}

void sample_cells(const InstDesc* d, uint32_t par, uint32_t mode, uint32_t ncells, hipStream_t st) {
  if (!ncells) return;
  hipLaunchKernelGGL(k_sample_cells, dim3(blocks_for(ncells, 256)), dim3(256), 0, st, d, par, mode);
}

void validate(const InstDesc* d, uint32_t par, uint32_t n, uint32_t* bad, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_validate, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, par, n, bad);
}

}  // namespace dsmc_launch
