// Host launch wrappers of the DSMC kernels (implemented in step.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../common/dsmc_dev.h"

namespace dsmc_launch {

using dsmc_dev::InstDesc;

constexpr uint32_t kScanMaxTiles = 4096;  // scan of up to 4096*4096 entries

void mt_seed(uint32_t* state, const uint32_t* seeds, uint32_t n_streams, hipStream_t st);
void mt_generate(const uint32_t* a, const uint32_t* b, const uint32_t* c, const uint32_t* seeds, uint32_t n_streams,
                 uint32_t n_draws, uint32_t* out, hipStream_t st);
void cell_ids(const InstDesc& I, const float* x, const float* y, const float* z, uint32_t n, uint32_t* keys,
              hipStream_t st);
void keys_count(const InstDesc* d, uint32_t par, uint32_t n, hipStream_t st);
// Exclusive scan of in[0..n) into out[0..n] (out[n] = total); cursor gets a
// copy of the offsets when non-null; in is cleared when clear_in; total_out
// (device) receives the total. scratch: kScanMaxTiles + 1 uint32.
int scan(uint32_t* in, uint32_t n, uint32_t* out, uint32_t* cursor, bool clear_in, uint32_t* scratch,
         uint32_t* total_out, hipStream_t st);
// Step kernels take an array of n_inst descriptors (grid.y = instance).
void scatter(const InstDesc* d, uint32_t par, uint32_t grid, hipStream_t st, uint32_t n_inst = 1);
// Scan of the step: which = 0 cell counts, 1 injection face counts; max_n =
// largest count array of the batch; scratch: n_inst * (kScanMaxTiles + 1).
void step_scan(const InstDesc* d, uint32_t n_inst, uint32_t which, uint32_t max_n, uint32_t* scratch,
               hipStream_t st);
// Stable sort operator: cell offsets (n_cells + 1) and the permutation of a
// stable sort by key (keys >= n_cells dropped, listed last). work:
// op_sort_work_words(n, n_cells) uint32. No allocation, no memset.
uint64_t op_sort_work_words(uint32_t n, uint32_t n_cells);
int op_sort(const uint32_t* keys, uint32_t n, uint32_t n_cells, uint32_t* perm, uint32_t* cell_start,
            uint32_t* work, hipStream_t st);
void tile(uint32_t mode, const InstDesc* d, uint32_t par, uint32_t flags, uint32_t n_vwg, hipStream_t st,
          uint32_t n_inst = 1);
// PHYSICS wall pass over the tiles the tile kernel queued (opar = the
// buffer the tile kernel wrote); `grid` workgroups per instance.
void walls(const InstDesc* d, uint32_t opar, uint32_t flags, uint32_t grid, hipStream_t st, uint32_t n_inst = 1);
void load_reference(const InstDesc* d, uint32_t par, uint32_t ppc, uint32_t n_threads, float sux, float suy,
                    float suz, float temp, hipStream_t st);
void move_zero(const InstDesc* d, uint32_t par, uint32_t n, hipStream_t st);
void load_count(const InstDesc* d, uint32_t n_streams, const float* fn, hipStream_t st);
void load_generate(const InstDesc* d, uint32_t par, uint32_t n_streams, const float* fu, const float* fv,
                   const float* fw, const float* fT, hipStream_t st);
void inject_count(const InstDesc* d, uint32_t n_threads, hipStream_t st, uint32_t n_inst = 1);
void inject_generate(const InstDesc* d, uint32_t opar, uint32_t n_threads, hipStream_t st, uint32_t n_inst = 1);
void sample_cells(const InstDesc* d, uint32_t par, uint32_t mode, uint32_t ncells, hipStream_t st);
void validate(const InstDesc* d, uint32_t par, uint32_t n, uint32_t* bad, hipStream_t st);

enum TileFlags : uint32_t { kFlagSample = 1u, kFlagKeepPerm = 2u, kFlagHits = 4u };

}  // namespace dsmc_launch
