// C entry points over the reference's own CPU sources (TEST INFRASTRUCTURE).
//
// Built by oracle/Makefile into oracle/_ref/libdsmc_ref.so from the files
// where they lie under /root/reference (src/genmtrand.c, src/aabb.cpp with
// include/tritest.h, src/reggrid.cpp, src/vector.cpp); nothing of theirs is
// copied here. Used only to pin the oracle and the engine's host geometry:
//   ref_genrand      Matsumoto-Nishimura genrand_mt (src/genmtrand.c:45-80)
//   ref_tri_box      AABB_t::tri_overlaps -> triBoxOverlap (src/aabb.cpp:38-50)
//   ref_create_grid  the loop of Model::CreateGrid (src/model.cpp:733-769) on
//                    reg_grid_t (src/reggrid.cpp:27-93), with encloseGrid
#include <cstdint>
#include <vector>

#include "aabb.h"
#include "reggrid.h"

extern "C" {
#include "dci.h"
}

extern "C" void ref_genrand(uint32_t aaa, uint32_t mask_b, uint32_t mask_c, uint32_t seed, uint32_t n, uint32_t* out) {
  uint32_t state[19];
  mt_struct m;
  m.aaa = aaa;
  m.mm = 9;
  m.nn = 19;
  m.rr = 1;
  m.ww = 32;
  m.wmask = 0xFFFFFFFFu;
  m.umask = 0xFFFFFFFEu;
  m.lmask = 0x1u;
  m.shift0 = 12;
  m.shift1 = 18;
  m.shiftB = 7;
  m.shiftC = 15;
  m.maskB = mask_b;
  m.maskC = mask_c;
  m.state = state;
  sgenrand_mt(seed, &m);
  for (uint32_t i = 0; i < n; ++i) out[i] = genrand_mt(&m);
}

extern "C" int ref_tri_box(const float c[3], const float h[3], const float a[3], const float b[3], const float cc[3]) {
  AABB_t box(vec3_t(c[0], c[1], c[2]), vec3_t(h[0], h[1], h[2]));
  return box.tri_overlaps(vec3_t(a[0], a[1], a[2]), vec3_t(b[0], b[1], b[2]), vec3_t(cc[0], cc[1], cc[2])) ? 1 : 0;
}

// bbox6 = min xyz, max xyz of the grid; outputs per cell flag (0/1), the
// triangle count per cell and the concatenated triangle lists.
extern "C" int ref_create_grid(const float* verts, const uint32_t* tris, uint32_t nt, const float bbox6[6], int gx,
                               int gy, int gz, float cell_out[3], uint32_t* flags, uint32_t* counts,
                               uint32_t* entries, uint32_t cap_entries, uint32_t* n_entries) {
  reg_grid_t grid;
  bbox_t bb(vec3_t(bbox6[0], bbox6[1], bbox6[2]), vec3_t(bbox6[3], bbox6[4], bbox6[5]));
  grid.reset(bb, vec3i_t(gx, gy, gz));
  vec3i_t idx;
  for (idx.x = 0; idx.x < gx; idx.x++)
    for (idx.y = 0; idx.y < gy; idx.y++)
      for (idx.z = 0; idx.z < gz; idx.z++) {
        AABB_t aabb = grid.getAABB(idx);
        uint32_t ci = grid.getCellIdx(idx);
        for (uint32_t t = 0; t < nt; ++t) {
          const float* a = &verts[3 * tris[3 * t]];
          const float* b = &verts[3 * tris[3 * t + 1]];
          const float* c = &verts[3 * tris[3 * t + 2]];
          if (aabb.tri_overlaps(vec3_t(a[0], a[1], a[2]), vec3_t(b[0], b[1], b[2]), vec3_t(c[0], c[1], c[2]))) {
            grid[ci].first.push_back(t);
            grid[ci].second = true;
          }
        }
      }
  grid.encloseGrid();
  cell_out[0] = grid.cellSize.x;
  cell_out[1] = grid.cellSize.y;
  cell_out[2] = grid.cellSize.z;
  uint32_t k = 0;
  for (uint32_t i = 0; i < grid.cellCount(); ++i) {
    flags[i] = grid[i].second ? 1u : 0u;
    counts[i] = (uint32_t)grid[i].first.size();
    for (uint32_t t : grid[i].first) {
      if (k < cap_entries) entries[k] = t;
      ++k;
    }
  }
  *n_entries = k;
  return k <= cap_entries ? 0 : -1;
}
