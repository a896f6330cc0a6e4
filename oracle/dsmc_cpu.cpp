// dsmc_cpu.cpp - CPU oracle of the DSMC step. TEST INFRASTRUCTURE ONLY:
// loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline as
// the checker; the engine never links it.
//
// What it restates (reference file:line):
//   MT streams ........ kernels/mtwister.cu:41-116 (seed chain, Knuth init,
//                       DCMT recurrence and tempering), as the generic block
//                       form of Matsumoto-Nishimura (genrand_mt)
//   cell keys ......... kernels/dsmc_kernel.cu:1701-1724 (GetCellID3D)
//   sort + bounds ..... kernels/dsmc_kernel.cu:2033-2051 (stable sort_by_key),
//                       :1734-1787 (cellStartEnd, REORDER gather)
//   NTC collisions .... kernels/dsmc_kernel.cu:1002-1089 (hard spheres, MT)
//   move + walls ...... kernels/dsmc_kernel.cu:1256-1314, 1596-1657;
//                       kernels/tri_intersect.cu:14-56; kernels/vectors.cu
//   box functors ...... kernels/dsmc_kernel.cu:1316-1555
//   sampling .......... kernels/dsmc_kernel.cu:856-890
//   loading ........... kernels/dsmc_kernel.cu:394-427, 479-546, 629-680
//   step order ........ include/gpu_dsmc.inl:221-274
// and the PHYSICS extensions (VHS NTC, nearest-hit walls by brute force over
// all triangles, diffuse re-emission, freestream faces and injection).
//
// The oracle visits the tiles in the engine's locality order (sorted by
// (z / B, y / B, z mod B, y mod B, x) of their first cell, B the largest
// power of two with B^2 x <= 128 tile_cells; with G = n_streams / threads a
// multiple of 8, patches of 128 positions dealt to 8 "XCDs" and taken by
// workgroup w as elements w / 8, w / 8 + G / 8, ... of its XCD's patches,
// element e being position ((e / 128) * 8 + w mod 8) * 128 + e mod 128;
// else positions w, w + G, ...). No draw depends on that order; it only
// groups the per-workgroup double sums of the collision counters.
// PHYSICS collisions, injection and the wall re-emission draw counter-based
// Philox4x32-10 (keyed by cell/slot, step and item, so no draw depends on
// the schedule); PHYSICS loading draws from MT stream (cell mod n_streams).
// REFERENCE mode keeps the reference's own assignment: cell c collides with
// stream c (ds_mtState[cellid], dsmc_kernel.cu:1030-1031) and loading
// thread X*Y*i + y*X + x draws from that stream (:633). The oracle follows
// that assignment with plain loops.
#include "dsmc_cpu.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../cudadsmc_amd/csrc/common/detmath.h"
#include "../cudadsmc_amd/csrc/common/pairing.h"
#include "../cudadsmc_amd/csrc/common/philox.h"

namespace {

const uint32_t INV = 0xFFFFFFFFu;
const uint32_t DEFER = 0xFFFFFFFEu;  // PHYSICS: particle left for the wall pass
const float BOLZ_REF = 1.3806e-23f;
const double BOLZ = 1.380649e-23;
const float PI_REF = 3.14159265358979f;
// candidate pairs per cell and step beyond which the engine clamps and flags
const uint32_t kMaxPairs = 1u << 20;

// ----------------------------------------------------------- Hybrid Taus
// HybridTausRng / HybridTausRngInt / TausStep / LCGStep
// (kernels/mtwister.cu:118-153): the collision RNG of
// ComputeBirdCollisionsNoMTDev (kernels/dsmc_kernel.cu:912-1000).
struct TausState {
  uint32_t x = 0, y = 0, z = 0, w = 0;
  static uint32_t step(uint32_t& v, int s1, int s2, int s3, unsigned long long mask) {
    uint32_t b = ((v << s1) ^ v) >> s2;
    v = (uint32_t)((((unsigned long long)v & mask) << s3) ^ b);
    return v;
  }
  uint32_t next() {
    uint32_t r = step(x, 13, 19, 12, 4294967294ULL) ^ step(y, 2, 25, 4, 4294967288ULL) ^
                 step(z, 3, 11, 17, 4294967280ULL);
    w = (uint32_t)((unsigned long long)w * 1664525ULL + 1013904223ULL);
    return r ^ w;
  }
  float uniform() { return (float)(2.3283064365387e-10 * (double)next()); }
};

// ------------------------------------------------------------------ MT
struct MtBlock {  // genrand_mt-style generator (own implementation)
  uint32_t n = 19, m = 9, a = 0, lower = 1, u = 12, s = 7, b = 0, t = 15, c = 0, l = 18;
  std::vector<uint32_t> st;
  std::vector<uint32_t> prev;  // the block before the last regen (for to_window)
  uint32_t i = 0;
  void seed(uint32_t sd) {
    st.assign(n, 0u);
    prev.clear();
    st[0] = sd;
    for (uint32_t k = 1; k < n; ++k) st[k] = 1812433253u * (st[k - 1] ^ (st[k - 1] >> 30)) + k;
    i = n;
  }
  uint32_t tw(uint32_t x0, uint32_t x1) const {
    uint32_t y = (x0 & ~lower) | (x1 & lower);
    return (y >> 1) ^ ((y & 1u) ? a : 0u);
  }
  void regen() {
    prev = st;
    for (uint32_t k = 0; k < n; ++k) {
      uint32_t nxt = st[(k + 1) % n];
      st[k] = st[(k + m) % n] ^ tw(st[k], nxt);
    }
    i = 0;
  }
  uint32_t next() {
    if (i >= n) regen();
    uint32_t y = st[i++];
    y ^= y >> u;
    y ^= (y << s) & b;
    y ^= (y << t) & c;
    y ^= y >> l;
    return y;
  }
  float uniform() { return ((float)next() + 1.0f) / 4294967296.0f; }
  // window of the last n words, oldest first (the engine's state layout)
  // Sliding window x_d .. x_{d+18} after d draws: with i words of the
  // current block consumed, the tail of the previous block then the head of
  // the current one.
  void to_window(uint32_t* w) const {
    for (uint32_t k = 0; k < n; ++k) w[k] = (i + k < n) ? prev[i + k] : st[i + k - n];
  }
  void from_window(const uint32_t* w) {
    st.assign(w, w + n);
    prev.clear();
    i = n;
  }
};

// In regen() the update of st[k] for k >= n - m reads st[k + m - n], which
// was already replaced in this pass, and st[n-1] reads the new st[0]: the
// modular loop above does exactly that because it runs in increasing k.

// --------------------------------------------------------------- geometry
struct V3 {
  float x, y, z;
};

bool mt_intersect(float ox, float oy, float oz, float dx, float dy, float dz, const V3& a, const V3& b, const V3& c,
                  float& t, float& u, float& v, bool rel = false) {
  // intersect_triangle (tri_intersect.cu:14-56); rel: PHYSICS mode's
  // parallel threshold 1e-9 (|e1|^2 + |e2|^2) instead of the absolute 1e-6
  float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
  float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
  float px = dy * e2z - dz * e2y, py = dz * e2x - dx * e2z, pz = dx * e2y - dy * e2x;
  float det = e1x * px + e1y * py + e1z * pz;
  const double EPSILON = rel ? 1.0e-9 * ((double)(e1x * e1x + e1y * e1y + e1z * e1z) +
                                         (double)(e2x * e2x + e2y * e2y + e2z * e2z))
                             : 0.000001;
  if (det > -EPSILON && det < EPSILON) return false;
  float inv_det = (float)(1.0 / det);
  float tx = ox - a.x, ty = oy - a.y, tz = oz - a.z;
  u = (tx * px + ty * py + tz * pz) * inv_det;
  if (u < 0.0 || u > 1.0) return false;
  float qx = ty * e1z - tz * e1y, qy = tz * e1x - tx * e1z, qz = tx * e1y - ty * e1x;
  v = (dx * qx + dy * qy + dz * qz) * inv_det;
  if (v < 0.0 || u + v > 1.0) return false;
  t = (e2x * qx + e2y * qy + e2z * qz) * inv_det;
  return true;
}

void normalize(float& x, float& y, float& z) {  // vectors.cu:24-31
  float l = std::sqrt(x * x + y * y + z * z);
  if (l > 0) {
    x = x / l;
    y = y / l;
    z = z / l;
  }
}

}  // namespace

struct oracle_sim {
  dsmc_settings s;
  bool phys;
  uint32_t X, Y, Z, C, TC, TT, S, G;
  float gmin[3], gmax[3], cell[3], cell_volume;
  float fnum, sigma_hs, vhs_k = 0, vhs_exp = 1, avg_conc, dt, mass, vm_wall, vm_inf, U[3], inj_lambda[6];
  uint32_t ppc = 0;
  // geometry
  bool has_geom = false;
  std::vector<V3> tv0, tv1, tv2, tn;
  float cgmin[3], cgmax[3], cgcell[3];
  uint32_t cgdim[3];
  std::vector<uint32_t> cgflag, cgstart, cgend, cgtri;
  std::vector<float> frac;
  std::vector<uint8_t> nearc;  // PHYSICS: near-body code per cell (the engine's, see near_body)
  // RNG streams
  std::vector<MtBlock> rng;
  std::vector<TausState> taus;  // per cell, REFERENCE mode with mt = 0 (empty otherwise)
  // state in storage order
  std::vector<float> x, y, z, vx, vy, vz;
  std::vector<uint32_t> key;
  std::vector<float> ntc;  // [C][2]
  std::vector<float> cstat;  // [C][4]
  std::vector<double> samp;  // [5][C]
  std::vector<double> acc;   // [G][7]
  std::vector<uint32_t> perm, cell_start, hits;
  uint32_t n_store = 0;
  uint32_t stepno = 0;

  // ---------------------------------------------------------- keys
  uint32_t cell_key(float px, float py, float pz) const {
    uint32_t ix = dm_f2u_sat((px - gmin[0]) / cell[0]);
    uint32_t iy = dm_f2u_sat((py - gmin[1]) / cell[1]);
    uint32_t iz = dm_f2u_sat((pz - gmin[2]) / cell[2]);
    return std::min(iz, Z - 1) * X * Y + std::min(iy, Y - 1) * X + std::min(ix, X - 1);
  }
  // PHYSICS: offset times the float reciprocal of the cell size
  uint32_t cell_key_phys(float px, float py, float pz) const {
    const float r0 = 1.0f / cell[0], r1 = 1.0f / cell[1], r2 = 1.0f / cell[2];
    uint32_t ix = dm_f2u_sat((px - gmin[0]) * r0);
    uint32_t iy = dm_f2u_sat((py - gmin[1]) * r1);
    uint32_t iz = dm_f2u_sat((pz - gmin[2]) * r2);
    return std::min(iz, Z - 1) * X * Y + std::min(iy, Y - 1) * X + std::min(ix, X - 1);
  }
  uint32_t key_of(float px, float py, float pz) const {
    return phys ? cell_key_phys(px, py, pz) : cell_key(px, py, pz);
  }
  uint32_t cg_cell(float px, float py, float pz) const {  // GetCollisionCellID
    if (px < cgmin[0] || py < cgmin[1] || pz < cgmin[2] || px > cgmax[0] || py > cgmax[1] || pz > cgmax[2])
      return INV;
    uint32_t ix = std::min(dm_f2u_sat((px - cgmin[0]) / cgcell[0]), cgdim[0] - 1);
    uint32_t iy = std::min(dm_f2u_sat((py - cgmin[1]) / cgcell[1]), cgdim[1] - 1);
    uint32_t iz = std::min(dm_f2u_sat((pz - cgmin[2]) / cgcell[2]), cgdim[2] - 1);
    return iz * cgdim[0] * cgdim[1] + iy * cgdim[0] + ix;
  }

  // ------------------------------------------------- REFERENCE move
  bool collide_geometry(float& px, float& py, float& pz, float& ux, float& uy, float& uz, float& ox, float& oy,
                        float& oz, uint32_t cid) const {
    float dx = ux, dy = uy, dz = uz;
    normalize(dx, dy, dz);
    float ex = px - ox, ey = py - oy, ez = pz - oz;
    float dist = std::sqrt(ex * ex + ey * ey + ez * ez);
    if (dist < 0.0001f) return false;
    for (uint32_t i = cgstart[cid]; i < cgend[cid]; ++i) {
      uint32_t tri = cgtri[i];
      const V3 &a = tv0[tri], &b = tv1[tri], &c = tv2[tri];
      float t, u, v;
      if (mt_intersect(ox, oy, oz, dx, dy, dz, a, b, c, t, u, v) && t >= 0 && t <= dist) {
        float uv = 1 - u - v;
        const V3& n = tn[tri];
        float wx = n.x * 0.0001f, wy = n.y * 0.0001f, wz = n.z * 0.0001f;
        float hx = a.x * uv + b.x * u + c.x * v;
        float hy = a.y * uv + b.y * u + c.y * v;
        float hz = a.z * uv + b.z * u + c.z * v;
        float d = ux * n.x + uy * n.y + uz * n.z;  // reflect4
        ux = ux - 2 * d * n.x;
        uy = uy - 2 * d * n.y;
        uz = uz - 2 * d * n.z;
        float rx = ux, ry = uy, rz = uz;
        normalize(rx, ry, rz);
        float od = dist - t;
        px = hx + rx * od;
        py = hy + ry * od;
        pz = hz + rz * od;
        ox = hx + wx;
        oy = hy + wy;
        oz = hz + wz;
        return true;
      }
    }
    return false;
  }

  void ref_box(float& px, float& py, float& pz, float& ux, float& uy, float& uz) const {
    const double D = 0.001;  // BOUNCE_DELTA
    const float *mn = gmin, *mx = gmax;
    switch (s.periodic_condition) {
      case 'x':
        if (px > mx[0]) px = (float)((double)mn[0] + D + px - mx[0]);
        if (py > mx[1]) { py = (float)(mx[1] - D); uy *= -1; }
        if (px < mn[0]) px = (float)((double)mx[0] - D + px - mn[0]);
        if (py < mn[1]) { py = (float)(mn[1] + D); uy *= -1; }
        if (pz < mn[2]) { pz = (float)(mn[2] + D); uz *= -1; }
        if (pz > mx[2]) { pz = (float)(mx[2] - D); uz *= -1; }
        break;
      case 'y':
        if (py > mx[1]) py = (float)((double)mn[1] + D + py - mx[1]);
        if (px > mx[0]) { px = (float)(mx[0] - D); ux *= -1; }
        if (py < mn[1]) py = (float)((double)mx[1] - D + py - mn[1]);
        if (px < mn[0]) { px = (float)(mn[0] + D); ux *= -1; }
        if (pz < mn[2]) { pz = (float)(mn[2] + D); uz *= -1; }
        if (pz > mx[2]) { pz = (float)(mx[2] - D); uz *= -1; }
        break;
      case 'z':
        if (pz > mx[2]) pz = (float)((double)mn[2] + D + pz - mx[2]);
        if (py > mx[1]) { py = (float)(mx[1] - D); uy *= -1; }
        if (pz < mn[2]) pz = (float)((double)mx[2] - D + pz - mn[2]);
        if (py < mn[1]) { py = (float)(mn[1] + D); uy *= -1; }
        if (px < mn[0]) { px = (float)(mn[0] + D); ux *= -1; }
        if (px > mx[0]) { px = (float)(mx[0] - D); ux *= -1; }
        break;
      default:
        if (py > mx[1]) { py = (float)(mx[1] - D); uy *= -1; }
        if (px > mx[0]) { px = (float)(mx[0] - D); ux *= -1; }
        if (py < mn[1]) { py = (float)(mn[1] + D); uy *= -1; }
        if (px < mn[0]) { px = (float)(mn[0] + D); ux *= -1; }
        if (pz < mn[2]) { pz = (float)(mn[2] + D); uz *= -1; }
        if (pz > mx[2]) { pz = (float)(mx[2] - D); uz *= -1; }
        break;
    }
  }

  uint32_t move_reference(float& px, float& py, float& pz, float& ux, float& uy, float& uz, float step) const {
    float ox = px, oy = py, oz = pz;
    float nx = ox + ux * step, ny = oy + uy * step, nz = oz + uz * step;
    uint32_t nh = 0;
    if (has_geom) {
      bool coll = false;
      int i = 0;
      do {
        uint32_t st = cg_cell(ox, oy, oz), en = cg_cell(nx, ny, nz);
        if (st != INV && cgflag[st] != 0) coll = collide_geometry(nx, ny, nz, ux, uy, uz, ox, oy, oz, st);
        if (!coll && st != en && en != INV && cgflag[en] != 0)
          coll = collide_geometry(nx, ny, nz, ux, uy, uz, ox, oy, oz, en);
        if (coll) ++nh;
      } while (coll && i++ < 10);
    }
    ref_box(nx, ny, nz, ux, uy, uz);
    px = nx;
    py = ny;
    pz = nz;
    return nh;
  }

  // --------------------------------------------------- PHYSICS move
  // Nearest front-facing hit with t in [0, dist]; ties on t go to the lower
  // triangle index. Small meshes: every triangle. Large meshes: the triangles
  // listed in the collision cells overlapping the segment's bounding box
  // (any triangle the segment hits touches the cell holding the hit point,
  // which lies in that box).
  bool test_tri(uint32_t tri, float ox, float oy, float oz, float dx, float dy, float dz, float dist, uint32_t skip,
                bool& found, float& best, uint32_t& bt) const {
    if (tri == skip) return false;
    const V3& n = tn[tri];
    if (!(dx * n.x + dy * n.y + dz * n.z < 0.0f)) return false;
    float t, u, v;
    if (mt_intersect(ox, oy, oz, dx, dy, dz, tv0[tri], tv1[tri], tv2[tri], t, u, v, true) && t >= 0.0f &&
        t <= dist) {
      if (!found || t < best || (t == best && tri < bt)) {
        found = true;
        best = t;
        bt = tri;
      }
    }
    return true;
  }

  bool nearest_hit(float ox, float oy, float oz, float dx, float dy, float dz, float dist, uint32_t skip, float& best,
                   uint32_t& bt) const {
    bool found = false;
    if (tn.size() <= 4096) {
      for (uint32_t tri = 0; tri < (uint32_t)tn.size(); ++tri)
        test_tri(tri, ox, oy, oz, dx, dy, dz, dist, skip, found, best, bt);
      return found;
    }
    float ex = ox + dx * dist, ey = oy + dy * dist, ez = oz + dz * dist;
    float lo[3] = {std::min(ox, ex), std::min(oy, ey), std::min(oz, ez)};
    float hi[3] = {std::max(ox, ex), std::max(oy, ey), std::max(oz, ez)};
    int a[3], b[3];
    for (int k = 0; k < 3; ++k) {
      if (hi[k] < cgmin[k] || lo[k] > cgmax[k]) return false;
      a[k] = std::max(0, (int)std::floor((lo[k] - cgmin[k]) / cgcell[k]) - 1);
      b[k] = std::min((int)cgdim[k] - 1, (int)std::floor((hi[k] - cgmin[k]) / cgcell[k]) + 1);
    }
    for (int z = a[2]; z <= b[2]; ++z)
      for (int y = a[1]; y <= b[1]; ++y)
        for (int x = a[0]; x <= b[0]; ++x) {
          uint32_t c = (uint32_t)z * cgdim[0] * cgdim[1] + (uint32_t)y * cgdim[0] + (uint32_t)x;
          if (cgflag[c] != 1) continue;
          for (uint32_t i = cgstart[c]; i < cgend[c]; ++i)
            test_tri(cgtri[i], ox, oy, oz, dx, dy, dz, dist, skip, found, best, bt);
        }
    return found;
  }

  static void tangents(float nx, float ny, float nz, float* t1, float* t2) {
    float ax = 0, ay = 0, az = 0;
    float fx = std::fabs(nx), fy = std::fabs(ny), fz = std::fabs(nz);
    if (fx <= fy && fx <= fz) ax = 1;
    else if (fy <= fz) ay = 1;
    else az = 1;
    float cx = ny * az - nz * ay, cy = nz * ax - nx * az, cz = nx * ay - ny * ax;
    normalize(cx, cy, cz);
    t1[0] = cx; t1[1] = cy; t1[2] = cz;
    t2[0] = ny * cz - nz * cy;
    t2[1] = nz * cx - nx * cz;
    t2[2] = nx * cy - ny * cx;
  }

  static bool face(float& q, float& v, float lo, float hi, int flo, int fhi) {
    if (q < lo) {
      if (flo == DSMC_FACE_FREESTREAM) return false;
      if (flo == DSMC_FACE_PERIODIC) q = q + (hi - lo);
      else { q = lo + (lo - q); v = -v; }
    } else if (q > hi) {
      if (fhi == DSMC_FACE_FREESTREAM) return false;
      if (fhi == DSMC_FACE_PERIODIC) q = q - (hi - lo);
      else { q = hi - (q - hi); v = -v; }
    }
    q = q < lo ? lo : (q > hi ? hi : q);
    return true;
  }

  // PHYSICS move, in two parts as the engine runs it: a particle that moves
  // a cell or more along some axis, or whose flight |u| dt reaches its
  // cell's near-body distance (code * smallest cell / 64; 255 = far), is
  // deferred to the wall pass (move_walls, after all tiles); all others
  // just fly.
  bool near_body(uint32_t c, float ux, float uy, float uz) const {
    if (!has_geom) return false;
    if (std::fabs(ux) * dt >= cell[0] || std::fabs(uy) * dt >= cell[1] || std::fabs(uz) * dt >= cell[2]) return true;
    const uint32_t code = nearc.empty() ? 0u : nearc[c];  // without the engine's codes: every cell at distance 0
    if (code == 255u) return false;
    const float unit = std::min(cell[0], std::min(cell[1], cell[2])) / 64.0f;
    const float s2 = (ux * ux + uy * uy + uz * uz) * (dt * dt);
    const float d = (float)code * unit;
    return s2 >= d * d;
  }

  uint32_t faces_key(float& px, float& py, float& pz, float& ux, float& uy, float& uz) const {
    if (!face(px, ux, gmin[0], gmax[0], s.face_bc[0], s.face_bc[1])) return INV;
    if (!face(py, uy, gmin[1], gmax[1], s.face_bc[2], s.face_bc[3])) return INV;
    if (!face(pz, uz, gmin[2], gmax[2], s.face_bc[4], s.face_bc[5])) return INV;
    return cell_key_phys(px, py, pz);
  }

  uint32_t move_free(float& px, float& py, float& pz, float& ux, float& uy, float& uz) const {
    px = px + ux * dt;
    py = py + uy * dt;
    pz = pz + uz * dt;
    return faces_key(px, py, pz, ux, uy, uz);
  }

  // Walls, then faces. Diffuse re-emission of bounce b draws Philox4x32-10 at
  // counter (slot, step, 0, b) under key (seed, 0x6A09E667).
  uint32_t move_walls(float& px, float& py, float& pz, float& ux, float& uy, float& uz, uint32_t slot,
                      uint32_t& nh, double* a) const {
    nh = 0;
    float tl = dt;
    uint32_t skip = INV;
    for (int bnc = 0; bnc < 8; ++bnc) {
      float sp = std::sqrt(ux * ux + uy * uy + uz * uz);
      float dist = sp * tl;
      if (!(dist > 0.0f)) break;
      float dx = ux / sp, dy = uy / sp, dz = uz / sp;
      float sh;
      uint32_t tri;
      if (!nearest_hit(px, py, pz, dx, dy, dz, dist, skip, sh, tri)) break;
      const V3& n = tn[tri];
      float hx = px + dx * sh, hy = py + dy * sh, hz = pz + dz * sh;
      float wx, wy, wz;
      if (vm_wall > 0.0f) {
        dsmc_rng::U32x4 c = {{slot, stepno, 0u, (uint32_t)bnc}};
        dsmc_rng::U32x4 rr = dsmc_rng::philox4x32_10(c, (uint32_t)s.seed, 0x6A09E667u);
        float r1 = dsmc_rng::px_unit(rr.v[0]), r2 = dsmc_rng::px_unit(rr.v[1]), r3 = dsmc_rng::px_unit(rr.v[2]);
        float vn = vm_wall * std::sqrt(-dm_logf(r1));
        float rt = vm_wall * std::sqrt(-dm_logf(r2));
        float sn, cs;
        dm_sincosf(6.28318530717958647692f * r3, &sn, &cs);
        float a1 = rt * cs, a2 = rt * sn;
        float t1[3], t2[3];
        tangents(n.x, n.y, n.z, t1, t2);
        wx = n.x * vn + t1[0] * a1 + t2[0] * a2;
        wy = n.y * vn + t1[1] * a1 + t2[1] * a2;
        wz = n.z * vn + t1[2] * a1 + t2[2] * a2;
      } else {
        float d = ux * n.x + uy * n.y + uz * n.z;
        wx = ux - 2.0f * d * n.x;
        wy = uy - 2.0f * d * n.y;
        wz = uz - 2.0f * d * n.z;
      }
      double m = mass;
      a[3] += m * ((double)ux - (double)wx);
      a[4] += m * ((double)uy - (double)wy);
      a[5] += m * ((double)uz - (double)wz);
      a[6] += 0.5 * m * (((double)ux * ux + (double)uy * uy + (double)uz * uz) - ((double)wx * wx + (double)wy * wy + (double)wz * wz));
      ++nh;
      tl = tl - sh / sp;
      px = hx + n.x * 1.0e-6f;
      py = hy + n.y * 1.0e-6f;
      pz = hz + n.z * 1.0e-6f;
      ux = wx;
      uy = wy;
      uz = wz;
      skip = tri;
    }
    if (tl > 0.0f) {
      px = px + ux * tl;
      py = py + uy * tl;
      pz = pz + uz * tl;
    }
    return faces_key(px, py, pz, ux, uy, uz);
  }

  // ------------------------------------------------------------- NTC
  template <class R>
  static void scatter(float* ux, float* uy, float* uz, uint32_t i, uint32_t j, float mag, R& r) {
    const float u1 = r.uniform();
    const float u2 = r.uniform();
    scatter_uv(ux, uy, uz, i, j, mag, u1, u2);
  }
  // isotropic post-collision velocities from two uniforms in (0, 1]:
  // cos(theta) = 2 u1 - 1, azimuth 2 pi u2 (dsmc_kernel.cu:1057-1078)
  static void scatter_uv(float* ux, float* uy, float* uz, uint32_t i, uint32_t j, float mag, float u1, float u2) {
    float v1x = ux[i], v1y = uy[i], v1z = uz[i], v2x = ux[j], v2y = uy[j], v2z = uz[j];
    float b = 2.0f * u1 - 1.0f;
    float a = std::sqrt(1 - b * b);
    float c = 2 * PI_REF * u2;
    float sn, cs;
    dm_sincosf(c, &sn, &cs);
    float ax = b * mag / 2.0f, ay = a * mag * cs / 2.0f, az = a * mag * sn / 2.0f;
    ux[i] = (v1x + v2x) / 2.0f + ax;
    uy[i] = (v1y + v2y) / 2.0f + ay;
    uz[i] = (v1z + v2z) / 2.0f + az;
    ux[j] = (v1x + v2x) / 2.0f - ax;
    uy[j] = (v1y + v2y) / 2.0f - ay;
    uz[j] = (v1z + v2z) / 2.0f - az;
  }

  template <class R>
  void ntc_reference(float* ux, float* uy, float* uz, int part, float* nt, float* st, R& r) const {
    if (part < 2) return;
    float fpc = (0.5 * part * avg_conc * fnum * nt[0] * dt) / cell_volume + nt[1];
    uint32_t pairs = dm_f2u_sat(fpc);
    st[0] += pairs;
    nt[1] = fpc - pairs;
    pairs = std::min(pairs, kMaxPairs);
    for (uint32_t k = 0; k < pairs; ++k) {
      uint32_t i = r.next() % (uint32_t)part;
      uint32_t j = r.next() % (uint32_t)part;
      float dx = ux[i] - ux[j], dy = uy[i] - uy[j], dz = uz[i] - uz[j];
      float mag = std::sqrt(dx * dx + dy * dy + dz * dz);
      float v = mag * sigma_hs;
      if (v / nt[0] > r.uniform()) {
        st[1]++;
        st[2] += mag;
        scatter(ux, uy, uz, i, j, mag, r);
      }
      if (v > nt[0]) nt[0] = v;
    }
  }

  // PHYSICS NTC: M = floor(n (n - 1) / 2 F (sigma c_r)max dt / V + rem)
  // candidates in rounds of disjoint pairs (pairing.h). Every pair of a round
  // is tested against the round's starting (sigma c_r)max, then the maximum
  // takes the largest sigma c_r of the round; VHS cross-section
  // sigma c_r = vhs_k c_r^(2 - 2 omega). Draws: Philox4x32-10 under
  // (seed, kNtcKey1) at (cell, step, round, item).
  void ntc_physics(float* ux, float* uy, float* uz, uint32_t n, float vol, float* nt, uint32_t cell,
                   double* a) const {
    using namespace dsmc_pairing;
    if (n < 2) return;
    float fpc = (float)(0.5 * (double)n * (double)(n - 1) * (double)fnum * (double)nt[0] * (double)dt / (double)vol +
                        (double)nt[1]);
    uint32_t pairs = dm_f2u_sat(fpc);
    nt[1] = fpc - (float)pairs;
    pairs = std::min(pairs, kMaxPairs);
    a[0] += pairs;
    const uint32_t rounds = rounds_of(n, pairs);
    for (uint32_t r = 0; r < rounds; ++r) {
      const uint32_t m = round_pairs(n, pairs, r);
      const dsmc_rng::U32x4 rk = dsmc_rng::philox4x32_10({{cell, stepno, r, kRoundItem}}, (uint32_t)s.seed, kNtcKey1);
      const float max0 = nt[0];
      float rmax = max0;
      for (uint32_t k = 0; k < m; ++k) {
        uint32_t i, j;
        pair_members(n, m, k, rk, i, j);
        const dsmc_rng::U32x4 d = dsmc_rng::philox4x32_10({{cell, stepno, r, k}}, (uint32_t)s.seed, kNtcKey1);
        float dx = ux[i] - ux[j], dy = uy[i] - uy[j], dz = uz[i] - uz[j];
        float mag = std::sqrt(dx * dx + dy * dy + dz * dz);
        float sc = vhs_exp == 1.0f ? vhs_k * mag : (mag > 0.0f ? vhs_k * dm_powf(mag, vhs_exp) : 0.0f);
        if (sc > rmax) rmax = sc;
        if (sc > dsmc_rng::px_unit(d.v[0]) * max0) {
          a[1] += 1.0;
          a[2] += (double)std::llrint((double)mag * 65536.0);  // c_r in units of 2^-16 m/s, summed exactly
          scatter_uv(ux, uy, uz, i, j, mag, dsmc_rng::px_unit(d.v[1]), dsmc_rng::px_unit(d.v[2]));
        }
      }
      nt[0] = rmax;
    }
  }

  bool inside_body(float px, float py, float pz) const {
    if (!has_geom) return false;
    uint32_t c = cg_cell(px, py, pz);
    if (c == INV) return false;
    if (cgflag[c] == 2) return true;
    if (cgflag[c] != 1) return false;
    float best = 3.0e38f, sd = 1.0f;
    for (uint32_t i = cgstart[c]; i < cgend[c]; ++i) {
      uint32_t t = cgtri[i];
      const V3 &a = tv0[t], &b = tv1[t], &cc = tv2[t], &n = tn[t];
      float cx = (a.x + b.x + cc.x) / 3.0f, cy = (a.y + b.y + cc.y) / 3.0f, cz = (a.z + b.z + cc.z) / 3.0f;
      float rx = px - cx, ry = py - cy, rz = pz - cz;
      float d2 = rx * rx + ry * ry + rz * rz;
      if (d2 < best) {
        best = d2;
        sd = rx * n.x + ry * n.y + rz * n.z;
      }
    }
    return sd < 0.0f;
  }

  bool geometry_in_cell(float cx, float cy, float cz) const {  // IsGeometryInCell (:479-546)
    const float sx = cell[0], sy = cell[1], sz = cell[2];
    float pts[8][3];
    float px = cx, py = cy, pz = cz;
    int k = 0;
    auto put = [&]() { pts[k][0] = px; pts[k][1] = py; pts[k][2] = pz; ++k; };
    put();
    px = px + sx; put();
    py = py + sy; put();
    pz = pz + sz; put();
    px = cx; py = cy; pz = cz;
    py = py + sy; put();
    pz = pz + sz; put();
    py = cy; put();
    px = px + sx; put();
    for (int q = 0; q < 8; ++q) {
      uint32_t c = cg_cell(pts[q][0], pts[q][1], pts[q][2]);
      if (c != INV && cgflag[c] != 0) return true;
    }
    return false;
  }

  // ------------------------------------------------------ tile schedule
  // Locality order of the tiles (see the header).
  std::vector<uint32_t> tile_order(uint32_t ntiles) const {
    uint32_t B = 1;
    while (uint64_t(4) * B * B * X <= uint64_t(128) * TC && 2 * B <= std::max(Y, Z)) B *= 2;
    const uint64_t yp = (Y + B - 1) / B;
    std::vector<std::pair<uint64_t, uint32_t>> keyed;
    keyed.reserve(ntiles);
    for (uint32_t t = 0; t < ntiles; ++t) {
      const uint64_t c = uint64_t(t) * TC;
      const uint64_t cx = c % X, cy = (c / X) % Y, cz = c / (uint64_t(X) * Y);
      const uint64_t patch = (cz / B) * yp + cy / B;
      keyed.push_back({((patch * B + cz % B) * B + cy % B) * X + cx, t});
    }
    std::sort(keyed.begin(), keyed.end());
    std::vector<uint32_t> out;
    out.reserve(ntiles);
    for (const auto& e : keyed) out.push_back(e.second);
    return out;
  }

  // Positions of the order that workgroup vw visits, in visiting order.
  std::vector<uint32_t> tile_positions(uint32_t vw, uint32_t ntiles) const {
    std::vector<uint32_t> js;
    if (G % 8 == 0) {
      const uint64_t P = 128, x = vw % 8;
      for (uint64_t e = vw / 8;; e += G / 8) {
        const uint64_t j = ((e / P) * 8 + x) * P + e % P;
        if (j >= ntiles) break;
        js.push_back((uint32_t)j);
      }
    } else {
      for (uint64_t j = vw; j < ntiles; j += G) js.push_back((uint32_t)j);
    }
    return js;
  }

  // ------------------------------------------------------------- step
  void step(bool sample) {
    ++stepno;  // the engine's step counter (counter of the wall draws)
    // stable counting sort of the storage by key (the sort_by_key + bounds)
    std::vector<uint32_t> cnt(C, 0);
    for (uint32_t i = 0; i < n_store; ++i)
      if (key[i] != INV) cnt[key[i]]++;
    cell_start.assign(C + 1, 0);
    for (uint32_t c = 0; c < C; ++c) cell_start[c + 1] = cell_start[c] + cnt[c];
    const uint32_t N = cell_start[C];
    perm.assign(N, 0);
    {
      std::vector<uint32_t> cur(cell_start.begin(), cell_start.end() - 1);
      for (uint32_t i = 0; i < n_store; ++i)
        if (key[i] != INV) perm[cur[key[i]]++] = i;
    }
    // REORDER gather
    std::vector<float> nx(N), ny(N), nz(N), nvx(N), nvy(N), nvz(N);
    for (uint32_t q = 0; q < N; ++q) {
      uint32_t src = perm[q];
      nx[q] = x[src]; ny[q] = y[src]; nz[q] = z[src];
      nvx[q] = vx[src]; nvy[q] = vy[src]; nvz[q] = vz[src];
    }
    std::vector<uint32_t> nkey(N);
    hits.assign(N, 0);
    const uint32_t ntiles = (C + TC - 1) / TC;
    const std::vector<uint32_t> order = tile_order(ntiles);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t vw = 0; vw < (int64_t)G; ++vw) {
      double* a = &acc[(size_t)vw * 7];
      for (const uint32_t j : tile_positions((uint32_t)vw, ntiles)) {
        const uint32_t tile = order[j];
        uint32_t c0 = tile * TC, nc = std::min(TC, C - c0);
        uint32_t p0 = cell_start[c0], p1 = cell_start[c0 + nc];
        // NTC: PHYSICS one stream per cell lane, REFERENCE the cell's own stream
        for (uint32_t l = 0; l < nc; ++l) {
          uint32_t c = c0 + l;
          uint32_t a0 = cell_start[c], b0 = cell_start[c + 1];
          if (phys) {
            float vol = cell_volume * (frac.empty() ? 1.0f : std::max(frac[c], 1.0f / 64.0f));
            ntc_physics(&nvx[a0], &nvy[a0], &nvz[a0], b0 - a0, vol, &ntc[2 * (size_t)c], c, a);
          } else if (!taus.empty()) {
            ntc_reference(&nvx[a0], &nvy[a0], &nvz[a0], (int)(b0 - a0), &ntc[2 * (size_t)c], &cstat[4 * (size_t)c],
                          taus[c]);
          } else {
            ntc_reference(&nvx[a0], &nvy[a0], &nvz[a0], (int)(b0 - a0), &ntc[2 * (size_t)c], &cstat[4 * (size_t)c],
                          rng[c]);
          }
        }
        // move (no draws: wall re-emission is counter-based)
        uint32_t cq = c0;
        for (uint32_t q = p0; q < p1; ++q) {
          while (q >= cell_start[cq + 1]) ++cq;
          uint32_t nh = 0, k;
          if (phys) {
            k = near_body(cq, nvx[q], nvy[q], nvz[q]) ? DEFER
                                                      : move_free(nx[q], ny[q], nz[q], nvx[q], nvy[q], nvz[q]);
          } else {
            nh = move_reference(nx[q], ny[q], nz[q], nvx[q], nvy[q], nvz[q], dt);
            k = cell_key(nx[q], ny[q], nz[q]);
          }
          nkey[q] = k;
          hits[q] = nh;
        }
        if (sample) {
          for (uint32_t l = 0; l < nc; ++l) {
            uint32_t c = c0 + l;
            uint32_t a0 = cell_start[c], b0 = cell_start[c + 1];
            if (b0 <= a0) continue;
            if (phys) {
              double sn = samp[c], sx = samp[C + c], sy = samp[2 * C + c], sz = samp[3 * C + c], s2 = samp[4 * C + c];
              for (uint32_t q = a0; q < b0; ++q) {
                double ux = nvx[q], uy = nvy[q], uz = nvz[q];
                sn += 1.0; sx += ux; sy += uy; sz += uz;
                s2 += ux * ux + uy * uy + uz * uz;
              }
              samp[c] = sn; samp[C + c] = sx; samp[2 * C + c] = sy; samp[3 * C + c] = sz; samp[4 * C + c] = s2;
            } else {
              float sn = (float)samp[c], sx = (float)samp[C + c], sy = (float)samp[2 * C + c],
                    sz = (float)samp[3 * C + c], s2 = (float)samp[4 * C + c];
              for (uint32_t q = a0; q < b0; ++q) {
                float ux = nvx[q], uy = nvy[q], uz = nvz[q];
                sn += 1.0f; sx += ux; sy += uy; sz += uz;
                s2 += ux * ux + uy * uy + uz * uz;
              }
              samp[c] = sn; samp[C + c] = sx; samp[2 * C + c] = sy; samp[3 * C + c] = sz; samp[4 * C + c] = s2;
            }
          }
        }
      }
    }
    // wall pass over the deferred particles, in slot order
    if (phys && has_geom) {
      double wf[4] = {0.0, 0.0, 0.0, 0.0};
#pragma omp parallel
      {
        double a7[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 4096)
        for (int64_t qq = 0; qq < (int64_t)N; ++qq) {
          const uint32_t q = (uint32_t)qq;
          if (nkey[q] != DEFER) continue;
          uint32_t nh = 0;
          nkey[q] = move_walls(nx[q], ny[q], nz[q], nvx[q], nvy[q], nvz[q], q, nh, a7);
          hits[q] = nh;
        }
#pragma omp critical
        for (int k = 0; k < 4; ++k) wf[k] += a7[3 + k];
      }
      for (int k = 0; k < 4; ++k) acc[3 + k] += wf[k];
    }
    x.swap(nx); y.swap(ny); z.swap(nz);
    vx.swap(nvx); vy.swap(nvy); vz.swap(nvz);
    key.swap(nkey);
    n_store = N;
    if (phys) inject();
  }

  // -------------------------------------------------------- injection
  void inject() {
    uint32_t off[7] = {0};
    for (int f = 0; f < 6; ++f) {
      uint32_t axis = f / 2, ta = axis == 0 ? 1 : 0, tb = axis == 2 ? 1 : 2;
      uint32_t d[3] = {X, Y, Z};
      off[f + 1] = off[f] + ((s.face_bc[f] == DSMC_FACE_FREESTREAM) ? d[ta] * d[tb] : 0u);
    }
    const uint32_t F = off[6];
    if (!F) return;
    std::vector<uint32_t> cnt(F);
    auto face_of = [&](uint32_t g) { uint32_t f = 0; while (f < 5 && g >= off[f + 1]) ++f; return f; };
    // Philox4x32-10 under key (seed, kInjectKey1): the count of face cell g
    // at counter (g, step, 0, 0); particle j of g at (g, step, 1 + j, b) for
    // rejection block b (two tries each) and (g, step, 1 + j, 2^31 + k)
    const uint32_t k0 = (uint32_t)s.seed, k1 = 0xBB67AE85u;
    auto draw = [&](uint32_t c0, uint32_t c2, uint32_t c3) {
      dsmc_rng::U32x4 c = {{c0, stepno, c2, c3}};
      return dsmc_rng::philox4x32_10(c, k0, k1);
    };
    for (uint32_t g = 0; g < F; ++g)
      cnt[g] = dm_f2u_sat(inj_lambda[face_of(g)] + dsmc_rng::px_unit(draw(g, 0u, 0u).v[0]));
    std::vector<uint32_t> start(F + 1, 0);
    for (uint32_t g = 0; g < F; ++g) start[g + 1] = start[g] + cnt[g];
    const uint32_t base = n_store, tot = start[F];
    x.resize(base + tot); y.resize(base + tot); z.resize(base + tot);
    vx.resize(base + tot); vy.resize(base + tot); vz.resize(base + tot);
    key.resize(base + tot);
    const float twopi = 6.28318530717958647692f;
    const float sd = vm_inf * 0.70710678118654752f;
    const float dims_cell[3] = {cell[0], cell[1], cell[2]};
    const uint32_t dims[3] = {X, Y, Z};
    for (uint32_t g = 0; g < F; ++g) {
      uint32_t k = cnt[g];
      if (!k) continue;
      uint32_t f = face_of(g), local = g - off[f];
      uint32_t axis = f >> 1, ta = axis == 0 ? 1u : 0u, tb = axis == 2 ? 1u : 2u;
      uint32_t ia = local % dims[ta], ib = local / dims[ta];
      float sign = (f & 1u) ? -1.0f : 1.0f;
      float un = sign * U[axis];
      float sr = un / vm_inf;
      float zlo = sr - 5.0f;
      zlo = zlo > 0.0f ? zlo : 0.0f;
      float zhi = sr + 5.0f;
      zhi = zhi > 1.0e-3f ? zhi : 1.0e-3f;
      float zs = 0.5f * (sr + std::sqrt(sr * sr + 2.0f));
      zs = zs < zlo ? zlo : (zs > zhi ? zhi : zs);
      float gmx = zs * dm_expf(-(zs - sr) * (zs - sr));
      for (uint32_t j = 0; j < k; ++j) {
        float zz = zlo;
        bool ok = false;
        for (uint32_t b = 0; b < 500u && !ok; ++b) {
          const dsmc_rng::U32x4 rr = draw(g, 1u + j, b);
          for (int h = 0; h < 2 && !ok; ++h) {
            zz = zlo + (zhi - zlo) * dsmc_rng::px_unit(rr.v[2 * h]);
            float gz = zz * dm_expf(-(zz - sr) * (zz - sr));
            ok = gz > gmx * dsmc_rng::px_unit(rr.v[2 * h + 1]);
          }
        }
        const dsmc_rng::U32x4 ra = draw(g, 1u + j, 0x80000000u), rb = draw(g, 1u + j, 0x80000001u);
        float vn = zz * vm_inf;
        float a1 = std::sqrt(-2.0f * dm_logf(dsmc_rng::px_unit(ra.v[0])));
        float s1, c1;
        dm_sincosf(twopi * dsmc_rng::px_unit(ra.v[1]), &s1, &c1);
        float p[3], v[3];
        v[axis] = sign * vn;
        v[ta] = U[ta] + sd * a1 * c1;
        v[tb] = U[tb] + sd * a1 * s1;
        p[axis] = (f & 1u) ? gmax[axis] : gmin[axis];
        p[ta] = gmin[ta] + ((float)ia + dsmc_rng::px_unit(ra.v[2])) * dims_cell[ta];
        p[tb] = gmin[tb] + ((float)ib + dsmc_rng::px_unit(ra.v[3])) * dims_cell[tb];
        float fr = dsmc_rng::px_unit(rb.v[0]) * dt;
        float px = p[0] + v[0] * fr, py = p[1] + v[1] * fr, pz = p[2] + v[2] * fr;
        px = px < gmin[0] ? gmin[0] : (px > gmax[0] ? gmax[0] : px);
        py = py < gmin[1] ? gmin[1] : (py > gmax[1] ? gmax[1] : py);
        pz = pz < gmin[2] ? gmin[2] : (pz > gmax[2] ? gmax[2] : pz);
        uint32_t idx = base + start[g] + j;
        x[idx] = px; y[idx] = py; z[idx] = pz;
        vx[idx] = v[0]; vy[idx] = v[1]; vz[idx] = v[2];
        key[idx] = cell_key_phys(px, py, pz);
      }
    }
    n_store = base + tot;
  }

  // ---------------------------------------------------------- loading
  void load(float ux, float uy, float uz) {
    if (!phys) {
      const uint32_t P = ppc, n = P * C;
      x.assign(n, 0); y.assign(n, 0); z.assign(n, 0);
      vx.assign(n, 0); vy.assign(n, 0); vz.assign(n, 0);
      const float twopi = 2.0f * PI_REF;
      for (uint32_t t = 0; t < X * Y * P; ++t) {
        uint32_t i = t / (X * Y), xy = t - i * X * Y, cy = xy / X, cx = xy - cy * X;
        MtBlock& r = rng[t];
        float bx = gmin[0] + cx * cell[0], by = gmin[1] + cy * cell[1];
        for (uint32_t cz = 0; cz < Z; ++cz) {
          uint32_t idx = (X * Y * cz + cy * X + cx) * P + i;
          float bz = gmin[2] + cz * cell[2];
          if (has_geom && geometry_in_cell(bx, by, bz)) continue;  // parked at 0
          float vv = std::sqrt(2 * BOLZ_REF * s.temp / mass * (-dm_logf(r.uniform())));
          vx[idx] = vv * dm_sinf(r.uniform() * twopi) + ux;
          vy[idx] = vv * dm_sinf(r.uniform() * twopi) + uy;
          vz[idx] = vv * dm_sinf(r.uniform() * twopi) + uz;
          x[idx] = bx + r.uniform() * cell[0];
          y[idx] = by + r.uniform() * cell[1];
          z[idx] = bz + r.uniform() * cell[2];
        }
      }
      for (uint32_t i = 0; i < n; ++i) move_reference(x[i], y[i], z[i], vx[i], vy[i], vz[i], 0.0f);
      n_store = n;
    } else {
      U[0] = ux; U[1] = uy; U[2] = uz;
      std::vector<uint32_t> cnt(C);
      for (uint32_t c = 0; c < C; ++c) {
        MtBlock& r = rng[c % S];
        float fr = frac.empty() ? 1.0f : frac[c];
        float lam = (float)((double)s.density * (double)cell_volume * (double)fr / (double)fnum);
        float rr = r.uniform();
        cnt[c] = fr > 0.0f ? dm_f2u_sat(lam + rr) : 0u;
        ntc[2 * (size_t)c + 1] = r.uniform();
      }
      std::vector<uint32_t> start(C + 1, 0);
      for (uint32_t c = 0; c < C; ++c) start[c + 1] = start[c] + cnt[c];
      const uint32_t n = start[C];
      x.assign(n, 0); y.assign(n, 0); z.assign(n, 0);
      vx.assign(n, 0); vy.assign(n, 0); vz.assign(n, 0);
      const float twopi = 6.28318530717958647692f;
      const float sdv = std::sqrt((float)(BOLZ * (double)s.temp / (double)mass));
      for (uint32_t c = 0; c < C; ++c) {
        uint32_t k = cnt[c];
        if (!k) continue;
        MtBlock& r = rng[c % S];
        uint32_t cz = c / (X * Y), rem = c - cz * X * Y, cy = rem / X, cx = rem - cy * X;
        float bx = gmin[0] + cx * cell[0], by = gmin[1] + cy * cell[1], bz = gmin[2] + cz * cell[2];
        bool cut = !frac.empty() && frac[c] < 1.0f;
        for (uint32_t j = 0; j < k; ++j) {
          float px, py, pz;
          int tries = 0;
          do {
            px = bx + r.uniform() * cell[0];
            py = by + r.uniform() * cell[1];
            pz = bz + r.uniform() * cell[2];
            ++tries;
          } while (cut && tries < 64 && inside_body(px, py, pz));
          float r1 = r.uniform(), r2 = r.uniform(), r3 = r.uniform(), r4 = r.uniform();
          float a1 = std::sqrt(-2.0f * dm_logf(r1)), a2 = std::sqrt(-2.0f * dm_logf(r3));
          float s1, c1, s2, c2;
          dm_sincosf(twopi * r2, &s1, &c1);
          dm_sincosf(twopi * r4, &s2, &c2);
          uint32_t idx = start[c] + j;
          x[idx] = px; y[idx] = py; z[idx] = pz;
          vx[idx] = ux + sdv * a1 * c1;
          vy[idx] = uy + sdv * a1 * s1;
          vz[idx] = uz + sdv * a2 * c2;
        }
      }
      n_store = n;
    }
    key.resize(n_store);
    for (uint32_t i = 0; i < n_store; ++i) key[i] = key_of(x[i], y[i], z[i]);
  }
};

extern "C" {

void oracle_mt_generic(uint32_t n, uint32_t m, uint32_t a, uint32_t lower_mask, uint32_t u, uint32_t s, uint32_t b,
                       uint32_t t, uint32_t c, uint32_t l, uint32_t seed, uint32_t n_out, uint32_t* out) {
  MtBlock g;
  g.n = n; g.m = m; g.a = a; g.lower = lower_mask; g.u = u; g.s = s; g.b = b; g.t = t; g.c = c; g.l = l;
  g.seed(seed);
  for (uint32_t k = 0; k < n_out; ++k) out[k] = g.next();
}

int oracle_tile_order(oracle_sim* o, uint32_t* out, uint32_t n) {
  if (!o || !out) return -1;
  const uint32_t ntiles = (o->C + o->TC - 1) / o->TC;
  if (n < ntiles) return -1;
  std::vector<uint32_t> ord = o->tile_order(ntiles);
  std::copy(ord.begin(), ord.end(), out);
  return (int)ntiles;
}

int oracle_tile_walk(oracle_sim* o, uint32_t vw, uint32_t* out, uint32_t n) {
  if (!o || !out || vw >= o->G) return -1;
  const uint32_t ntiles = (o->C + o->TC - 1) / o->TC;
  std::vector<uint32_t> js = o->tile_positions(vw, ntiles);
  if (js.size() > n) return -1;
  std::copy(js.begin(), js.end(), out);
  return (int)js.size();
}

void oracle_mt_streams(const uint32_t* a, const uint32_t* b, const uint32_t* c, const uint32_t* seeds,
                       uint32_t n_streams, uint32_t n_draws, uint32_t* out) {
  for (uint32_t sidx = 0; sidx < n_streams; ++sidx) {
    MtBlock g;
    g.a = a[sidx]; g.b = b[sidx]; g.c = c[sidx];
    g.seed(seeds[sidx]);
    for (uint32_t d = 0; d < n_draws; ++d) out[(size_t)d * n_streams + sidx] = g.next();
  }
}

void oracle_cell_ids(const dsmc_settings* s, const float* x, const float* y, const float* z, uint32_t n,
                     uint32_t* keys) {
  float mn[3], cl[3];
  uint32_t d[3];
  for (int k = 0; k < 3; ++k) {
    mn[k] = s->bmin[k];
    d[k] = (uint32_t)s->grid_dim[k];
    cl[k] = (s->bmax[k] - s->bmin[k]) / (float)s->grid_dim[k];
  }
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t ix = std::min(dm_f2u_sat((x[i] - mn[0]) / cl[0]), d[0] - 1);
    uint32_t iy = std::min(dm_f2u_sat((y[i] - mn[1]) / cl[1]), d[1] - 1);
    uint32_t iz = std::min(dm_f2u_sat((z[i] - mn[2]) / cl[2]), d[2] - 1);
    keys[i] = iz * d[0] * d[1] + iy * d[0] + ix;
  }
}

void oracle_stable_sort(const uint32_t* keys, uint32_t n, uint32_t n_cells, uint32_t* perm, uint32_t* cell_start) {
  std::vector<uint32_t> idx;
  idx.reserve(n);
  for (uint32_t i = 0; i < n; ++i)
    if (keys[i] < n_cells) idx.push_back(i);  // keys >= n_cells are dropped
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t p, uint32_t q) { return keys[p] < keys[q]; });
  std::copy(idx.begin(), idx.end(), perm);
  for (uint32_t c = 0; c <= n_cells; ++c) {
    auto it = std::lower_bound(idx.begin(), idx.end(), c, [&](uint32_t p, uint32_t v) { return keys[p] < v; });
    cell_start[c] = (uint32_t)(it - idx.begin());
  }
}

int oracle_intersect(const float o[3], const float d[3], const float v0[3], const float v1[3], const float v2[3],
                     float out_tuv[3]) {
  V3 a = {v0[0], v0[1], v0[2]}, b = {v1[0], v1[1], v1[2]}, c = {v2[0], v2[1], v2[2]};
  float t = 0, u = 0, v = 0;
  bool hit = mt_intersect(o[0], o[1], o[2], d[0], d[1], d[2], a, b, c, t, u, v);
  out_tuv[0] = t; out_tuv[1] = u; out_tuv[2] = v;
  return hit ? 1 : 0;
}

void oracle_detmath(int f, const float* x, const float* y, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    switch (f) {
      case 0: out[i] = dm_sinf(x[i]); break;
      case 1: out[i] = dm_cosf(x[i]); break;
      case 2: out[i] = dm_logf(x[i]); break;
      case 3: out[i] = dm_exp2f(x[i]); break;
      default: out[i] = dm_powf(x[i], y[i]); break;
    }
  }
}

oracle_sim* oracle_create(const dsmc_settings* s, uint32_t tile_cells, uint32_t tile_threads, uint32_t n_streams,
                          const uint32_t* a, const uint32_t* b, const uint32_t* c) {
  if (!s || !a || !b || !c || tile_cells == 0 || tile_threads == 0 || n_streams % tile_threads) return nullptr;
  auto* o = new oracle_sim();
  o->s = *s;
  o->phys = s->mode == DSMC_MODE_PHYSICS;
  o->X = (uint32_t)s->grid_dim[0];
  o->Y = (uint32_t)s->grid_dim[1];
  o->Z = (uint32_t)s->grid_dim[2];
  o->C = o->X * o->Y * o->Z;
  o->TC = tile_cells;
  o->TT = tile_threads;
  o->S = n_streams;
  o->G = n_streams / tile_threads;
  for (int k = 0; k < 3; ++k) {
    o->gmin[k] = s->bmin[k];
    o->gmax[k] = s->bmax[k];
    o->cell[k] = (s->bmax[k] - s->bmin[k]) / (float)s->grid_dim[k];
  }
  o->cell_volume = o->cell[0] * o->cell[1] * o->cell[2];
  o->fnum = s->fnum;
  o->sigma_hs = PI_REF * s->diameter * s->diameter;
  o->avg_conc = (float)s->part_per_cell;
  o->ppc = s->part_per_cell;
  o->dt = s->dt;
  o->mass = s->molmass;
  o->vm_wall = s->wall_temp > 0.0f ? (float)std::sqrt(2.0 * BOLZ * s->wall_temp / s->molmass) : 0.0f;
  o->vm_inf = (float)std::sqrt(2.0 * BOLZ * s->temp / s->molmass);
  for (int k = 0; k < 3; ++k) o->U[k] = s->stream_velocity[k];
  for (int f = 0; f < 6; ++f) o->inj_lambda[f] = 0.0f;
  o->rng.resize(n_streams);
  for (uint32_t i = 0; i < n_streams; ++i) {
    o->rng[i].a = a[i];
    o->rng[i].b = b[i];
    o->rng[i].c = c[i];
    o->rng[i].st.assign(19, 0u);
    o->rng[i].i = 19;
  }
  o->ntc.assign(2 * (size_t)o->C, 0.0f);
  o->cstat.assign(4 * (size_t)o->C, 0.0f);
  o->samp.assign(5 * (size_t)o->C, 0.0);
  o->acc.assign(7 * (size_t)o->G, 0.0);
  return o;
}

void oracle_destroy(oracle_sim* o) { delete o; }

int oracle_set_geometry(oracle_sim* o, const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                        const float* normals, const float cg_bounds[9], const uint32_t cg_dim[3],
                        const uint32_t* flags, const uint32_t* start, const uint32_t* end, const uint32_t* entries,
                        uint32_t n_entries, const float* fluid_frac) {
  if (!o) return -1;
  o->tv0.resize(nt); o->tv1.resize(nt); o->tv2.resize(nt); o->tn.resize(nt);
  for (uint32_t i = 0; i < nt; ++i) {
    const float* p = &verts[3 * tris[3 * i]];
    const float* q = &verts[3 * tris[3 * i + 1]];
    const float* r = &verts[3 * tris[3 * i + 2]];
    o->tv0[i] = {p[0], p[1], p[2]};
    o->tv1[i] = {q[0], q[1], q[2]};
    o->tv2[i] = {r[0], r[1], r[2]};
    o->tn[i] = {normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]};
  }
  (void)nv;
  uint32_t N = cg_dim[0] * cg_dim[1] * cg_dim[2];
  for (int k = 0; k < 3; ++k) {
    o->cgmin[k] = cg_bounds[k];
    o->cgmax[k] = cg_bounds[3 + k];
    o->cgcell[k] = cg_bounds[6 + k];
    o->cgdim[k] = cg_dim[k];
  }
  o->cgflag.assign(flags, flags + N);
  o->cgstart.assign(start, start + N);
  o->cgend.assign(end, end + N);
  o->cgtri.assign(entries, entries + n_entries);
  o->has_geom = true;
  if (fluid_frac) o->frac.assign(fluid_frac, fluid_frac + o->C);
  return 0;
}

int oracle_set_near_cells(oracle_sim* o, const uint8_t* flags) {
  if (!o) return -1;
  if (flags) o->nearc.assign(flags, flags + o->C);
  else o->nearc.clear();
  return 0;
}

int oracle_set_scalars(oracle_sim* o, float fnum, float vhs_k, float vhs_exp, const float inj_lambda[6]) {
  if (!o) return -1;
  o->fnum = fnum;
  o->vhs_k = vhs_k;
  o->vhs_exp = vhs_exp;
  for (int f = 0; f < 6; ++f) o->inj_lambda[f] = inj_lambda ? inj_lambda[f] : 0.0f;
  return 0;
}

int oracle_set_rng(oracle_sim* o, const uint32_t* w, uint32_t n) {
  if (!o || n < 20u * o->S) return -1;
  for (uint32_t i = 0; i < o->S; ++i) o->rng[i].from_window(&w[(size_t)i * 20]);
  return 0;
}

int oracle_get_rng(oracle_sim* o, uint32_t* w, uint32_t n) {
  if (!o || n < 20u * o->S) return -1;
  for (uint32_t i = 0; i < o->S; ++i) {
    o->rng[i].to_window(&w[(size_t)i * 20]);
    w[(size_t)i * 20 + 19] = 0u;
  }
  return 0;
}

int oracle_set_taus(oracle_sim* o, const uint32_t* w, uint32_t n) {
  if (!o || n < 4u * o->C) return -1;
  o->taus.resize(o->C);
  for (uint32_t c = 0; c < o->C; ++c) o->taus[c] = {w[4 * c], w[4 * c + 1], w[4 * c + 2], w[4 * c + 3]};
  return 0;
}

int oracle_get_taus(oracle_sim* o, uint32_t* w, uint32_t n) {
  if (!o || o->taus.empty() || n < 4u * o->C) return -1;
  for (uint32_t c = 0; c < o->C; ++c) {
    w[4 * c] = o->taus[c].x;
    w[4 * c + 1] = o->taus[c].y;
    w[4 * c + 2] = o->taus[c].z;
    w[4 * c + 3] = o->taus[c].w;
  }
  return 0;
}

// Window (20 words, the last zero) of one MT(19,9) stream after d draws.
void oracle_mt_window_after(uint32_t a, uint32_t b, uint32_t c, uint32_t seed, uint32_t d, uint32_t* out) {
  MtBlock m;
  m.a = a;
  m.b = b;
  m.c = c;
  m.seed(seed);
  for (uint32_t k = 0; k < d; ++k) m.next();
  m.to_window(out);
  out[19] = 0u;
}

// One Philox4x32-10 block (the wall re-emission generator).
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  dsmc_rng::U32x4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
  dsmc_rng::U32x4 r = dsmc_rng::philox4x32_10(c, key[0], key[1]);
  for (int k = 0; k < 4; ++k) out[k] = r.v[k];
}

// Draws of one Hybrid-Taus stream (the combined generator, or with
// component = 1..3 one Tausworthe component alone, 4 the LCG alone).
void oracle_taus_draws(const uint32_t seed[4], int component, uint32_t n, uint32_t* out) {
  TausState t{seed[0], seed[1], seed[2], seed[3]};
  for (uint32_t i = 0; i < n; ++i) {
    switch (component) {
      case 1: out[i] = TausState::step(t.x, 13, 19, 12, 4294967294ULL); break;
      case 2: out[i] = TausState::step(t.y, 2, 25, 4, 4294967288ULL); break;
      case 3: out[i] = TausState::step(t.z, 3, 11, 17, 4294967280ULL); break;
      case 4: t.w = (uint32_t)((unsigned long long)t.w * 1664525ULL + 1013904223ULL); out[i] = t.w; break;
      default: out[i] = t.next();
    }
  }
}

// Period of one Tausworthe component from state v (steps until the state
// returns; 0 if it has not after `limit` steps).
uint64_t oracle_taus_period(int component, uint32_t v, uint64_t limit) {
  static const int sh[3][3] = {{13, 19, 12}, {2, 25, 4}, {3, 11, 17}};
  static const unsigned long long mk[3] = {4294967294ULL, 4294967288ULL, 4294967280ULL};
  if (component < 1 || component > 3) return 0;
  const int* q = sh[component - 1];
  uint32_t z = v;
  TausState::step(z, q[0], q[1], q[2], mk[component - 1]);  // enter the cycle (low bits masked off)
  const uint32_t z0 = z;
  for (uint64_t k = 1; k <= limit; ++k) {
    TausState::step(z, q[0], q[1], q[2], mk[component - 1]);
    if (z == z0) return k;
  }
  return 0;
}

int oracle_set_ntc(oracle_sim* o, const float* ntc, uint32_t n) {
  if (!o || n < o->C) return -1;
  o->ntc.assign(ntc, ntc + 2 * (size_t)o->C);
  return 0;
}

int oracle_get_ntc(oracle_sim* o, float* ntc, uint32_t n) {
  if (!o || n < o->C) return -1;
  std::copy(o->ntc.begin(), o->ntc.end(), ntc);
  return 0;
}

int oracle_set_state(oracle_sim* o, const float* pos, const float* vel, uint32_t n) {
  if (!o) return -1;
  o->x.assign(pos, pos + n);
  o->y.assign(pos + n, pos + 2 * (size_t)n);
  o->z.assign(pos + 2 * (size_t)n, pos + 3 * (size_t)n);
  o->vx.assign(vel, vel + n);
  o->vy.assign(vel + n, vel + 2 * (size_t)n);
  o->vz.assign(vel + 2 * (size_t)n, vel + 3 * (size_t)n);
  o->key.resize(n);
  for (uint32_t i = 0; i < n; ++i) o->key[i] = o->key_of(o->x[i], o->y[i], o->z[i]);
  o->n_store = n;
  return 0;
}

uint32_t oracle_n_stored(oracle_sim* o) { return o ? o->n_store : 0u; }

int oracle_get_state(oracle_sim* o, float* pos, float* vel, uint32_t* keys, uint32_t cap) {
  if (!o || cap < o->n_store) return -1;
  uint32_t n = o->n_store;
  if (pos) {
    std::copy(o->x.begin(), o->x.begin() + n, pos);
    std::copy(o->y.begin(), o->y.begin() + n, pos + cap);
    std::copy(o->z.begin(), o->z.begin() + n, pos + 2 * (size_t)cap);
  }
  if (vel) {
    std::copy(o->vx.begin(), o->vx.begin() + n, vel);
    std::copy(o->vy.begin(), o->vy.begin() + n, vel + cap);
    std::copy(o->vz.begin(), o->vz.begin() + n, vel + 2 * (size_t)cap);
  }
  if (keys) std::copy(o->key.begin(), o->key.begin() + n, keys);
  return 0;
}

int oracle_load(oracle_sim* o, float ux, float uy, float uz) {
  if (!o) return -1;
  o->load(ux, uy, uz);
  return 0;
}

int oracle_ntc_cell(oracle_sim* o, uint32_t cell, uint32_t stepno, float* vx, float* vy, float* vz, uint32_t n,
                    float vol, float* nt, double* acc3) {
  if (!o || !o->phys) return -1;
  const uint32_t saved = o->stepno;
  o->stepno = stepno;
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  o->ntc_physics(vx, vy, vz, n, vol, nt, cell, a);
  o->stepno = saved;
  for (int k = 0; k < 3; ++k) acc3[k] = a[k];
  return 0;
}

uint32_t oracle_pair_perm(uint32_t n, const uint32_t keys[4], uint32_t x) {
  dsmc_rng::U32x4 k = {{keys[0], keys[1], keys[2], keys[3]}};
  return dsmc_pairing::pp_apply(dsmc_pairing::pp_make(n, k), x);
}

int oracle_step(oracle_sim* o, int sample) {
  if (!o) return -1;
  o->step(sample != 0);
  return 0;
}

int oracle_get_perm(oracle_sim* o, uint32_t* out, uint32_t n) {
  if (!o) return -1;
  std::copy(o->perm.begin(), o->perm.begin() + std::min<size_t>(n, o->perm.size()), out);
  return 0;
}

int oracle_get_cell_start(oracle_sim* o, uint32_t* out, uint32_t n) {
  if (!o || n < o->cell_start.size()) return -1;
  std::copy(o->cell_start.begin(), o->cell_start.end(), out);
  return 0;
}

int oracle_get_hits(oracle_sim* o, uint32_t* out, uint32_t n) {
  if (!o) return -1;
  std::copy(o->hits.begin(), o->hits.begin() + std::min<size_t>(n, o->hits.size()), out);
  return 0;
}

int oracle_get_samples(oracle_sim* o, double* out, uint32_t n) {
  if (!o || n < o->C) return -1;
  std::copy(o->samp.begin(), o->samp.end(), out);
  return 0;
}

int oracle_get_cstat(oracle_sim* o, float* out, uint32_t n) {
  if (!o || n < o->C) return -1;
  std::copy(o->cstat.begin(), o->cstat.end(), out);
  return 0;
}

int oracle_get_acc(oracle_sim* o, double* out, uint32_t n) {
  if (!o) return -1;
  for (uint32_t k = 0; k < n && k < 7; ++k) {
    double t = 0;
    for (uint32_t w = 0; w < o->G; ++w) t += o->acc[(size_t)w * 7 + k];
    out[k] = (k == 2 && o->phys) ? t / 65536.0 : t;  // PHYSICS c_r sums are in 2^-16 m/s
  }
  return 0;
}

int oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

}  // extern "C"
