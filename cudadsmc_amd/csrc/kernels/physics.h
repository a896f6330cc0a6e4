// Per-particle and per-cell physics of the DSMC step (device functions).
//
// REFERENCE mode reproduces the semantics of kernels/dsmc_kernel.cu of the
// reference: cell keys (GetCellID3D :1701-1710), collision-grid lookup
// (GetCollisionCellID :154-204), Moller-Trumbore (tri_intersect.cu:14-56),
// first-hit specular walls (CollideWithGeometry :1256-1314), the move loop
// (ProcessCollisionsDev :1596-1657), the box functors (:1316-1555) and the
// hard-sphere NTC cell loop (ComputeBirdCollisionsDev :1002-1089).
// PHYSICS mode adds VHS cross-sections, nearest-hit walls with diffuse
// re-emission at the wall temperature and open freestream faces.
//
// FP policy: no contraction (-ffp-contract=off); the expression order of each
// reference line is kept; transcendental functions come from detmath.h.
#pragma once
#include <hip/hip_runtime.h>

#include "../common/dsmc_dev.h"
#include "../common/pairing.h"
#include "../common/philox.h"

namespace dsmc_dev {

// ------------------------------------------------------------------ keys
// GetCellID3D (dsmc_kernel.cu:1701-1710) with the GPU's saturating
// float->uint conversion made explicit.
__device__ __forceinline__ uint32_t cell_key(const InstDesc& I, float x, float y, float z) {
  uint32_t ix = dm_f2u_sat((x - I.gmin[0]) / I.cell[0]);
  uint32_t iy = dm_f2u_sat((y - I.gmin[1]) / I.cell[1]);
  uint32_t iz = dm_f2u_sat((z - I.gmin[2]) / I.cell[2]);
  ix = min(ix, I.dim[0] - 1u);
  iy = min(iy, I.dim[1] - 1u);
  iz = min(iz, I.dim[2] - 1u);
  return iz * I.dim[0] * I.dim[1] + iy * I.dim[0] + ix;
}

// PHYSICS cell key: the offset times the float reciprocal of the cell size
// (one multiply per axis instead of an IEEE divide).
__device__ __forceinline__ uint32_t cell_key_phys(const InstDesc& I, float x, float y, float z) {
  uint32_t ix = dm_f2u_sat((x - I.gmin[0]) * I.inv_cell[0]);
  uint32_t iy = dm_f2u_sat((y - I.gmin[1]) * I.inv_cell[1]);
  uint32_t iz = dm_f2u_sat((z - I.gmin[2]) * I.inv_cell[2]);
  ix = min(ix, I.dim[0] - 1u);
  iy = min(iy, I.dim[1] - 1u);
  iz = min(iz, I.dim[2] - 1u);
  return iz * I.dim[0] * I.dim[1] + iy * I.dim[0] + ix;
}

// GetCollisionCellID (dsmc_kernel.cu:154-204).
__device__ __forceinline__ uint32_t cg_cell(const InstDesc& I, float x, float y, float z) {
  if (x < I.cg_min[0] || y < I.cg_min[1] || z < I.cg_min[2] || x > I.cg_max[0] || y > I.cg_max[1] ||
      z > I.cg_max[2])
    return kInvalidKey;
  uint32_t ix = dm_f2u_sat((x - I.cg_min[0]) / I.cg_cell[0]);
  uint32_t iy = dm_f2u_sat((y - I.cg_min[1]) / I.cg_cell[1]);
  uint32_t iz = dm_f2u_sat((z - I.cg_min[2]) / I.cg_cell[2]);
  ix = min(ix, I.cg_dim[0] - 1u);
  iy = min(iy, I.cg_dim[1] - 1u);
  iz = min(iz, I.cg_dim[2] - 1u);
  return iz * I.cg_dim[0] * I.cg_dim[1] + iy * I.cg_dim[0] + ix;
}

// ----------------------------------------------------------- Moller-Trumbore
// intersect_triangle (tri_intersect.cu:14-56): non-culling, EPSILON and
// inv_det in double, the rest in float.
// The reference's parallel test |det| < EPSILON is absolute (m^2), so it
// rejects every hit on triangles of a few mm at oblique incidence; PHYSICS
// mode (REL = true) scales the threshold with the triangle instead.
// Edges e1 = b - a, e2 = c - a.
template <bool REL = false>
__device__ __forceinline__ bool intersect_edges(float ox, float oy, float oz, float dx, float dy, float dz, float4 a,
                                                float e1x, float e1y, float e1z, float e2x, float e2y, float e2z,
                                                float& t, float& u, float& v) {
  float px = dy * e2z - dz * e2y;
  float py = dz * e2x - dx * e2z;
  float pz = dx * e2y - dy * e2x;
  float det = e1x * px + e1y * py + e1z * pz;
  const double eps = REL ? kTriRelEpsilon * ((double)(e1x * e1x + e1y * e1y + e1z * e1z) +
                                            (double)(e2x * e2x + e2y * e2y + e2z * e2z))
                         : kTriEpsilon;
  if ((double)det > -eps && (double)det < eps) return false;
  float inv_det = (float)(1.0 / (double)det);
  float tx = ox - a.x, ty = oy - a.y, tz = oz - a.z;
  u = (tx * px + ty * py + tz * pz) * inv_det;
  if ((double)u < 0.0 || (double)u > 1.0) return false;
  float qx = ty * e1z - tz * e1y;
  float qy = tz * e1x - tx * e1z;
  float qz = tx * e1y - ty * e1x;
  v = (dx * qx + dy * qy + dz * qz) * inv_det;
  if ((double)v < 0.0 || (double)(u + v) > 1.0) return false;
  t = (e2x * qx + e2y * qy + e2z * qz) * inv_det;
  return true;
}

template <bool REL = false>
__device__ __forceinline__ bool intersect_tri(float ox, float oy, float oz, float dx, float dy, float dz,
                                              float4 a, float4 b, float4 c, float& t, float& u, float& v) {
  return intersect_edges<REL>(ox, oy, oz, dx, dy, dz, a, b.x - a.x, b.y - a.y, b.z - a.z, c.x - a.x, c.y - a.y,
                              c.z - a.z, t, u, v);
}

// normalize (vectors.cu:24-31)
__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
  float l = sqrtf(x * x + y * y + z * z);
  if (l > 0.0f) {
    x = x / l;
    y = y / l;
    z = z / l;
  }
}

// CollideWithGeometry (dsmc_kernel.cu:1256-1314): first hit in list order.
__device__ __forceinline__ bool ref_collide_geometry(const InstDesc& I, float& px, float& py, float& pz,
                                                     float& vx, float& vy, float& vz, float& ox, float& oy,
                                                     float& oz, uint32_t cell) {
  uint32_t s = gmem(I.cg_start)[cell], e = gmem(I.cg_end)[cell];
  float dx = vx, dy = vy, dz = vz;
  normalize3(dx, dy, dz);
  float ddx = px - ox, ddy = py - oy, ddz = pz - oz;
  float dist = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);  // distance4 (vectors.cu:101-106)
  if (dist < kGeomEps) return false;
  for (uint32_t i = s; i < e; ++i) {
    uint32_t tri = gmem(I.cg_tri)[i];
    float4 a = gload(gmem(I.tv0) + tri), b = gload(gmem(I.tv1) + tri), c = gload(gmem(I.tv2) + tri);
    float t, u, v;
    if (intersect_tri(ox, oy, oz, dx, dy, dz, a, b, c, t, u, v)) {
      if (t >= 0.0f && t <= dist) {
        float uv = 1.0f - u - v;
        float4 n = gload(gmem(I.tn) + tri);
        float wx = n.x * kModelWidth, wy = n.y * kModelWidth, wz = n.z * kModelWidth;
        float hx = a.x * uv + b.x * u + c.x * v;
        float hy = a.y * uv + b.y * u + c.y * v;
        float hz = a.z * uv + b.z * u + c.z * v;
        // reflect4 (vectors.cu:101-107... :96-100)
        float dt = vx * n.x + vy * n.y + vz * n.z;
        vx = vx - 2.0f * dt * n.x;
        vy = vy - 2.0f * dt * n.y;
        vz = vz - 2.0f * dt * n.z;
        float rx = vx, ry = vy, rz = vz;
        normalize3(rx, ry, rz);
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
  }
  return false;
}

// Box functors of the reference (dsmc_kernel.cu:1316-1555). BOUNCE_DELTA is
// a double literal, so the wrap/clamp expressions are evaluated in double.
__device__ __forceinline__ float bd_add(float a, float b, float c) {
  // a + BOUNCE_DELTA + b - c, evaluated left to right in double
  return (float)((((double)a + kBounceDelta) + (double)b) - (double)c);
}
__device__ __forceinline__ float bd_sub(float a, float b, float c) {
  return (float)((((double)a - kBounceDelta) + (double)b) - (double)c);
}

__device__ __forceinline__ void ref_box(const InstDesc& I, float& px, float& py, float& pz, float& vx,
                                        float& vy, float& vz) {
  const float* mn = I.gmin;
  const float* mx = I.gmax;
  const double D = kBounceDelta;
  switch (I.periodic) {
    case 'x':  // PeriodicXBoundaryCondition (:1392-1456)
      if (px > mx[0]) px = bd_add(mn[0], px, mx[0]);
      if (py > mx[1]) { py = (float)((double)mx[1] - D); vy *= -1.0f; }
      if (px < mn[0]) px = bd_sub(mx[0], px, mn[0]);
      if (py < mn[1]) { py = (float)((double)mn[1] + D); vy *= -1.0f; }
      if (pz < mn[2]) { pz = (float)((double)mn[2] + D); vz *= -1.0f; }
      if (pz > mx[2]) { pz = (float)((double)mx[2] - D); vz *= -1.0f; }
      break;
    case 'y':  // PeriodicYBoundaryCondition (:1458-1522)
      if (py > mx[1]) py = bd_add(mn[1], py, mx[1]);
      if (px > mx[0]) { px = (float)((double)mx[0] - D); vx *= -1.0f; }
      if (py < mn[1]) py = bd_sub(mx[1], py, mn[1]);
      if (px < mn[0]) { px = (float)((double)mn[0] + D); vx *= -1.0f; }
      if (pz < mn[2]) { pz = (float)((double)mn[2] + D); vz *= -1.0f; }
      if (pz > mx[2]) { pz = (float)((double)mx[2] - D); vz *= -1.0f; }
      break;
    case 'z':  // PeriodicZBoundaryCondition (:1316-1390)
      if (pz > mx[2]) pz = bd_add(mn[2], pz, mx[2]);
      if (py > mx[1]) { py = (float)((double)mx[1] - D); vy *= -1.0f; }
      if (pz < mn[2]) pz = bd_sub(mx[2], pz, mn[2]);
      if (py < mn[1]) { py = (float)((double)mn[1] + D); vy *= -1.0f; }
      if (px < mn[0]) { px = (float)((double)mn[0] + D); vx *= -1.0f; }
      if (px > mx[0]) { px = (float)((double)mx[0] - D); vx *= -1.0f; }
      break;
    default:  // NoPeriodicBoundaryCondition (:1524-1555)
      if (py > mx[1]) { py = (float)((double)mx[1] - D); vy *= -1.0f; }
      if (px > mx[0]) { px = (float)((double)mx[0] - D); vx *= -1.0f; }
      if (py < mn[1]) { py = (float)((double)mn[1] + D); vy *= -1.0f; }
      if (px < mn[0]) { px = (float)((double)mn[0] + D); vx *= -1.0f; }
      if (pz < mn[2]) { pz = (float)((double)mn[2] + D); vz *= -1.0f; }
      if (pz > mx[2]) { pz = (float)((double)mx[2] - D); vz *= -1.0f; }
      break;
  }
}

// ProcessCollisionsDev / IntegrateDev body (dsmc_kernel.cu:1557-1657).
// Returns the number of wall reflections.
__device__ __forceinline__ uint32_t move_reference(const InstDesc& I, float& px, float& py, float& pz,
                                                   float& vx, float& vy, float& vz, float dt) {
  float ox = px, oy = py, oz = pz;
  float nx = ox + vx * dt, ny = oy + vy * dt, nz = oz + vz * dt;
  uint32_t hits = 0;
  if (I.has_geom) {
    bool coll = false;  // not reset inside the loop, as in the reference
    int i = 0;
    do {
      uint32_t st = cg_cell(I, ox, oy, oz);
      uint32_t en = cg_cell(I, nx, ny, nz);
      if (st != kInvalidKey) {
        if (gmem(I.cg_flag)[st] != 0u) coll = ref_collide_geometry(I, nx, ny, nz, vx, vy, vz, ox, oy, oz, st);
      }
      if (!coll && st != en && en != kInvalidKey) {
        if (gmem(I.cg_flag)[en] != 0u) coll = ref_collide_geometry(I, nx, ny, nz, vx, vy, vz, ox, oy, oz, en);
      }
      if (coll) ++hits;
    } while (coll && i++ < kCollisionLimit);
  }
  ref_box(I, nx, ny, nz, vx, vy, vz);
  px = nx;
  py = ny;
  pz = nz;
  return hits;
}

// The same flight for the 64 lanes of a wavefront together (lanes without a
// particle pass act = false). Each call of CollideWithGeometry on a cell
// searches that cell's triangle list for its first hit in list order; here
// the lists of all lanes that search in the same call form one pool, each
// lane tests one entry of it per round, and a lane's first hit is the
// smallest entry index among its hits (an LDS atomic minimum over the
// wavefront's `first` array, 64 entries). The hit is then recomputed and
// applied by its lane alone, exactly as the one-lane loop would.
__device__ __forceinline__ bool ref_collide_wave(const InstDesc& I, bool want, float& px, float& py, float& pz,
                                                 float& vx, float& vy, float& vz, float& ox, float& oy, float& oz,
                                                 uint32_t cell, uint32_t* first) {
  const uint32_t lane = __lane_id();
  uint32_t s = 0u, n = 0u;
  float dx = vx, dy = vy, dz = vz;
  normalize3(dx, dy, dz);
  const float ddx = px - ox, ddy = py - oy, ddz = pz - oz;
  const float dist = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);  // distance4 (vectors.cu:101-106)
  if (want && !(dist < kGeomEps)) {
    s = gmem(I.cg_start)[cell];
    n = gmem(I.cg_end)[cell] - s;
  }
  uint32_t incl = n;
#pragma unroll
  for (uint32_t o = 1; o < 64u; o <<= 1) {
    const uint32_t x = __shfl_up(incl, o);
    if (lane >= o) incl += x;
  }
  const uint32_t total = __shfl(incl, 63);
  if (total == 0u) return false;  // uniform over the wavefront
  const uint32_t off = incl - n;
  if (n) first[lane] = 0xFFFFFFFFu;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
  for (uint32_t k0 = 0; k0 < total; k0 += 64u) {
    const uint32_t k = k0 + lane < total ? k0 + lane : total - 1u;
    uint32_t o = 0;  // owner: the last lane whose first entry number is <= k
#pragma unroll
    for (uint32_t step = 32; step; step >>= 1) {
      const uint32_t m = o + step;
      if (m < 64u && (uint32_t)__shfl(off, (int)(m & 63u)) <= k) o = m;
    }
    const uint32_t e = __shfl(s, (int)o) + (k - __shfl(off, (int)o));
    const float qx = __shfl(ox, (int)o), qy = __shfl(oy, (int)o), qz = __shfl(oz, (int)o);
    const float ex = __shfl(dx, (int)o), ey = __shfl(dy, (int)o), ez = __shfl(dz, (int)o);
    const float qd = __shfl(dist, (int)o);
    if (k0 + lane < total) {
      const uint32_t tri = gmem(I.cg_tri)[e];
      float t, u, v;
      const float4 a = gload(gmem(I.tv0) + tri), b = gload(gmem(I.tv1) + tri), c = gload(gmem(I.tv2) + tri);
      if (intersect_tri(qx, qy, qz, ex, ey, ez, a, b, c, t, u, v) && t >= 0.0f && t <= qd)
        atomicMin(&first[o], e);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
  if (!n) return false;
  const uint32_t e = first[lane];
  if (e == 0xFFFFFFFFu) return false;
  const uint32_t tri = gmem(I.cg_tri)[e];
  const float4 a = gload(gmem(I.tv0) + tri), b = gload(gmem(I.tv1) + tri), c = gload(gmem(I.tv2) + tri);
  float t, u, v;
  intersect_tri(ox, oy, oz, dx, dy, dz, a, b, c, t, u, v);
  const float uv = 1.0f - u - v;
  const float4 nrm = gload(gmem(I.tn) + tri);
  const float wx = nrm.x * kModelWidth, wy = nrm.y * kModelWidth, wz = nrm.z * kModelWidth;
  const float hx = a.x * uv + b.x * u + c.x * v;
  const float hy = a.y * uv + b.y * u + c.y * v;
  const float hz = a.z * uv + b.z * u + c.z * v;
  const float dtn = vx * nrm.x + vy * nrm.y + vz * nrm.z;
  vx = vx - 2.0f * dtn * nrm.x;
  vy = vy - 2.0f * dtn * nrm.y;
  vz = vz - 2.0f * dtn * nrm.z;
  float rx = vx, ry = vy, rz = vz;
  normalize3(rx, ry, rz);
  const float od = dist - t;
  px = hx + rx * od;
  py = hy + ry * od;
  pz = hz + rz * od;
  ox = hx + wx;
  oy = hy + wy;
  oz = hz + wz;
  return true;
}

// move_reference for the 64 lanes of a wavefront together (see
// ref_collide_wave); lanes without a particle pass act = false.
__device__ __forceinline__ uint32_t move_reference_wave(const InstDesc& I, bool act, float& px, float& py,
                                                        float& pz, float& vx, float& vy, float& vz, float dt,
                                                        uint32_t* first) {
  float ox = px, oy = py, oz = pz;
  float nx = ox + vx * dt, ny = oy + vy * dt, nz = oz + vz * dt;
  uint32_t hits = 0;
  if (I.has_geom) {
    bool coll = false;  // not reset inside the loop, as in the reference
    int i = 0;
    bool live = act;
    while (__ballot(live) != 0ull) {
      uint32_t st = kInvalidKey, en = kInvalidKey;
      if (live) {
        st = cg_cell(I, ox, oy, oz);
        en = cg_cell(I, nx, ny, nz);
      }
      const bool w1 = live && st != kInvalidKey && gmem(I.cg_flag)[st] != 0u;
      const bool c1 = ref_collide_wave(I, w1, nx, ny, nz, vx, vy, vz, ox, oy, oz, st, first);
      if (w1) coll = c1;
      const bool w2 = live && !coll && st != en && en != kInvalidKey && gmem(I.cg_flag)[en] != 0u;
      const bool c2 = ref_collide_wave(I, w2, nx, ny, nz, vx, vy, vz, ox, oy, oz, en, first);
      if (w2) coll = c2;
      if (live) {
        if (coll) ++hits;
        live = coll && i++ < kCollisionLimit;
      }
    }
  }
  if (act) {
    ref_box(I, nx, ny, nz, vx, vy, vz);
    px = nx;
    py = ny;
    pz = nz;
  }
  return hits;
}

// ---------------------------------------------------------------- physics
// Nearest front-facing triangle hit along o + s*d, 0 <= s <= dist, over the
// collision-grid cells the segment crosses (3-D DDA). Ties on s go to the
// lower triangle index.
struct DdaAxis {
  int c, step;
  float tmax, tdel;
};

__device__ __forceinline__ bool dda_clip(float o, float d, float lo, float hi, float& t0, float& t1) {
  if (d == 0.0f) return !(o < lo || o > hi);
  float ta = (lo - o) / d;
  float tb = (hi - o) / d;
  if (ta > tb) { float tmp = ta; ta = tb; tb = tmp; }
  t0 = ta > t0 ? ta : t0;
  t1 = tb < t1 ? tb : t1;
  return true;
}

__device__ __forceinline__ DdaAxis dda_axis(float o, float d, float t0, float lo, float cell, uint32_t dim) {
  DdaAxis r;
  float p = o + d * t0;
  float f = (p - lo) / cell;
  int ci = (int)dm_floorf(f);
  ci = ci < 0 ? 0 : (ci >= (int)dim ? (int)dim - 1 : ci);
  r.c = ci;
  if (d > 0.0f) {
    r.step = 1;
    float bnd = lo + (float)(ci + 1) * cell;
    r.tmax = (bnd - o) / d;
    r.tdel = cell / d;
  } else if (d < 0.0f) {
    r.step = -1;
    float bnd = lo + (float)ci * cell;
    r.tmax = (bnd - o) / d;
    r.tdel = -cell / d;
  } else {
    r.step = 0;
    r.tmax = 3.0e38f;
    r.tdel = 3.0e38f;
  }
  return r;
}

// The wall pass runs the DDA of a wavefront's 64 particles in lockstep: at
// each step every live lane has one collision-grid cell, and the triangle
// entries of all those cells are tested together, one entry per lane (the
// entries of cell c of lane o are numbered after those of lanes < o), so a
// cell with hundreds of entries costs a few wavefront-wide rounds rather
// than hundreds of dependent loads on one lane. The nearest hit of each
// lane in the step is the minimum of (s, triangle index) over its cell's
// entries, kept as a 64-bit key in LDS (s >= 0, whose bits order as
// integers; -0 is keyed as +0, as s == best treats them).
struct WaveHitLds {
  unsigned long long* key;  // [64] of the calling wavefront
};

__device__ __forceinline__ uint64_t hit_key(float t, uint32_t tri) {
  return ((uint64_t)__float_as_uint(t + 0.0f) << 32) | tri;
}

// Entry k of a cell's records tested against the segment o + s d, s in
// [0, dist]: front-facing, not the triangle just left.
__device__ __forceinline__ bool test_entry(const InstDesc& I, uint32_t k, float ox, float oy, float oz, float dx,
                                           float dy, float dz, float dist, uint32_t skip, float& t, uint32_t& tri) {
  const gptr<const float4> rec = gmem<const float4>(I.cg_rec) + 4u * k;
  const float4 r0 = gload(rec), r1 = gload(rec + 1), r2 = gload(rec + 2), n = gload(rec + 3);
  tri = __float_as_uint(r0.w);
  if (tri == skip) return false;
  if (!(dx * n.x + dy * n.y + dz * n.z < 0.0f)) return false;
  float u, v;
  if (!intersect_edges<true>(ox, oy, oz, dx, dy, dz, r0, r1.x, r1.y, r1.z, r2.x, r2.y, r2.z, t, u, v)) return false;
  return t >= 0.0f && t <= dist;
}

// Nearest front-facing triangle hit along o + s*d, 0 <= s <= dist, over the
// collision-grid cells the segment crosses (3-D DDA), for the lanes with
// `act`; ties on s go to the lower triangle index. Called by all 64 lanes
// of the wavefront together.
__device__ __forceinline__ bool wave_nearest_hit(const InstDesc& I, bool act, float ox, float oy, float oz,
                                                 float dx, float dy, float dz, float dist, uint32_t skip,
                                                 float& s_hit, uint32_t& tri_hit, const WaveHitLds& L) {
  const uint32_t lane = __lane_id();
  float t0 = 0.0f, t1 = dist;
  bool on = act;
  if (on) {
    on = dda_clip(ox, dx, I.cg_min[0], I.cg_max[0], t0, t1) && dda_clip(oy, dy, I.cg_min[1], I.cg_max[1], t0, t1) &&
         dda_clip(oz, dz, I.cg_min[2], I.cg_max[2], t0, t1) && !(t0 > t1);
  }
  DdaAxis ax = {0, 0, 0.0f, 0.0f}, ay = ax, az = ax;
  if (on) {
    ax = dda_axis(ox, dx, t0, I.cg_min[0], I.cg_cell[0], I.cg_dim[0]);
    ay = dda_axis(oy, dy, t0, I.cg_min[1], I.cg_cell[1], I.cg_dim[1]);
    az = dda_axis(oz, dz, t0, I.cg_min[2], I.cg_cell[2], I.cg_dim[2]);
  }
  bool found = false;
  float best = 0.0f;
  uint32_t best_tri = 0;
  const int max_iter = (int)(I.cg_dim[0] + I.cg_dim[1] + I.cg_dim[2]) + 3;
  int it = 0;
  while (__ballot(on) != 0ull) {
    // this step's cell of each live lane and its entry range
    uint32_t s = 0u, n = 0u;
    if (on) {
      const uint32_t cell =
          (uint32_t)az.c * I.cg_dim[0] * I.cg_dim[1] + (uint32_t)ay.c * I.cg_dim[0] + (uint32_t)ax.c;
      const uint2 se = gload(gmem<const uint2>(I.cg_se) + cell);
      s = se.x;
      n = se.y - se.x;
    }
    uint32_t incl = n;
#pragma unroll
    for (uint32_t o = 1; o < 64u; o <<= 1) {
      const uint32_t x = __shfl_up(incl, o);
      if (lane >= o) incl += x;
    }
    const uint32_t off = incl - n;
    const uint32_t total = __shfl(incl, 63);
    if (total != 0u) {  // uniform over the wavefront
      if (n) L.key[lane] = ~0ull;
      // the keys are set before any lane's minimum reaches them
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      __builtin_amdgcn_wave_barrier();
      for (uint32_t k0 = 0; k0 < total; k0 += 64u) {
        const uint32_t k = k0 + lane < total ? k0 + lane : total - 1u;
        // owner: the last lane whose first entry number is <= k
        uint32_t o = 0;
#pragma unroll
        for (uint32_t step = 32; step; step >>= 1) {
          const uint32_t m = o + step;
          const uint32_t om = __shfl(off, (int)(m & 63u));
          if (m < 64u && om <= k) o = m;
        }
        const uint32_t e = __shfl(s, (int)o) + (k - __shfl(off, (int)o));
        const float qx = __shfl(ox, (int)o), qy = __shfl(oy, (int)o), qz = __shfl(oz, (int)o);
        const float ex = __shfl(dx, (int)o), ey = __shfl(dy, (int)o), ez = __shfl(dz, (int)o);
        const float qd = __shfl(dist, (int)o);
        const uint32_t qs = __shfl(skip, (int)o);
        float t;
        uint32_t tri;
        if (k0 + lane < total && test_entry(I, e, qx, qy, qz, ex, ey, ez, qd, qs, t, tri))
          atomicMin(&L.key[o], hit_key(t, tri));
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      __builtin_amdgcn_wave_barrier();
      if (n) {
        const unsigned long long key = L.key[lane];
        if (key != ~0ull) {
          const float t = __uint_as_float((uint32_t)(key >> 32));
          const uint32_t tri = (uint32_t)key;
          if (!found || t < best || (t == best && tri < best_tri)) {
            found = true;
            best = t;
            best_tri = tri;
            if (t == 0.0f) {
              // the sign of a zero s as computed (the key holds +0)
              for (uint32_t q = s; q < s + n; ++q) {
                float tq;
                uint32_t trq;
                if (test_entry(I, q, ox, oy, oz, dx, dy, dz, dist, skip, tq, trq) && trq == tri) {
                  best = tq;
                  break;
                }
              }
            }
          }
        }
      }
    }
    if (on) {
      const bool xmin = ax.tmax < ay.tmax ? ax.tmax < az.tmax : false;
      const bool ymin = !xmin && (ay.tmax < az.tmax);
      const float texit = xmin ? ax.tmax : (ymin ? ay.tmax : az.tmax);
      if ((found && best <= texit) || texit > t1) {
        on = false;
      } else if (xmin) {
        ax.c += ax.step;
        ax.tmax = ax.tmax + ax.tdel;
        on = ax.c >= 0 && ax.c < (int)I.cg_dim[0];
      } else if (ymin) {
        ay.c += ay.step;
        ay.tmax = ay.tmax + ay.tdel;
        on = ay.c >= 0 && ay.c < (int)I.cg_dim[1];
      } else {
        az.c += az.step;
        az.tmax = az.tmax + az.tdel;
        on = az.c >= 0 && az.c < (int)I.cg_dim[2];
      }
      if (++it >= max_iter) on = false;
    }
  }
  s_hit = best;
  tri_hit = best_tri;
  return found;
}

// Orthonormal tangent pair of a unit normal.
__device__ __forceinline__ void tangents(float nx, float ny, float nz, float* t1, float* t2) {
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  float fx = nx < 0.0f ? -nx : nx, fy = ny < 0.0f ? -ny : ny, fz = nz < 0.0f ? -nz : nz;
  if (fx <= fy && fx <= fz) ax = 1.0f;
  else if (fy <= fz) ay = 1.0f;
  else az = 1.0f;
  float cx = ny * az - nz * ay, cy = nz * ax - nx * az, cz = nx * ay - ny * ax;
  normalize3(cx, cy, cz);
  t1[0] = cx; t1[1] = cy; t1[2] = cz;
  t2[0] = ny * cz - nz * cy;
  t2[1] = nz * cx - nx * cz;
  t2[2] = nx * cy - ny * cx;
}

// One axis of the PHYSICS box: returns false when the particle leaves
// through a freestream face.
__device__ __forceinline__ bool phys_face(float& q, float& v, float lo, float hi, uint32_t f_lo, uint32_t f_hi) {
  if (q < lo) {
    if (f_lo == kFaceFreestream) return false;
    if (f_lo == kFacePeriodic) q = q + (hi - lo);
    else { q = lo + (lo - q); v = -v; }
  } else if (q > hi) {
    if (f_hi == kFaceFreestream) return false;
    if (f_hi == kFacePeriodic) q = q - (hi - lo);
    else { q = hi - (q - hi); v = -v; }
  }
  q = q < lo ? lo : (q > hi ? hi : q);
  return true;
}

// Wall sums in 64-bit fixed point (momentum change in units of
// m v_ref / 2^36, energy change in units of m v_ref^2 / 2^34): integer
// addition does not depend on the order of the terms, so the sums are the
// same for any schedule of the wall pass.
struct WallAcc {
  long long fx, fy, fz, heat;
};

// PHYSICS move. A particle is "near" when it moves a cell or more along some
// axis in one step, or when its cell has a near-body code (the distance from
// the cell to the body's collision-grid cells, in units of near_unit) and
// its flight |v| dt is at least that distance. Its wall interaction is then
// deferred to the wall pass (k_walls), which runs after the tile kernel
// with the particle's state as left there. Every other particle can reach
// no triangle and only flies; for a near particle that hits nothing the
// wall pass computes exactly the same flight, so the split changes no
// result.
__device__ __forceinline__ bool phys_near(const InstDesc& I, uint32_t code, float vx, float vy, float vz,
                                          float dt) {
  if (!I.has_geom) return false;
  if (fabsf(vx) * dt >= I.cell[0] || fabsf(vy) * dt >= I.cell[1] || fabsf(vz) * dt >= I.cell[2]) return true;
  if (code == 255u) return false;
  const float s2 = (vx * vx + vy * vy + vz * vz) * (dt * dt);
  const float d = (float)code * I.near_unit;
  return s2 >= d * d;
}

// Box faces after the flight: the new cell key, or kInvalidKey when the
// particle left through a freestream face.
__device__ __forceinline__ uint32_t phys_faces_key(const InstDesc& I, float& px, float& py, float& pz, float& vx,
                                                   float& vy, float& vz) {
  if (!phys_face(px, vx, I.gmin[0], I.gmax[0], I.face_bc[0], I.face_bc[1])) return kInvalidKey;
  if (!phys_face(py, vy, I.gmin[1], I.gmax[1], I.face_bc[2], I.face_bc[3])) return kInvalidKey;
  if (!phys_face(pz, vz, I.gmin[2], I.gmax[2], I.face_bc[4], I.face_bc[5])) return kInvalidKey;
  return cell_key_phys(I, px, py, pz);
}

// Free flight of a particle that is not near the body.
__device__ __forceinline__ uint32_t move_free(const InstDesc& I, float& px, float& py, float& pz, float& vx,
                                              float& vy, float& vz, float dt) {
  px = px + vx * dt;
  py = py + vy * dt;
  pz = pz + vz * dt;
  return phys_faces_key(I, px, py, pz, vx, vy, vz);
}

// Flight of a near particle: nearest-hit walls (diffuse at wall_temp, else
// specular), up to kPhysMaxBounces, then the box faces. The diffuse
// re-emission of bounce b draws Philox4x32-10 at counter (slot, step, 0, b)
// under key (seed, instance salt): the draws depend on which particle
// bounces when, not on which thread handles it. Called by all 64 lanes of
// the wavefront together; lanes without a particle pass act = false (and
// get kInvalidKey).
__device__ __forceinline__ uint32_t move_walls(const InstDesc& I, bool act, float& px, float& py, float& pz,
                                               float& vx, float& vy, float& vz, float dt, uint32_t slot,
                                               uint32_t stepno, uint32_t& hits, WallAcc& acc, const WaveHitLds& L) {
  hits = 0;
  float tl = dt;
  uint32_t skip = kInvalidKey;
  bool fly = act;
  for (int b = 0; b < kPhysMaxBounces; ++b) {
    float sp = 0.0f, dist = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f;
    if (fly) {
      sp = sqrtf(vx * vx + vy * vy + vz * vz);
      dist = sp * tl;
      fly = dist > 0.0f;
      if (fly) {
        dx = vx / sp;
        dy = vy / sp;
        dz = vz / sp;
      }
    }
    if (__ballot(fly) == 0ull) break;  // uniform over the wavefront
    float s;
    uint32_t tri;
    const bool hit = wave_nearest_hit(I, fly, px, py, pz, dx, dy, dz, dist, skip, s, tri, L);
    if (!fly) continue;
    if (!hit) {
      fly = false;
      continue;
    }
    float4 n = gload(gmem<const float4>(I.tn) + tri);
    float hx = px + dx * s, hy = py + dy * s, hz = pz + dz * s;
    float ux, uy, uz;
    if (I.vm_wall > 0.0f) {
      dsmc_rng::U32x4 c = {{slot, stepno, 0u, (uint32_t)b}};
      dsmc_rng::U32x4 r = dsmc_rng::philox4x32_10(c, I.wall_key[0], I.wall_key[1]);
      float r1 = dsmc_rng::px_unit(r.v[0]), r2 = dsmc_rng::px_unit(r.v[1]), r3 = dsmc_rng::px_unit(r.v[2]);
      float vn = I.vm_wall * sqrtf(-dm_logf(r1));
      float rt = I.vm_wall * sqrtf(-dm_logf(r2));
      float sn, cs;
      dm_sincosf(6.28318530717958647692f * r3, &sn, &cs);
      float a1 = rt * cs, a2 = rt * sn;
      float t1[3], t2[3];
      tangents(n.x, n.y, n.z, t1, t2);
      ux = n.x * vn + t1[0] * a1 + t2[0] * a2;
      uy = n.y * vn + t1[1] * a1 + t2[1] * a2;
      uz = n.z * vn + t1[2] * a1 + t2[2] * a2;
    } else {
      float dd = vx * n.x + vy * n.y + vz * n.z;
      ux = vx - 2.0f * dd * n.x;
      uy = vy - 2.0f * dd * n.y;
      uz = vz - 2.0f * dd * n.z;
    }
    acc.fx += __double2ll_rn(((double)vx - (double)ux) * I.wall_scale[0]);
    acc.fy += __double2ll_rn(((double)vy - (double)uy) * I.wall_scale[0]);
    acc.fz += __double2ll_rn(((double)vz - (double)uz) * I.wall_scale[0]);
    acc.heat += __double2ll_rn((((double)vx * vx + (double)vy * vy + (double)vz * vz) -
                                ((double)ux * ux + (double)uy * uy + (double)uz * uz)) * I.wall_scale[1]);
    ++hits;
    tl = tl - s / sp;
    const float off = 1.0e-6f;
    px = hx + n.x * off;
    py = hy + n.y * off;
    pz = hz + n.z * off;
    vx = ux;
    vy = uy;
    vz = uz;
    skip = tri;
  }
  if (!act) return kInvalidKey;
  if (tl > 0.0f) {
    px = px + vx * tl;
    py = py + vy * tl;
    pz = pz + vz * tl;
  }
  return phys_faces_key(I, px, py, pz, vx, vy, vz);
}

// ------------------------------------------------------------------- NTC
// Isotropic post-collision velocities (dsmc_kernel.cu:1057-1078).
// cos(theta) = 2 u1 - 1 and azimuth 2 pi u2 from uniforms u1, u2 in (0, 1].
__device__ __forceinline__ void scatter_uv(float* vx, float* vy, float* vz, uint32_t i, uint32_t j, float mag,
                                           float u1, float u2) {
  float v1x = vx[i], v1y = vy[i], v1z = vz[i];
  float v2x = vx[j], v2y = vy[j], v2z = vz[j];
  float b = 2.0f * u1 - 1.0f;
  float a = sqrtf(1.0f - b * b);
  float c = 6.28318548202514648438f * u2;  // 2*DSMC_PI in float
  float sn, cs;
  dm_sincosf(c, &sn, &cs);
  float ax = b * mag / 2.0f, ay = a * mag * cs / 2.0f, az = a * mag * sn / 2.0f;
  vx[i] = (v1x + v2x) / 2.0f + ax;
  vy[i] = (v1y + v2y) / 2.0f + ay;
  vz[i] = (v1z + v2z) / 2.0f + az;
  vx[j] = (v1x + v2x) / 2.0f - ax;
  vy[j] = (v1y + v2y) / 2.0f - ay;
  vz[j] = (v1z + v2z) / 2.0f - az;
}

template <class R>
__device__ __forceinline__ void scatter_pair(float* vx, float* vy, float* vz, uint32_t i, uint32_t j, float mag,
                                             R& rng) {
  const float u1 = rng_f(rng);
  const float u2 = rng_f(rng);
  scatter_uv(vx, vy, vz, i, j, mag, u1, u2);
}

// ComputeBirdCollisionsDev / ComputeBirdCollisionsNoMTDev cell loop
// (dsmc_kernel.cu:1002-1089 with MT, 912-1000 with Hybrid Taus) on the n
// particles of one cell (velocity arrays already offset to the cell).
template <class R>
__device__ __forceinline__ void ntc_reference(const InstDesc& I, float* vx, float* vy, float* vz, uint32_t n,
                                              float2& nt, float4& st, R& rng) {
  int part = (int)n;
  if (part < 2) return;
  float fpc = (float)((0.5 * (double)part * (double)I.avg_conc * (double)I.fnum * (double)nt.x *
                       (double)I.dt) / (double)I.cell_volume + (double)nt.y);
  uint32_t pairs = dm_f2u_sat(fpc);
  st.x += (float)pairs;
  nt.y = fpc - (float)pairs;
  if (pairs > kMaxPairsPerCell) {
    pairs = kMaxPairsPerCell;
    atomicOr(&I.ctr[3], kStatusPairs);
  }
  for (uint32_t k = 0; k < pairs; ++k) {
    uint32_t i = rng_u32(rng) % (uint32_t)part;
    uint32_t j = rng_u32(rng) % (uint32_t)part;
    float dx = vx[i] - vx[j], dy = vy[i] - vy[j], dz = vz[i] - vz[j];
    float mag = sqrtf(dx * dx + dy * dy + dz * dz);
    float v = mag * I.sigma_hs;
    if (v / nt.x > rng_f(rng)) {
      st.y += 1.0f;
      st.z += mag;
      scatter_pair(vx, vy, vz, i, j, mag, rng);
    }
    if (v > nt.x) nt.x = v;
  }
}

// Collision counters of a lane: candidates, collisions and the sum of the
// collisions' relative speeds in units of 2^-16 m/s (kCrScale). Integer sums
// do not depend on the order of their terms, so the totals are the same for
// any assignment of tiles to workgroups.
struct CollAcc {
  unsigned long long sel, coll;
  long long cr;
};


// ---------------------------------------------------------- body interior
// Point inside the body: collision-grid cell filled, or, in a surface cell,
// behind the triangle of that cell whose centroid is nearest.
__device__ __forceinline__ bool inside_body(const InstDesc& I, float x, float y, float z) {
  if (!I.has_geom) return false;
  uint32_t c = cg_cell(I, x, y, z);
  if (c == kInvalidKey) return false;
  uint32_t f = gmem(I.cg_flag)[c];
  if (f == kCgFilled) return true;
  if (f != kCgSurface) return false;
  float best = 3.0e38f, sd = 1.0f;
  for (uint32_t i = gmem(I.cg_start)[c]; i < gmem(I.cg_end)[c]; ++i) {
    uint32_t tri = gmem(I.cg_tri)[i];
    float4 a = gload(gmem(I.tv0) + tri), b = gload(gmem(I.tv1) + tri), cc = gload(gmem(I.tv2) + tri);
    float4 n = gload(gmem(I.tn) + tri);
    float cx = (a.x + b.x + cc.x) / 3.0f, cy = (a.y + b.y + cc.y) / 3.0f, cz = (a.z + b.z + cc.z) / 3.0f;
    float rx = x - cx, ry = y - cy, rz = z - cz;
    float d2 = rx * rx + ry * ry + rz * rz;
    if (d2 < best) {
      best = d2;
      sd = rx * n.x + ry * n.y + rz * n.z;
    }
  }
  return sd < 0.0f;
}

}  // namespace dsmc_dev
