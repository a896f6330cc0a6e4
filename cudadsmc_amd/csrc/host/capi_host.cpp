// Host-only part of the C ABI: mesh generators, the .raw loader and the
// collision-grid builder. None of these touch the GPU.
#include <algorithm>
#include <cstring>
#include <string>

#include "dsmc/dsmc_c.h"
#include "host_internal.h"

using namespace dsmc_host;

static int export_mesh(const Mesh& m, float* verts, uint32_t vcap, uint32_t* tris, uint32_t tcap, float* normals,
                       uint32_t* nv, uint32_t* nt) {
  if (nv) *nv = m.nv();
  if (nt) *nt = m.nt();
  if ((verts && vcap < m.nv()) || (tris && tcap < m.nt()) || (normals && tcap < m.nt()))
    return set_global_error(DSMC_ERR_CAPACITY, "mesh buffers too small");
  if (verts) std::copy(m.v.begin(), m.v.end(), verts);
  if (tris) std::copy(m.t.begin(), m.t.end(), tris);
  if (normals) std::copy(m.n.begin(), m.n.end(), normals);
  return DSMC_OK;
}

extern "C" {

int dsmc_mesh_icosphere(float cx, float cy, float cz, float r, uint32_t subdiv, float* verts, uint32_t vcap,
                        uint32_t* tris, uint32_t tcap, float* normals, uint32_t* nv, uint32_t* nt) {
  if (!(r > 0.0f) || subdiv > 8) return set_global_error(DSMC_ERR_INVALID, "bad icosphere parameters");
  Mesh m;
  mesh_icosphere(m, cx, cy, cz, r, subdiv);
  return export_mesh(m, verts, vcap, tris, tcap, normals, nv, nt);
}

int dsmc_mesh_sphere_cone(float x0, float cy, float cz, float rn, float half_angle_deg, float rb, uint32_t n_az,
                          uint32_t n_profile, float* verts, uint32_t vcap, uint32_t* tris, uint32_t tcap,
                          float* normals, uint32_t* nv, uint32_t* nt) {
  if (!(rn > 0.0f) || !(rb > rn * 0.5f) || !(half_angle_deg > 0.0f && half_angle_deg < 90.0f) || n_az < 3 ||
      n_profile < 6)
    return set_global_error(DSMC_ERR_INVALID, "bad sphere-cone parameters");
  Mesh m;
  mesh_sphere_cone(m, x0, cy, cz, rn, half_angle_deg, rb, n_az, n_profile);
  return export_mesh(m, verts, vcap, tris, tcap, normals, nv, nt);
}

int dsmc_mesh_load_raw(const char* path, float* verts, uint32_t vcap, uint32_t* tris, uint32_t tcap,
                       float* normals, uint32_t* nv, uint32_t* nt) {
  if (!path) return set_global_error(DSMC_ERR_INVALID, "null path");
  Mesh m;
  std::string err;
  if (!mesh_load_raw(m, path, err)) return set_global_error(DSMC_ERR_IO, err);
  return export_mesh(m, verts, vcap, tris, tcap, normals, nv, nt);
}

int dsmc_tri_box_overlap(const float center[3], const float half[3], const float a[3], const float b[3],
                         const float c[3]) {
  return tri_box_overlap(center, half, a, b, c) ? 1 : 0;
}

int dsmc_cgrid_build(int mode, const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                     const float* normals, uint32_t gx, uint32_t gy, uint32_t gz, float bounds[9],
                     uint32_t dims[3], uint32_t* flags, uint32_t* start, uint32_t* end, uint32_t* entries,
                     uint32_t cap_cells, uint32_t cap_entries, uint32_t* n_entries) {
  if (!verts || !tris || nv == 0 || nt == 0 || gx == 0 || gy == 0 || gz == 0)
    return set_global_error(DSMC_ERR_INVALID, "empty mesh or grid");
  Mesh m;
  m.v.assign(verts, verts + 3ull * nv);
  m.t.assign(tris, tris + 3ull * nt);
  for (uint32_t i : m.t)
    if (i >= nv) return set_global_error(DSMC_ERR_INVALID, "triangle index out of range");
  if (normals) m.n.assign(normals, normals + 3ull * nt);
  else m.compute_normals();
  CGrid g;
  if (mode == DSMC_MODE_PHYSICS) build_cgrid_physics(m, gx, gy, gz, g);
  else build_cgrid_reference(m, gx, gy, gz, g);
  if (bounds)
    for (int k = 0; k < 3; ++k) {
      bounds[k] = g.mn[k];
      bounds[3 + k] = g.mx[k];
      bounds[6 + k] = g.cell[k];
    }
  if (dims)
    for (int k = 0; k < 3; ++k) dims[k] = g.dim[k];
  if (n_entries) *n_entries = (uint32_t)g.tri.size();
  if ((flags || start || end) && cap_cells < g.ncells()) return set_global_error(DSMC_ERR_CAPACITY, "cell buffers too small");
  if (entries && cap_entries < g.tri.size()) return set_global_error(DSMC_ERR_CAPACITY, "entry buffer too small");
  if (flags) std::copy(g.flag.begin(), g.flag.end(), flags);
  if (start) std::copy(g.start.begin(), g.start.end(), start);
  if (end) std::copy(g.end.begin(), g.end.end(), end);
  if (entries) std::copy(g.tri.begin(), g.tri.end(), entries);
  return DSMC_OK;
}

}  // extern "C"
