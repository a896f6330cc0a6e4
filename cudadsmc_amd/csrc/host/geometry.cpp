// Geometry pre-processing on the host: body meshes and the uniform
// collision grid the move kernel walks (the reference's Model/reg_grid_t
// path, src/model.cpp:348-401 and 709-769, src/reggrid.cpp:27-93,
// src/aabb.cpp:38-50).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <thread>

#include "host_internal.h"

namespace dsmc_host {

void Mesh::compute_normals() {
  n.assign(t.size(), 0.0f);
  for (uint32_t i = 0; i < nt(); ++i) {
    const float* a = &v[3 * t[3 * i]];
    const float* b = &v[3 * t[3 * i + 1]];
    const float* c = &v[3 * t[3 * i + 2]];
    float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float x = e1[1] * e2[2] - e1[2] * e2[1];
    float y = e1[2] * e2[0] - e1[0] * e2[2];
    float z = e1[0] * e2[1] - e1[1] * e2[0];
    float l = std::sqrt(x * x + y * y + z * z);
    if (l > 0.0f) {
      x = x / l;
      y = y / l;
      z = z / l;
    }
    n[3 * i] = x;
    n[3 * i + 1] = y;
    n[3 * i + 2] = z;
  }
}

void Mesh::bbox(float mn[3], float mx[3]) const {
  for (int a = 0; a < 3; ++a) {
    mn[a] = 3.0e38f;
    mx[a] = -3.0e38f;
  }
  // over the vertices the triangles use (meshes such as data/ares.raw carry
  // unreferenced vertices at the origin; the bbox stored in the file is the
  // box of the referenced ones)
  for (uint32_t k = 0; k < t.size(); ++k) {
    uint32_t i = t[k];
    for (int a = 0; a < 3; ++a) {
      mn[a] = std::min(mn[a], v[3 * i + a]);
      mx[a] = std::max(mx[a], v[3 * i + a]);
    }
  }
}

void Mesh::translate(float dx, float dy, float dz) {
  for (uint32_t i = 0; i < nv(); ++i) {
    v[3 * i] += dx;
    v[3 * i + 1] += dy;
    v[3 * i + 2] += dz;
  }
}

// Orient every triangle so its normal points away from `inside`.
static void orient_outward(Mesh& m, const double inside[3]) {
  for (uint32_t i = 0; i < m.nt(); ++i) {
    const float* a = &m.v[3 * m.t[3 * i]];
    const float* b = &m.v[3 * m.t[3 * i + 1]];
    const float* c = &m.v[3 * m.t[3 * i + 2]];
    double e1[3], e2[3], cen[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = (double)b[k] - a[k];
      e2[k] = (double)c[k] - a[k];
      cen[k] = ((double)a[k] + b[k] + c[k]) / 3.0 - inside[k];
    }
    double nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2], nz = e1[0] * e2[1] - e1[1] * e2[0];
    if (nx * cen[0] + ny * cen[1] + nz * cen[2] < 0.0) std::swap(m.t[3 * i + 1], m.t[3 * i + 2]);
  }
  m.compute_normals();
}

void mesh_icosphere(Mesh& m, float cx, float cy, float cz, float r, uint32_t subdiv) {
  // icosahedron with unit circumradius, then midpoint subdivision projected
  // onto the sphere (20 * 4^subdiv triangles)
  const double p = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<double> V = {-1, p, 0, 1, p, 0, -1, -p, 0, 1, -p, 0, 0, -1, p, 0, 1, p,
                           0, -1, -p, 0, 1, -p, p, 0, -1, p, 0, 1, -p, 0, -1, -p, 0, 1};
  std::vector<uint32_t> T = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                             11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                             3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  auto norm = [&](uint32_t i) {
    double l = std::sqrt(V[3 * i] * V[3 * i] + V[3 * i + 1] * V[3 * i + 1] + V[3 * i + 2] * V[3 * i + 2]);
    V[3 * i] /= l;
    V[3 * i + 1] /= l;
    V[3 * i + 2] /= l;
  };
  for (uint32_t i = 0; i < 12; ++i) norm(i);
  for (uint32_t s = 0; s < subdiv; ++s) {
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
    auto midpoint = [&](uint32_t a, uint32_t b) {
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      uint32_t idx = (uint32_t)(V.size() / 3);
      for (int k = 0; k < 3; ++k) V.push_back(0.5 * (V[3 * a + k] + V[3 * b + k]));
      norm(idx);
      mid[key] = idx;
      return idx;
    };
    std::vector<uint32_t> T2;
    T2.reserve(T.size() * 4);
    for (size_t f = 0; f < T.size(); f += 3) {
      uint32_t a = T[f], b = T[f + 1], c = T[f + 2];
      uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      uint32_t nt[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      T2.insert(T2.end(), nt, nt + 12);
    }
    T.swap(T2);
  }
  m.v.resize(V.size());
  for (size_t i = 0; i < V.size() / 3; ++i) {
    m.v[3 * i] = (float)(cx + r * V[3 * i]);
    m.v[3 * i + 1] = (float)(cy + r * V[3 * i + 1]);
    m.v[3 * i + 2] = (float)(cz + r * V[3 * i + 2]);
  }
  m.t = T;
  double c[3] = {cx, cy, cz};
  orient_outward(m, c);
}

void mesh_sphere_cone(Mesh& m, float x0, float cy, float cz, float rn, float half_angle_deg, float rb,
                      uint32_t n_az, uint32_t n_profile) {
  // meridian profile (x, r): nose arc from the tip to the tangency point at
  // polar angle 90 - theta, straight cone to the base radius, base disk.
  const double th = half_angle_deg * M_PI / 180.0;
  const double xc = x0 + rn;                         // sphere centre
  const double at = M_PI / 2.0 - th;                 // tangency angle from the axis
  const double xt = xc - rn * std::cos(at), rt = rn * std::sin(at);
  const double xb = xt + (rb - rt) / std::tan(th);   // base plane
  const double l_arc = rn * at, l_cone = (rb - rt) / std::sin(th), l_base = rb;
  const double L = l_arc + l_cone + l_base;
  uint32_t na = std::max<uint32_t>(2, (uint32_t)std::lround(n_profile * l_arc / L));
  uint32_t nc = std::max<uint32_t>(2, (uint32_t)std::lround(n_profile * l_cone / L));
  uint32_t nb = std::max<uint32_t>(2, (uint32_t)std::lround(n_profile * l_base / L));
  std::vector<std::pair<double, double>> prof;  // excludes the two poles
  for (uint32_t k = 1; k <= na; ++k) {
    double a = at * k / na;
    prof.push_back({xc - rn * std::cos(a), rn * std::sin(a)});
  }
  for (uint32_t k = 1; k <= nc; ++k) {
    double f = (double)k / nc;
    prof.push_back({xt + f * (xb - xt), rt + f * (rb - rt)});
  }
  for (uint32_t k = 1; k < nb; ++k) {
    double f = (double)k / nb;
    prof.push_back({xb, rb * (1.0 - f)});
  }
  std::vector<double> V;
  auto add = [&](double x, double y, double z) {
    V.push_back(x);
    V.push_back(y);
    V.push_back(z);
    return (uint32_t)(V.size() / 3 - 1);
  };
  uint32_t tip = add(x0, cy, cz);
  std::vector<uint32_t> ring_start;
  for (auto& pr : prof) {
    ring_start.push_back((uint32_t)(V.size() / 3));
    for (uint32_t j = 0; j < n_az; ++j) {
      double ph = 2.0 * M_PI * j / n_az;
      add(pr.first, cy + pr.second * std::cos(ph), cz + pr.second * std::sin(ph));
    }
  }
  uint32_t back = add(xb, cy, cz);
  std::vector<uint32_t> T;
  for (uint32_t j = 0; j < n_az; ++j) {
    uint32_t j1 = (j + 1) % n_az;
    T.insert(T.end(), {tip, ring_start[0] + j, ring_start[0] + j1});
  }
  for (size_t r = 0; r + 1 < ring_start.size(); ++r) {
    for (uint32_t j = 0; j < n_az; ++j) {
      uint32_t j1 = (j + 1) % n_az;
      uint32_t a = ring_start[r] + j, b = ring_start[r] + j1, c = ring_start[r + 1] + j, d = ring_start[r + 1] + j1;
      T.insert(T.end(), {a, c, b, b, c, d});
    }
  }
  uint32_t last = ring_start.back();
  for (uint32_t j = 0; j < n_az; ++j) {
    uint32_t j1 = (j + 1) % n_az;
    T.insert(T.end(), {last + j, back, last + j1});
  }
  m.v.resize(V.size());
  for (size_t i = 0; i < V.size(); ++i) m.v[i] = (float)V[i];
  m.t = T;
  double inside[3] = {0.5 * (x0 + xb), cy, cz};
  orient_outward(m, inside);
}

bool mesh_load_raw(Mesh& m, const std::string& path, std::string& err) {
  std::ifstream in(path, std::ios::binary);
  if (!in) {
    err = "cannot open " + path;
    return false;
  }
  char mg[2];
  in.read(mg, 2);
  if (!in || mg[0] != 'R' || mg[1] != 'W') {
    err = path + ": not a raw model (magic)";
    return false;
  }
  int32_t ver[2];
  float bb[6];
  in.read((char*)ver, sizeof(ver));
  in.read((char*)bb, sizeof(bb));
  auto read_vec = [&](std::vector<char>& out, size_t elsize) -> bool {
    uint32_t n = 0;
    in.read((char*)&n, 4);
    if (!in) return false;
    out.resize((size_t)n * elsize);
    if (n) in.read(out.data(), (std::streamsize)out.size());
    return (bool)in;
  };
  std::vector<char> verts, tex, norms, fnorm, cols, idx;
  if (!read_vec(verts, 12) || !read_vec(tex, 8) || !read_vec(norms, 12) || !read_vec(fnorm, 12) ||
      !read_vec(cols, 12) || !read_vec(idx, 4)) {
    err = path + ": truncated mesh section";
    return false;
  }
  m.v.resize(verts.size() / 4);
  std::memcpy(m.v.data(), verts.data(), verts.size());
  m.t.resize(idx.size() / 4);
  std::memcpy(m.t.data(), idx.data(), idx.size());
  if (m.t.size() % 3 != 0) {
    err = path + ": index count not a multiple of 3";
    return false;
  }
  for (uint32_t i : m.t)
    if (i >= m.nv()) {
      err = path + ": index out of range";
      return false;
    }
  if (fnorm.size() / 12 == m.nt()) {
    m.n.resize(fnorm.size() / 4);
    std::memcpy(m.n.data(), fnorm.data(), fnorm.size());
  } else {
    m.compute_normals();
  }
  return true;
}

// ------------------------------------------------------------------- SAT
// Triangle/box overlap by the separating axis theorem (T. Akenine-Moller,
// "Fast 3D Triangle-Box Overlap Testing", JGT 2001). Axes: the nine cross
// products of the box axes with the triangle edges (for each the two
// vertices with distinct projections are tested), the three box normals,
// and the triangle plane (vertex of the box farthest along the normal).
static inline bool edge_axis(float pa, float pb, float rad) {
  float lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
  return !(lo > rad || hi < -rad);
}

bool tri_box_overlap(const float center[3], const float half[3], const float a[3], const float b[3],
                     const float c[3]) {
  float v[3][3];
  for (int k = 0; k < 3; ++k) {
    v[0][k] = a[k] - center[k];
    v[1][k] = b[k] - center[k];
    v[2][k] = c[k] - center[k];
  }
  float e[3][3];
  for (int k = 0; k < 3; ++k) {
    e[0][k] = v[1][k] - v[0][k];
    e[1][k] = v[2][k] - v[1][k];
    e[2][k] = v[0][k] - v[2][k];
  }
  // edge i runs from vertex i to vertex (i+1)%3; its opposite vertex is (i+2)%3
  for (int i = 0; i < 3; ++i) {
    const float* p = v[i];
    const float* q = v[(i + 2) % 3];
    float ex = e[i][0], ey = e[i][1], ez = e[i][2];
    float fx = std::fabs(ex), fy = std::fabs(ey), fz = std::fabs(ez);
    // x-axis cross edge: (0, -ez, ey)
    if (!edge_axis(ez * p[1] - ey * p[2], ez * q[1] - ey * q[2], fz * half[1] + fy * half[2])) return false;
    // y-axis cross edge: (ez, 0, -ex)
    if (!edge_axis(-ez * p[0] + ex * p[2], -ez * q[0] + ex * q[2], fz * half[0] + fx * half[2])) return false;
    // z-axis cross edge: (-ey, ex, 0)
    if (!edge_axis(ey * p[0] - ex * p[1], ey * q[0] - ex * q[1], fy * half[0] + fx * half[1])) return false;
  }
  // box normals
  for (int k = 0; k < 3; ++k) {
    float lo = std::min(v[0][k], std::min(v[1][k], v[2][k]));
    float hi = std::max(v[0][k], std::max(v[1][k], v[2][k]));
    if (lo > half[k] || hi < -half[k]) return false;
  }
  // triangle plane against the box
  float nrm[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                  e[0][0] * e[1][1] - e[0][1] * e[1][0]};
  float vmin[3], vmax[3];
  for (int k = 0; k < 3; ++k) {
    if (nrm[k] > 0.0f) {
      vmin[k] = -half[k] - v[0][k];
      vmax[k] = half[k] - v[0][k];
    } else {
      vmin[k] = half[k] - v[0][k];
      vmax[k] = -half[k] - v[0][k];
    }
  }
  if (nrm[0] * vmin[0] + nrm[1] * vmin[1] + nrm[2] * vmin[2] > 0.0f) return false;
  if (nrm[0] * vmax[0] + nrm[1] * vmax[1] + nrm[2] * vmax[2] >= 0.0f) return true;
  return false;
}

// Fill lists for a grid whose bounds and dims are set in g.
static void fill_lists(const Mesh& m, CGrid& g, float inflate) {
  const uint32_t N = g.ncells();
  std::vector<std::vector<uint32_t>> lists(N);
  for (uint32_t t = 0; t < m.nt(); ++t) {
    const float* a = &m.v[3 * m.t[3 * t]];
    const float* b = &m.v[3 * m.t[3 * t + 1]];
    const float* c = &m.v[3 * m.t[3 * t + 2]];
    int lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      float mn = std::min(a[k], std::min(b[k], c[k])), mx = std::max(a[k], std::max(b[k], c[k]));
      lo[k] = (int)std::floor((mn - g.mn[k]) / g.cell[k]) - 1;
      hi[k] = (int)std::floor((mx - g.mn[k]) / g.cell[k]) + 1;
      lo[k] = std::max(lo[k], 0);
      hi[k] = std::min(hi[k], (int)g.dim[k] - 1);
    }
    for (int z = lo[2]; z <= hi[2]; ++z)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          // reg_grid_t::getAABB + AABB_t::from_bbox (src/reggrid.cpp:81-88, include/aabb.h:72-80)
          float bmin[3] = {g.mn[0] + (float)x * g.cell[0], g.mn[1] + (float)y * g.cell[1],
                           g.mn[2] + (float)z * g.cell[2]};
          float half[3], cen[3];
          for (int k = 0; k < 3; ++k) {
            float bmax = bmin[k] + g.cell[k];
            half[k] = (bmax - bmin[k]) / 2.0f;
            cen[k] = bmin[k] + half[k];
            half[k] = half[k] * (1.0f + inflate) + (inflate > 0.0f ? 1e-9f : 0.0f);
          }
          if (tri_box_overlap(cen, half, a, b, c)) lists[(uint32_t)z * g.dim[0] * g.dim[1] + (uint32_t)y * g.dim[0] + (uint32_t)x].push_back(t);
        }
  }
  g.flag.assign(N, 0u);
  g.start.assign(N, 0xFFFFFFFFu);
  g.end.assign(N, 0xFFFFFFFFu);
  g.tri.clear();
  for (uint32_t i = 0; i < N; ++i) {
    if (lists[i].empty()) continue;
    g.flag[i] = 1u;
    g.start[i] = (uint32_t)g.tri.size();
    g.tri.insert(g.tri.end(), lists[i].begin(), lists[i].end());
    g.end[i] = (uint32_t)g.tri.size();
  }
}

void build_cgrid_reference(const Mesh& m, uint32_t gx, uint32_t gy, uint32_t gz, CGrid& g) {
  m.bbox(g.mn, g.mx);
  g.dim[0] = gx;
  g.dim[1] = gy;
  g.dim[2] = gz;
  // reg_grid_t::reset (src/reggrid.cpp:27-47)
  for (int k = 0; k < 3; ++k) g.cell[k] = (g.mx[k] - g.mn[k]) / (float)g.dim[k];
  fill_lists(m, g, 0.0f);
  // reg_grid_t::encloseGrid (src/reggrid.cpp:49-79): along each z column,
  // flag the cells from one flagged cell up to the next flagged cell, in pairs
  for (uint32_t x = 0; x < gx; ++x)
    for (uint32_t y = 0; y < gy; ++y) {
      bool open = false;
      uint32_t first = 0;
      for (uint32_t z = 0; z < gz; ++z) {
        uint32_t i = z * gx * gy + y * gx + x;
        if (g.flag[i] == 0u) continue;
        if (!open) {
          first = z;
          open = true;
        } else {
          for (uint32_t zz = first; zz < z; ++zz) {
            uint32_t j = zz * gx * gy + y * gx + x;
            if (g.flag[j] == 0u) g.flag[j] = 2u;
          }
          open = false;
        }
      }
    }
}

void build_cgrid_physics(const Mesh& m, uint32_t gx, uint32_t gy, uint32_t gz, CGrid& g) {
  float mn[3], mx[3];
  m.bbox(mn, mx);
  uint32_t d[3] = {gx, gy, gz};
  for (int k = 0; k < 3; ++k) {
    float c = (mx[k] - mn[k]) / (float)d[k];
    if (!(c > 0.0f)) c = 1e-6f;
    g.cell[k] = c;
    g.mn[k] = mn[k] - c;
    g.dim[k] = d[k] + 2u;
    g.mx[k] = g.mn[k] + c * (float)g.dim[k];
  }
  fill_lists(m, g, 1e-4f);
  // flood fill of empty cells from the outer layer (6-connected)
  const uint32_t X = g.dim[0], Y = g.dim[1], Z = g.dim[2], N = g.ncells();
  std::vector<uint8_t> outside(N, 0);
  std::vector<uint32_t> stack;
  for (uint32_t z = 0; z < Z; ++z)
    for (uint32_t y = 0; y < Y; ++y)
      for (uint32_t x = 0; x < X; ++x)
        if (x == 0 || y == 0 || z == 0 || x == X - 1 || y == Y - 1 || z == Z - 1) {
          uint32_t i = z * X * Y + y * X + x;
          if (g.flag[i] == 0u && !outside[i]) {
            outside[i] = 1;
            stack.push_back(i);
          }
        }
  while (!stack.empty()) {
    uint32_t i = stack.back();
    stack.pop_back();
    uint32_t z = i / (X * Y), r = i - z * X * Y, y = r / X, x = r - y * X;
    const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (auto& o : nb) {
      int xx = (int)x + o[0], yy = (int)y + o[1], zz = (int)z + o[2];
      if (xx < 0 || yy < 0 || zz < 0 || xx >= (int)X || yy >= (int)Y || zz >= (int)Z) continue;
      uint32_t j = (uint32_t)zz * X * Y + (uint32_t)yy * X + (uint32_t)xx;
      if (g.flag[j] == 0u && !outside[j]) {
        outside[j] = 1;
        stack.push_back(j);
      }
    }
  }
  for (uint32_t i = 0; i < N; ++i)
    if (g.flag[i] == 0u && !outside[i]) g.flag[i] = 2u;
}

bool point_inside(const Mesh& m, const CGrid& g, float x, float y, float z) {
  if (x < g.mn[0] || y < g.mn[1] || z < g.mn[2] || x > g.mx[0] || y > g.mx[1] || z > g.mx[2]) return false;
  uint32_t ix = std::min((uint32_t)((x - g.mn[0]) / g.cell[0]), g.dim[0] - 1);
  uint32_t iy = std::min((uint32_t)((y - g.mn[1]) / g.cell[1]), g.dim[1] - 1);
  uint32_t iz = std::min((uint32_t)((z - g.mn[2]) / g.cell[2]), g.dim[2] - 1);
  uint32_t c = iz * g.dim[0] * g.dim[1] + iy * g.dim[0] + ix;
  if (g.flag[c] == 2u) return true;
  if (g.flag[c] != 1u) return false;
  float best = 3.0e38f, sd = 1.0f;
  for (uint32_t k = g.start[c]; k < g.end[c]; ++k) {
    uint32_t t = g.tri[k];
    const float* a = &m.v[3 * m.t[3 * t]];
    const float* b = &m.v[3 * m.t[3 * t + 1]];
    const float* cc = &m.v[3 * m.t[3 * t + 2]];
    float cx = (a[0] + b[0] + cc[0]) / 3.0f, cy = (a[1] + b[1] + cc[1]) / 3.0f, cz = (a[2] + b[2] + cc[2]) / 3.0f;
    float rx = x - cx, ry = y - cy, rz = z - cz;
    float d2 = rx * rx + ry * ry + rz * rz;
    if (d2 < best) {
      best = d2;
      sd = rx * m.n[3 * t] + ry * m.n[3 * t + 1] + rz * m.n[3 * t + 2];
    }
  }
  return sd < 0.0f;
}

// Near-body code of every flow cell: 255 when no occupied collision-grid
// cell lies within the cell's box widened by one cell per axis (a particle
// that moves less than a cell along every axis cannot reach the body), else
// the Euclidean distance from the cell's box to the nearest occupied
// collision-grid cell, in units of (smallest cell size) / 64, rounded down
// and capped at 254 (so the code never overstates the distance).
void near_cells(const CGrid& g, const float gmin[3], const float cell[3], const uint32_t dim[3],
                std::vector<uint8_t>& out) {
  const uint64_t C = (uint64_t)dim[0] * dim[1] * dim[2];
  out.assign(C, 255);
  const float unit = std::min(cell[0], std::min(cell[1], cell[2])) / 64.0f;
  // collision-grid index range covered by sim cells [i - 1, i + 2) per axis
  auto range = [&](int k, uint32_t i, int& lo, int& hi) {
    double a = (double)gmin[k] + ((double)i - 1.0) * cell[k];
    double b = (double)gmin[k] + ((double)i + 2.0) * cell[k];
    lo = (int)std::floor((a - g.mn[k]) / g.cell[k]) - 1;
    hi = (int)std::floor((b - g.mn[k]) / g.cell[k]) + 1;
    lo = std::max(lo, 0);
    hi = std::min(hi, (int)g.dim[k] - 1);
  };
  // gap between [a0, a1] and [b0, b1] along one axis (0 when they overlap)
  auto gap = [](double a0, double a1, double b0, double b1) { return std::max(0.0, std::max(b0 - a1, a0 - b1)); };
  std::vector<uint8_t> occ(g.ncells());
  for (uint32_t c = 0; c < g.ncells(); ++c) occ[c] = g.flag[c] != 0u;
  for (uint32_t z = 0; z < dim[2]; ++z) {
    int z0, z1;
    range(2, z, z0, z1);
    if (z0 > z1) continue;
    for (uint32_t y = 0; y < dim[1]; ++y) {
      int y0, y1;
      range(1, y, y0, y1);
      if (y0 > y1) continue;
      for (uint32_t x = 0; x < dim[0]; ++x) {
        int x0, x1;
        range(0, x, x0, x1);
        if (x0 > x1) continue;
        const double bx0 = (double)gmin[0] + (double)x * cell[0], bx1 = bx0 + cell[0];
        const double by0 = (double)gmin[1] + (double)y * cell[1], by1 = by0 + cell[1];
        const double bz0 = (double)gmin[2] + (double)z * cell[2], bz1 = bz0 + cell[2];
        double best = -1.0;
        for (int cz = z0; cz <= z1; ++cz) {
          const double qz0 = (double)g.mn[2] + (double)cz * g.cell[2];
          const double gz = gap(bz0, bz1, qz0, qz0 + g.cell[2]);
          for (int cy = y0; cy <= y1; ++cy) {
            const double qy0 = (double)g.mn[1] + (double)cy * g.cell[1];
            const double gy = gap(by0, by1, qy0, qy0 + g.cell[1]);
            for (int cx = x0; cx <= x1; ++cx) {
              if (!occ[(uint64_t)cz * g.dim[0] * g.dim[1] + (uint64_t)cy * g.dim[0] + cx]) continue;
              const double qx0 = (double)g.mn[0] + (double)cx * g.cell[0];
              const double gx = gap(bx0, bx1, qx0, qx0 + g.cell[0]);
              const double d = std::sqrt(gx * gx + gy * gy + gz * gz);
              if (best < 0.0 || d < best) best = d;
            }
          }
        }
        if (best >= 0.0)
          out[(uint64_t)z * dim[0] * dim[1] + (uint64_t)y * dim[0] + x] =
              (uint8_t)std::min(254.0, std::floor(best / (double)unit));
      }
    }
  }
}

void fluid_fractions(const Mesh& m, const CGrid& g, const float gmin[3], const float cell[3], const uint32_t dim[3],
                     std::vector<float>& frac) {
  const uint64_t C = (uint64_t)dim[0] * dim[1] * dim[2];
  frac.assign(C, 1.0f);
  int lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = std::max(0, (int)std::floor((g.mn[k] - gmin[k]) / cell[k]) - 1);
    hi[k] = std::min((int)dim[k] - 1, (int)std::floor((g.mx[k] - gmin[k]) / cell[k]) + 1);
  }
  if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return;
  const int S = 4;
  auto work = [&](int z0, int z1) {
    for (int z = z0; z < z1; ++z)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          int outside = 0;
          for (int a = 0; a < S; ++a)
            for (int b = 0; b < S; ++b)
              for (int c = 0; c < S; ++c) {
                float px = gmin[0] + ((float)x + (a + 0.5f) / S) * cell[0];
                float py = gmin[1] + ((float)y + (b + 0.5f) / S) * cell[1];
                float pz = gmin[2] + ((float)z + (c + 0.5f) / S) * cell[2];
                if (!point_inside(m, g, px, py, pz)) ++outside;
              }
          frac[(uint64_t)z * dim[0] * dim[1] + (uint64_t)y * dim[0] + x] = (float)outside / (float)(S * S * S);
        }
  };
  int nz = hi[2] - lo[2] + 1;
  int nth = (int)std::min<unsigned>(8u, std::max(1u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int t = 0; t < nth; ++t) {
    int z0 = lo[2] + nz * t / nth, z1 = lo[2] + nz * (t + 1) / nth;
    th.emplace_back(work, z0, z1);
  }
  for (auto& x : th) x.join();
}

}  // namespace dsmc_host
