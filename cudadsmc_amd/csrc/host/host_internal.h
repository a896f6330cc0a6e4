// Host-side helpers shared by the solver, settings and geometry code.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace dsmc_host {

// Error string of calls that have no solver handle (thread-local).
int set_global_error(int code, const std::string& msg);
const char* global_error();

// Triangle mesh: 3 floats per vertex, 3 indices and a unit normal per triangle.
struct Mesh {
  std::vector<float> v;
  std::vector<uint32_t> t;
  std::vector<float> n;
  uint32_t nv() const { return (uint32_t)(v.size() / 3); }
  uint32_t nt() const { return (uint32_t)(t.size() / 3); }
  void compute_normals();  // n = normalize((b - a) x (c - a)), FP32
  void bbox(float mn[3], float mx[3]) const;
  void translate(float dx, float dy, float dz);
};

void mesh_icosphere(Mesh& m, float cx, float cy, float cz, float r, uint32_t subdiv);
// Sphere-capped cone along +x, nose tip at (x0, cy, cz): nose radius rn,
// half-angle (degrees), base radius rb, closed by a flat base disk.
void mesh_sphere_cone(Mesh& m, float x0, float cy, float cz, float rn, float half_angle_deg, float rb,
                      uint32_t n_az, uint32_t n_profile);
// Model::LoadRaw (src/model.cpp:348-401): vertices, face normals and indices.
bool mesh_load_raw(Mesh& m, const std::string& path, std::string& err);

// Uniform collision grid over a box (reg_grid_t, include/reggrid.h:32-56).
struct CGrid {
  float mn[3], mx[3], cell[3];
  uint32_t dim[3];
  std::vector<uint32_t> flag;   // 0 empty, 1 triangles, 2 filled interior
  std::vector<uint32_t> start;  // first entry or 0xFFFFFFFF
  std::vector<uint32_t> end;
  std::vector<uint32_t> tri;    // triangle indices, ascending per cell
  uint32_t ncells() const { return dim[0] * dim[1] * dim[2]; }
};

// REFERENCE: grid over the mesh bounding box, SAT per (cell, triangle),
// encloseGrid z-fill (src/model.cpp:733-769, src/reggrid.cpp:27-93).
void build_cgrid_reference(const Mesh& m, uint32_t gx, uint32_t gy, uint32_t gz, CGrid& g);
// PHYSICS: grid padded by one cell around the mesh, conservative SAT, and
// interior cells found by flood fill from the outside.
void build_cgrid_physics(const Mesh& m, uint32_t gx, uint32_t gy, uint32_t gz, CGrid& g);

// Akenine-Moller triangle/box overlap test (13 separating axes).
bool tri_box_overlap(const float center[3], const float half[3], const float a[3], const float b[3],
                     const float c[3]);

// Point inside the body (flag 2, or behind the nearest-centroid triangle of
// a flag-1 cell).
bool point_inside(const Mesh& m, const CGrid& g, float x, float y, float z);

// Fluid volume fraction of each simulation cell (4x4x4 stratified points).
// PHYSICS: 1 for every simulation cell whose box, widened by one cell on
// each side, meets a collision-grid cell holding triangles or interior. A
// particle in an unflagged cell that moves less than one cell per axis in a
// step cannot reach a triangle.
void near_cells(const CGrid& g, const float gmin[3], const float cell[3], const uint32_t dim[3],
                std::vector<uint8_t>& out);
void fluid_fractions(const Mesh& m, const CGrid& g, const float gmin[3], const float cell[3], const uint32_t dim[3],
                     std::vector<float>& frac);

// MT parameter rows: built-in table (1024 rows) or a 16-byte-record file.
const uint32_t* builtin_mt_table(uint32_t* rows);
bool load_mt_rows(const std::string& path, std::vector<uint32_t>& a, std::vector<uint32_t>& b,
                  std::vector<uint32_t>& c, std::string& err);

}  // namespace dsmc_host
