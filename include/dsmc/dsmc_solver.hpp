// dsmc_solver.hpp - C++ mirror of the reference's solver interface over the
// C ABI (dsmc_c.h). Method names and meaning follow dsmc::DSMCSolver
// (include/dsmc_base.h:40-80), settings_t (include/sim_settings.h:32-69) and
// dsmc::CreateSolver (kernels/dsmc_base.cu:221-233), so code written against
// the reference's GPU solver compiles against this header; errors that the
// reference reported with checkCUDAErrorAndThrow (dsmc_kernel.cu:2066-2072)
// become std::runtime_error here as well.
#ifndef DSMC_SOLVER_HPP_
#define DSMC_SOLVER_HPP_

#include <cstring>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dsmc_c.h"

namespace dsmc {

inline void throw_on(int rc, const dsmc_solver* h = nullptr) {
  if (rc != DSMC_OK)
    throw std::runtime_error(std::string("dsmc: ") + (h ? dsmc_last_error(h) : dsmc_global_last_error()));
}

// settings_t with readFromFile / readFromCmdLine.
struct settings_t : dsmc_settings {
  settings_t() { dsmc_settings_default(this); }
  void readFromFile(const char* filename) { throw_on(dsmc_settings_read_file(this, filename)); }
  void readFromCmdLine(int argc, char** argv) {
    throw_on(dsmc_settings_read_cmdline(this, argc, const_cast<const char* const*>(argv)));
  }
  uint32_t getCellsCount() const { return (uint32_t)grid_dim[0] * grid_dim[1] * grid_dim[2]; }
  // main's particle count (kernels/dsmc.cu:324-330)
  void deriveParticles() { throw_on(dsmc_settings_derive_particles(this)); }
};

// statistics_t (include/statistics.h:31-37), widened to double.
using statistics_t = dsmc_cell_stats;

struct float3 {
  float x, y, z;
};

// A triangle mesh handed to SetModel (the reference's Model keeps its own
// vertices and grid; here the mesh and the collision-grid size are given).
struct Model {
  std::vector<float> vertices;
  std::vector<uint32_t> triangles;
  std::vector<float> normals;
  uint32_t grid[3] = {5, 5, 20};

  bool LoadRaw(const char* fname) {
    uint32_t nv = 0, nt = 0;
    if (dsmc_mesh_load_raw(fname, nullptr, 0, nullptr, 0, nullptr, &nv, &nt) != DSMC_OK) return false;
    vertices.resize(3ull * nv);
    triangles.resize(3ull * nt);
    normals.resize(3ull * nt);
    return dsmc_mesh_load_raw(fname, vertices.data(), nv, triangles.data(), nt, normals.data(), &nv, &nt) == DSMC_OK;
  }
  void CreateGrid(uint32_t gx, uint32_t gy, uint32_t gz) {
    grid[0] = gx;
    grid[1] = gy;
    grid[2] = gz;
  }
  void MoveModel(const float3& t) {
    for (size_t i = 0; i < vertices.size(); i += 3) {
      vertices[i] += t.x;
      vertices[i + 1] += t.y;
      vertices[i + 2] += t.z;
    }
  }
};

enum SolverType { SolverGPU = DSMC_SOLVER_GPU, SolverGPUVBO = DSMC_SOLVER_GPU_VBO, SolverCPU = DSMC_SOLVER_CPU };

class DSMCSolver {
 public:
  explicit DSMCSolver(dsmc_solver* h) : h_(h) {}
  DSMCSolver(const DSMCSolver&) = delete;
  DSMCSolver& operator=(const DSMCSolver&) = delete;
  virtual ~DSMCSolver() { dsmc_destroy(h_); }

  void Init(const settings_t& s) {
    m_settings = s;
    throw_on(dsmc_init(h_, &s), h_);
  }
  void SetModel(Model* model) {
    throw_on(dsmc_set_geometry(h_, model->vertices.data(), (uint32_t)(model->vertices.size() / 3),
                               model->triangles.data(), (uint32_t)(model->triangles.size() / 3),
                               model->normals.empty() ? nullptr : model->normals.data(), model->grid[0],
                               model->grid[1], model->grid[2]),
             h_);
  }
  uint32_t InitSystemState(const float3& streamVel) {
    uint32_t n = 0;
    throw_on(dsmc_init_state(h_, streamVel.x, streamVel.y, streamVel.z, &n), h_);
    return n;
  }
  void RunSimulationStep(float dt) { throw_on(dsmc_step(h_, dt), h_); }
  void DoBenchmark(uint32_t itn, float dt) { throw_on(dsmc_run(h_, itn, dt), h_); }
  void DumpStatistics(const char* path) { throw_on(dsmc_dump_statistics(h_, path), h_); }
  // per-cell sums; collStat = {selections, collisions, sum c_r, 0}
  std::vector<statistics_t> SampleStatistics(double collStat[4]) {
    throw_on(dsmc_get_collision_stats(h_, collStat), h_);
    throw_on(dsmc_sample(h_), h_);
    std::vector<statistics_t> out(m_settings.getCellsCount());
    throw_on(dsmc_get_cell_stats(h_, out.data(), (uint32_t)out.size()), h_);
    return out;
  }
  void Synchronize() { throw_on(dsmc_synchronize(h_), h_); }
  dsmc_solver* handle() { return h_; }

 protected:
  settings_t m_settings;
  dsmc_solver* h_;
};

// CreateSolver returns nullptr for the solver types the engine does not
// provide, as the reference does for SolverCPU.
inline DSMCSolver* CreateSolver(SolverType type) {
  dsmc_solver* h = nullptr;
  if (dsmc_create((int)type, &h) != DSMC_OK) return nullptr;
  return new DSMCSolver(h);
}

}  // namespace dsmc

#endif  // DSMC_SOLVER_HPP_
