// dsmc_solver.hpp - C++ mirror of the reference's solver interface over the
// C ABI (dsmc_c.h). dsmc::DSMCSolver keeps the public and protected members
// of the reference's class (include/dsmc_base.h:40-80) with their
// signatures: Init, DoBenchmark, RunSimulationStep, InitSystemState,
// RegisterSimState, DumpStatistics(std::ostream&), SetModel,
// RunBenchmarkSimulationStep, SampleStatistics(float4&) returning a new[]
// array the caller delete[]s, plus simulation_state_t
// (include/draw_helper.h:40-78), statistics_t (include/statistics.h:31-37),
// Model's LoadRaw/CreateGrid/MoveModel (include/model.h:46-82) and
// dsmc::CreateSolver (kernels/dsmc_base.cu:221-233). The rendering members
// are not provided. settings_t keeps readFromFile/readFromCmdLine/
// getCellsCount, but its fields are those of dsmc_settings (plain arrays,
// e.g. grid_dim[0] for the reference's grid_dim.x). Errors that the
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

// Per-cell sums widened to double (dsmc_get_cell_stats).
using cell_stats_t = dsmc_cell_stats;

struct float2 {
  float x, y;
};
struct float3 {
  float x, y, z;
};
struct float4 {
  float x, y, z, w;
};

// statistics_t (include/statistics.h:31-37): FP32 per-cell sums, as
// SampleStatistics returns them.
struct statistics_t {
  float numInSample;
  float3 xyzSum;
  float velMag;
  float pad_;
  float2 padd;
};
static_assert(sizeof(statistics_t) == 8 * sizeof(float), "statistics_t is 8 floats (dsmc_sample_statistics)");

// simulation_state_t (include/draw_helper.h:40-78) without the interactive
// rendering toggles' side effects.
struct simulation_state_t {
  bool paused = true, drawGrid = false, drawVectorField = false, demoMode = false, processCollisions = true,
       sampleConcentration = false;
  float simTime = 0.0f, realTime = 0.0f;
  uint32_t slice = 0;
  float fps = 0.0f;
  uint32_t stepCount = 0, samplesCount = 0;
  void advanceSimTime(float dt) { simTime += dt; }
  void advanceRealTime(float dt) { realTime += dt; }
  void advanceSimStep() { stepCount++; }
  void advanceSamplesCount() { samplesCount++; }
};

struct vec3_t {
  float x, y, z;
  vec3_t(float x_ = 0, float y_ = 0, float z_ = 0) : x(x_), y(y_), z(z_) {}
};
struct vec3i_t {
  uint32_t x, y, z;
  vec3i_t(uint32_t x_ = 0, uint32_t y_ = 0, uint32_t z_ = 0) : x(x_), y(y_), z(z_) {}
};

// A triangle mesh handed to SetModel (the reference's Model keeps its own
// vertices and grid; here the mesh and the collision-grid size are given).
struct Model {
  std::vector<float> vertices;
  std::vector<uint32_t> triangles;
  std::vector<float> normals;
  uint32_t grid[3] = {5, 5, 20};
  // CreateGrid then MoveModel: the grid is built over the vertices as they
  // were at CreateGrid and shifts with the model (src/model.cpp:709-769)
  std::vector<float> grid_vertices;
  float grid_move[3] = {0.0f, 0.0f, 0.0f};

  // Model::LoadRaw (src/model.cpp:348-401); there are no render buffers to create
  bool LoadRaw(const char* fname, bool createBuffers = true) {
    (void)createBuffers;
    uint32_t nv = 0, nt = 0;
    if (dsmc_mesh_load_raw(fname, nullptr, 0, nullptr, 0, nullptr, &nv, &nt) != DSMC_OK) return false;
    vertices.resize(3ull * nv);
    triangles.resize(3ull * nt);
    normals.resize(3ull * nt);
    return dsmc_mesh_load_raw(fname, vertices.data(), nv, triangles.data(), nt, normals.data(), &nv, &nt) == DSMC_OK;
  }
  // Model::CreateGrid (src/model.cpp:733-769): the grid is built by SetModel
  void CreateGrid(const vec3i_t& dim) {
    grid[0] = dim.x;
    grid[1] = dim.y;
    grid[2] = dim.z;
    grid_vertices = vertices;
    grid_move[0] = grid_move[1] = grid_move[2] = 0.0f;
  }
  void MoveModel(const vec3_t& t) {
    for (size_t i = 0; i < vertices.size(); i += 3) {
      vertices[i] += t.x;
      vertices[i + 1] += t.y;
      vertices[i + 2] += t.z;
    }
    grid_move[0] += t.x;
    grid_move[1] += t.y;
    grid_move[2] += t.z;
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
  // DoBenchmark (kernels/dsmc_base.cu:64-84): the steps are enqueued as one
  // batch (HIP graph replay), without the per-step progress printout
  void DoBenchmark(uint32_t itn, float dt) {
    throw_on(dsmc_run(h_, itn, dt), h_);
    sync_state();
  }
  virtual void RunSimulationStep(float dt) {
    throw_on(dsmc_step(h_, dt), h_);
    sync_state();
  }
  virtual uint32_t InitSystemState(const float3& streamVel) {
    uint32_t n = 0;
    throw_on(dsmc_init_state(h_, streamVel.x, streamVel.y, streamVel.z, &n), h_);
    return n;
  }
  // The solver advances the registered state's stepCount, samplesCount and
  // simTime as the reference's step does (include/gpu_dsmc.inl:221-274);
  // realTime stays the caller's.
  void RegisterSimState(simulation_state_t* state) {
    m_state = state;
    sync_state();
  }
  // SampleStatistics + ProcessAndOutputStatistics (kernels/dsmc_base.cu:165-174)
  void DumpStatistics(std::ostream& out) {
    float4 collStat;
    statistics_t* stat = SampleStatistics(collStat);
    delete[] stat;
    dsmc_sim_state st{};
    throw_on(dsmc_get_sim_state(h_, &st), h_);
    if (m_state) {
      st.sim_time = m_state->simTime;
      st.real_time = m_state->realTime;
      st.samples_count = m_state->samplesCount;
    }
    uint64_t len = 0;
    throw_on(dsmc_format_statistics(h_, &st, nullptr, 0, &len), h_);
    std::string text(len + 1, '\0');
    throw_on(dsmc_format_statistics(h_, &st, &text[0], len + 1, &len), h_);
    text.resize(len);
    out << text;
  }
  // the same table written to a file by the engine (statistics.txt)
  void DumpStatistics(const char* path) { throw_on(dsmc_dump_statistics(h_, path), h_); }
  void SetModel(Model* model) {
    const float* nrm = model->normals.empty() ? nullptr : model->normals.data();
    const uint32_t nt = (uint32_t)(model->triangles.size() / 3);
    if (model->grid_vertices.size() == model->vertices.size())  // built by CreateGrid, maybe moved since
      throw_on(dsmc_set_geometry_moved(h_, model->grid_vertices.data(), (uint32_t)(model->grid_vertices.size() / 3),
                                       model->triangles.data(), nt, nrm, model->grid[0], model->grid[1],
                                       model->grid[2], model->grid_move),
               h_);
    else
      throw_on(dsmc_set_geometry(h_, model->vertices.data(), (uint32_t)(model->vertices.size() / 3),
                                 model->triangles.data(), nt, nrm, model->grid[0], model->grid[1], model->grid[2]),
               h_);
  }
  // per-cell sums in double without sampling again
  std::vector<cell_stats_t> CellStatistics() {
    std::vector<cell_stats_t> out(m_settings.getCellsCount());
    throw_on(dsmc_get_cell_stats(h_, out.data(), (uint32_t)out.size()), h_);
    return out;
  }
  void Synchronize() { throw_on(dsmc_synchronize(h_), h_); }
  dsmc_solver* handle() { return h_; }

 protected:
  virtual void RunBenchmarkSimulationStep(float dt) { RunSimulationStep(dt); }
  // GPUDSMCSolver::SampleStatistics (include/gpu_dsmc.inl:277-282): collision
  // totals, one more (uncounted) sample, and a new[] array of the per-cell
  // sums that the caller delete[]s
  virtual statistics_t* SampleStatistics(float4& collStat) {
    const uint32_t n = m_settings.getCellsCount();
    statistics_t* stat = new statistics_t[n];
    float coll[4];
    int rc = dsmc_sample_statistics(h_, coll, reinterpret_cast<float*>(stat), n);
    if (rc != DSMC_OK) {
      delete[] stat;
      throw_on(rc, h_);
    }
    collStat = {coll[0], coll[1], coll[2], coll[3]};
    return stat;
  }
  void sync_state() {
    if (!m_state) return;
    dsmc_sim_state st{};
    throw_on(dsmc_get_sim_state(h_, &st), h_);
    m_state->simTime = st.sim_time;
    m_state->stepCount = st.step_count;
    m_state->samplesCount = st.samples_count;
  }

  settings_t m_settings;
  simulation_state_t* m_state = nullptr;
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
