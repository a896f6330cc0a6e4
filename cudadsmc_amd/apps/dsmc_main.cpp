// dsmc_main - command-line driver with the reference's benchmark flow
// (kernels/dsmc.cu:308-470 without the OpenGL path): read settings from the
// command line (-f config.xml -benchmark -run-for T -dt DT ...), derive the
// particle count, create the GPU solver, optionally load data/ares.raw with
// a 5x5x20 collision grid moved by (0,0,-3), load the particles, run
// maxRunTime/dt + 1 steps and write statistics.txt. The stepping wall time
// goes into simulation_state_t::realTime, as the reference's interactive
// loop records it (kernels/dsmc.cu:270); its benchmark path leaves it 0.
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>

#include "dsmc/dsmc_solver.hpp"

int main(int argc, char** argv) {
  try {
    dsmc::settings_t settings;
    settings.readFromCmdLine(argc, argv);
    settings.deriveParticles();
    std::cout << "Expected number of particles per cell: " << settings.part_per_cell << " density is "
              << settings.density << "\n";
    std::unique_ptr<dsmc::DSMCSolver> solver(dsmc::CreateSolver(dsmc::SolverGPU));
    if (!solver) {
      std::cerr << "Device initialization failed: " << dsmc_global_last_error() << "\n";
      return 1;
    }
    solver->Init(settings);
    if (settings.geometry) {
      const char* raw = std::getenv("DSMC_MODEL") ? std::getenv("DSMC_MODEL") : "data/ares.raw";
      dsmc::Model model;
      if (!model.LoadRaw(raw, !settings.benchmark)) throw std::runtime_error("Unable to load model");
      model.CreateGrid(dsmc::vec3i_t(5, 5, 20));
      model.MoveModel(dsmc::vec3_t(0, 0, -3));
      solver->SetModel(&model);
    }
    dsmc::simulation_state_t sim_state;
    solver->RegisterSimState(&sim_state);
    auto t0 = std::chrono::steady_clock::now();
    uint32_t n = solver->InitSystemState({settings.stream_velocity[0], settings.stream_velocity[1],
                                          settings.stream_velocity[2]});
    solver->Synchronize();
    auto t1 = std::chrono::steady_clock::now();
    std::cout << n << " samples were generated in "
              << std::chrono::duration<double, std::milli>(t1 - t0).count() << " ms\n";
    if (settings.max_run_time > 0.0f) {
      uint32_t iters = (uint32_t)(settings.max_run_time / settings.dt + 1);
      std::cout << "Starting benchmarking. Will run " << iters << " iterations\n";
      t0 = std::chrono::steady_clock::now();
      solver->DoBenchmark(iters, settings.dt);
      solver->Synchronize();
      t1 = std::chrono::steady_clock::now();
      const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
      sim_state.advanceRealTime((float)(ms / 1000.0));
      std::cout << "Benchmark is done in " << ms << " ms\n";
    }
    std::ofstream fout("statistics.txt");
    solver->DumpStatistics(fout);
  } catch (std::exception& e) {
    std::cerr << "Exception: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
