// The per-node work of an HSS sweep or factorisation step: phases (one
// launch each, in order) of tasks (one workgroup each, independent within a
// phase) of ops (small GEMMs or copies, run in order by the task's
// workgroup).  All descriptors go to the device in one staged copy.
#pragma once
#include <utility>
#include <vector>

#include "kernels.hpp"

namespace hmat {

struct SweepProgram {
  std::vector<kern::BGemm> ops;
  std::vector<kern::NodeTask> tasks;
  std::vector<std::pair<int, int>> phases;  // [first task, end)
  void begin_phase() { phases.push_back({(int)tasks.size(), (int)tasks.size()}); }
  void begin_task() { tasks.push_back(kern::NodeTask{(int32_t)ops.size(), 0}); }
  void op(const kern::BGemm& g) {
    ops.push_back(g);
    ++tasks.back().nops;
  }
  // C = alpha A + beta C for an m x n block
  void copy(const double* A, int lda, double* C, int ldc, int m, int n, double alpha = 1.0, double beta = 0.0) {
    op(kern::BGemm{A, nullptr, C, m, n, 0, lda, 0, ldc, 0, 0, alpha, beta});
  }
  void end_task() {
    if (tasks.back().nops == 0) tasks.pop_back();
    phases.back().second = (int)tasks.size();
  }
  void run(hipStream_t s) {
    if (tasks.empty()) return;
    StageGroup g;
    const int io = g.add(ops), it = g.add(tasks);
    g.put(s);
    const kern::BGemm* dops = g.get<kern::BGemm>(io);
    const kern::NodeTask* dt = g.get<kern::NodeTask>(it);
    for (const auto& p : phases) kern::node_tasks(dt + p.first, p.second - p.first, dops, s);
  }
};

}  // namespace hmat
