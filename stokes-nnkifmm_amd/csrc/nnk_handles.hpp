// Definitions behind the opaque C handles of include/nnkifmm.h.
#pragma once
#include <memory>

#include "nnk_fmm.hpp"
#include "nnk_tree.hpp"

struct nnk_tree {
  std::shared_ptr<nnk::Tree> t;
};
struct nnk_fmm {
  std::unique_ptr<nnk::Fmm> f;
  int64_t npot, ntgt;
  bool graph = false;  // apply through the captured HIP graph
  void apply(const double *d_f, double *d_u, hipStream_t s) {
    if (graph) f->apply_graph(d_f, d_u, s);
    else f->apply(d_f, d_u, s);
  }
};
