// One launch sequence captured into a HIP graph and replayed with fixed
// staging buffers: run(x, y) copies x into `in`, launches the graph and
// copies `out` to y, all ordered on the caller's stream.  The captured body
// must only launch kernels on the stream it is given (no allocation, host
// synchronisation or memset), so it is run once eagerly before capture by
// the owner to settle lazy allocations.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nnk_common.hpp"

namespace nnk {

struct GraphOp {
  hipStream_t cs = nullptr;
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  DBuf in, out;
  int64_t nin = 0, nout = 0;

  GraphOp() = default;
  GraphOp(const GraphOp &) = delete;
  GraphOp &operator=(const GraphOp &) = delete;
  ~GraphOp() { reset(); }

  bool ready() const { return ge != nullptr; }

  void reset() {
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    if (cs) (void)hipStreamDestroy(cs);
    ge = nullptr;
    g = nullptr;
    cs = nullptr;
  }

  // body(const double *in, double *out, hipStream_t stream)
  template <class F> void capture(int64_t n_in, int64_t n_out, F &&body) {
    reset();
    nin = n_in;
    nout = n_out;
    if (!in.p || in.bytes < sizeof(double) * (size_t)n_in) in.alloc(sizeof(double) * (size_t)n_in);
    if (!out.p || out.bytes < sizeof(double) * (size_t)n_out) out.alloc(sizeof(double) * (size_t)n_out);
    NNK_HIP(hipDeviceSynchronize());
    NNK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    NNK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
      body(in.as<const double>(), out.as<double>(), cs);
    } catch (...) {
      hipGraph_t dead = nullptr;
      (void)hipStreamEndCapture(cs, &dead);
      if (dead) (void)hipGraphDestroy(dead);
      reset();
      throw;
    }
    NNK_HIP(hipStreamEndCapture(cs, &g));
    NNK_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
  }

  void run(const double *x, double *y, hipStream_t s) {
    NNK_HIP(hipMemcpyAsync(in.p, x, sizeof(double) * (size_t)nin, hipMemcpyDeviceToDevice, s));
    NNK_HIP(hipGraphLaunch(ge, s));
    NNK_HIP(hipMemcpyAsync(y, out.p, sizeof(double) * (size_t)nout, hipMemcpyDeviceToDevice, s));
  }
};

}  // namespace nnk
