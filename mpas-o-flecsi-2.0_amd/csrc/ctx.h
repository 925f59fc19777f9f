// Internal state behind the opaque mpas_ctx handle and the error plumbing
// of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "dmesh.h"
#include "mesh_setup.h"
#include "mpas_hip.h"

namespace mpasx {

void set_error(const std::string &msg);

// Fails the calling entry point with `code` and a message.
#define MPAS_FAIL(code, msg)    \
  do {                          \
    ::mpasx::set_error(msg);    \
    return (code);              \
  } while (0)

#define MPAS_HIP(expr)                                                        \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      MPAS_FAIL(MPAS_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// Device-resident state of the fused shallow-water run (apps/sw/state.hh:74-87).
struct SwState {
  bool ready = false;
  double elapsed = 0;
  // prognostic: prev (= curr) and the RK4 "next" accumulators
  double *u = nullptr, *h = nullptr, *tr = nullptr;
  double *nu = nullptr, *nh = nullptr, *ntr = nullptr;
  // provisional state and its diagnostics, ping-pong pair [2]
  double *pu[2] = {}, *ph[2] = {}, *ptr[2] = {};
  double *h_edge[2] = {}, *ke[2] = {}, *pv_edge[2] = {}, *vort[2] = {}, *div[2] = {};
  double *pv_vertex = nullptr, *h_s = nullptr;
  // FCT intermediates
  double *anti = nullptr;   // [E][nL][nT] antidiffusive flux
  double *rfac = nullptr;   // [C][nL][nT][2] limiter factors (in, out)
  double *tlo = nullptr;    // [C][nL][nT] upwind tendency
  unsigned long long *nonfinite = nullptr;  // counter of mpas_sw_state_check
  int cur = 0;              // which ping-pong slot holds the current state
};

// Per-kernel event timing of the fused step (mpas_prof_enable).
struct Prof {
  bool on = false;
  struct Pending {
    int kernel;
    hipEvent_t start, stop;
    double bytes;
    hipStream_t stream;
  };
  std::vector<std::string> names;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  int id(const char *name) {
    for (size_t i = 0; i < names.size(); i++)
      if (names[i] == name) return int(i);
    names.push_back(name);
    return int(names.size()) - 1;
  }
};

struct AdjState;  // tape and buffers of the tangent-linear and adjoint runs

}  // namespace mpasx

struct mpas_host_mesh : mpasx::HostMesh {};

struct mpas_ctx {
  mpas_params params{};
  mpas_host_mesh host;
  mpasx::DMesh dm{};
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::vector<void *> mesh_bufs;  // static mesh arrays
  std::set<void *> user_bufs;     // mpas_buf_alloc
  std::unique_ptr<mpasx::SwState> sw;
  mpasx::AdjState *adj = nullptr;  // mpas_adj_init; freed by adj_free
  bool aligned_c2c = false;       // cellsOnCell[j] is across edgesOnCell[j]
  // list lengths summed over the mesh (edgesOnEdge, edgesOnCell,
  // advCellsForEdge): the per-launch byte counts of the fused step
  double sum_eoe = 0, sum_eoc = 0, sum_adv = 0;
  mpasx::Prof prof;
};

namespace mpasx {

inline int blocks_for(int64_t n, int bs) { return int((n + bs - 1) / bs); }

// Checks the context and the current device; returns 0 or a failure code.
int check_ctx(const mpas_ctx *ctx);

// Kernel launch epilogue: reports launch errors as MPAS_EHIP.
int after_launch(const char *what);

// Fails with MPAS_ENONFINITE, naming `what` and the index, when one of the
// n host values is NaN or Inf.
int check_finite(const double *p, size_t n, const char *what);

// Event bracket of one launch on stream st while profiling is on (no-ops
// otherwise, and for a null context; st = null means the context's stream).
int prof_begin(mpas_ctx *ctx, const char *kernel, double bytes, hipStream_t st = nullptr);
int prof_end(mpas_ctx *ctx, hipStream_t st = nullptr);

// Frees the tangent-linear / adjoint tape of a context (no-op without one).
void adj_free(mpas_ctx *ctx);

// Zeroes freshly allocated device memory and waits for it: the context's,
// the ensemble's and the replay streams are non-blocking, so nothing else
// would order a null-stream fill before their kernels (a fill still queued
// when a kernel on another stream writes the buffer would clobber it).
inline hipError_t zero_now(void *p, size_t bytes) {
  const hipError_t e = hipMemsetAsync(p, 0, bytes, nullptr);
  return e != hipSuccess ? e : hipStreamSynchronize(nullptr);
}

}  // namespace mpasx
