// mpas_rk4_step_noalloc: the fused RK4 step over a state the caller holds in
// device memory, with every intermediate of the step (provisional states,
// RK4 accumulators, diagnostics, FCT fluxes and limiter factors) in one
// fixed device array of this file.  The call allocates nothing, so callers
// that may not allocate (code under HIP graph capture, serving loops) can
// use it.  It has no counterpart in the reference.
#include <mutex>

#include "ctx.h"
#include "stage.h"

namespace mpasx {
namespace {

constexpr size_t kScratchDoubles = size_t(MPAS_NOALLOC_SCRATCH_BYTES) / sizeof(double);
constexpr size_t kAlign = 256 / sizeof(double);  // buffers start on 256-byte boundaries

// The scratch array.  In host code its name is a host-side stand-in, not a
// device address: kernels and copies take the address from
// hipGetSymbolAddress (scratch_ptr), one per device.
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (looked up once per device,
// with the device current).
int scratch_ptr(int device, double **out) {
  if (device < 0 || device >= kMaxDevices)
    MPAS_FAIL(MPAS_EINVAL, "mpas_rk4_step_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      MPAS_FAIL(MPAS_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[device] = static_cast<double *>(p);
  }
  *out = g_ptr[device];
  return 0;
}

__global__ __launch_bounds__(256) void k_zero(double *p, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = 0.0;
}

// Carves the step's intermediates out of `base` (nullptr: only counts).
// Returns the doubles used.
size_t carve(const DMesh &M, double *base, SwState &W) {
  const size_t E = size_t(M.nE) * M.nL, C = size_t(M.nC) * M.nL, V = size_t(M.nV) * M.nL;
  const size_t T = C * M.nT;
  size_t off = 0;
  auto take = [&](size_t n) -> double * {
    double *p = base ? base + off : nullptr;
    off += (n + kAlign - 1) / kAlign * kAlign;
    return n ? p : nullptr;
  };
  W.nu = take(E);
  W.pu[0] = take(E);
  W.pu[1] = take(E);
  W.h_edge[0] = take(E);
  W.h_edge[1] = take(E);
  W.pv_edge[0] = take(E);
  W.pv_edge[1] = take(E);
  W.nh = take(C);
  W.ph[0] = take(C);
  W.ph[1] = take(C);
  W.ke[0] = take(C);
  W.ke[1] = take(C);
  W.div[0] = take(C);
  W.div[1] = take(C);
  W.ntr = take(T);
  W.ptr[0] = take(T);
  W.ptr[1] = take(T);
  W.tlo = take(T);
  W.rfac = take(2 * T);
  W.anti = take(E * M.nT);
  W.vort[0] = take(V);
  W.vort[1] = take(V);
  W.pv_vertex = take(V);
  return off;
}

}  // namespace
}  // namespace mpasx

using namespace mpasx;

extern "C" {

int mpas_rk4_step_noalloc(mpas_ctx *ctx, double dt, int nsteps, double *h, double *u,
                          double *tracers, const double *h_s) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!h || !u || !h_s) MPAS_FAIL(MPAS_EINVAL, "mpas_rk4_step_noalloc: h, u and h_s are required");
  const DMesh &M = ctx->dm;
  if (M.nT && !tracers) MPAS_FAIL(MPAS_EINVAL, "mpas_rk4_step_noalloc: tracers is NULL");
  if (nsteps < 0) MPAS_FAIL(MPAS_EINVAL, "mpas_rk4_step_noalloc: nsteps < 0");
  if (!(dt > 0)) MPAS_FAIL(MPAS_EINVAL, "mpas_rk4_step_noalloc: dt must be positive");
  if (M.nT && ctx->params.tracer_scheme == MPAS_TRACER_FCT && (M.mlc || !ctx->aligned_c2c))
    MPAS_FAIL(MPAS_EINVAL, "the fused FCT step needs every level active and cellsOnCell[j] "
                           "across edgesOnCell[j]");
  SwState W;
  const size_t need = carve(M, nullptr, W);
  if (need > kScratchDoubles)
    MPAS_FAIL(MPAS_ESHAPE, "mpas_rk4_step_noalloc: the step needs " +
                               std::to_string(need * sizeof(double)) +
                               " bytes of scratch, above the " +
                               std::to_string(size_t(MPAS_NOALLOC_SCRATCH_BYTES)) +
                               " the library holds per device");
  double *base = nullptr;
  if (int rc = scratch_ptr(ctx->device, &base)) return rc;
  carve(M, base, W);
  W.u = u;
  W.h = h;
  W.tr = M.nT ? tracers : nullptr;
  W.h_s = const_cast<double *>(h_s);
  W.cur = 0;
  // the array keeps the previous call's values: clear the part this call
  // uses (a kernel, not hipMemsetAsync, so that the call can be captured)
  hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, ctx->stream, base, need);
  if (int rc = after_launch("k_zero")) return rc;
  // apps/sw/run.cc:88-92: diagnostics of the initial state
  if (int rc = diagnostics_dm(M, ctx->stream, ctx->params.visc2 != 0.0, W.u, W.h, W.ke[0],
                              W.div[0], W.vort[0], W.pv_vertex, W.h_edge[0], W.pv_edge[0]))
    return rc;
  return sw_rk4_steps_on(ctx, W, dt, nsteps);
}

}  // extern "C"
