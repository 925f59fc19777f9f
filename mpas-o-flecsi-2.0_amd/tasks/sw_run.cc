// apps/sw/run.cc restated over the C++ task layer (one task call per
// FleCSI execute<>).
#include "sw_run.hh"

#include <array>
#include <cstdint>

namespace mpas { namespace hip { namespace sw {

fields::fields(mesh &m)
    : circulation(m, entity::vertices, m.levels()),
      vorticity(m, entity::vertices, m.levels()),
      pv_vertex(m, entity::vertices, m.levels()),
      tend_h(m, entity::cells, m.levels()),
      tend_u(m, entity::edges, m.levels()),
      tend_tracers(m, entity::cells, m.levels() * m.tracers()),
      bottomDepth(m, entity::cells, 1) {
  for (int v = 0; v < num_timeversions; v++) {
    h[v] = field(m, entity::cells, m.levels());
    u[v] = field(m, entity::edges, m.levels());
    h_edge[v] = field(m, entity::edges, m.levels());
    h_vertex[v] = field(m, entity::vertices, m.levels());
    ke[v] = field(m, entity::cells, m.levels());
    pv_edge[v] = field(m, entity::edges, m.levels());
    tracers[v] = field(m, entity::cells, m.levels() * m.tracers());
  }
}

void advance_timestep(mesh &m, fields &f, double dt) {
  namespace t = task;
  t::init_provis(m, f.u[prev], f.h[prev], f.tracers[prev], f.h_edge[prev], f.ke[prev],
                 f.pv_edge[prev], f.u[provis], f.h[provis], f.tracers[provis], f.h_edge[provis],
                 f.ke[provis], f.pv_edge[provis]);
  t::setup_timestep(m, f.u[prev], f.h[prev], f.tracers[prev], f.u[next], f.h[next],
                    f.tracers[next]);
  const std::array<double, 4> rk_weights{{dt / 6, dt / 3, dt / 3, dt / 6}};
  const std::array<double, 4> rk_substep_weights{{dt / 2, dt / 2, dt, 0}};
  for (std::size_t rkstep = 0; rkstep < rk_weights.size(); rkstep++) {
    t::compute_tend(m, f.h_edge[provis], f.h[provis], f.bottomDepth, f.u[provis], f.ke[provis],
                    f.pv_edge[provis], f.tend_h, f.tend_u);
    t::compute_scalar_tend(m, f.h_edge[provis], f.u[provis], f.tracers[provis], f.tend_tracers);
    if (rkstep < 3) {
      t::compute_substep(m, f.u[provis], f.u[prev], f.tend_u, f.h[provis], f.h[prev], f.tend_h,
                         f.tracers[provis], f.tracers[prev], f.tend_tracers,
                         rk_substep_weights[rkstep]);
      t::compute_solve_diagnostics(m, f.h[provis], f.u[provis], f.h_edge[provis],
                                   f.h_vertex[provis], f.circulation, f.vorticity,
                                   f.ke[provis], f.pv_edge[provis], f.pv_vertex);
    }
    t::accumulate_update(m, f.u[next], f.tend_u, f.h[next], f.tend_h, f.tracers[next],
                         f.tend_tracers, rk_weights[rkstep]);
  }
  t::finalize_timestep(m, f.u[prev], f.u[next], f.h[next], f.tracers[next]);
  t::compute_solve_diagnostics(m, f.h[next], f.u[next], f.h_edge[next], f.h_vertex[next],
                               f.circulation, f.vorticity, f.ke[next], f.pv_edge[next],
                               f.pv_vertex);
  t::timeshift(m, f.u[prev], f.h[prev], f.tracers[prev], f.h_edge[prev], f.ke[prev],
               f.pv_edge[prev], f.u[next], f.h[next], f.tracers[next], f.h_edge[next],
               f.ke[next], f.pv_edge[next]);
}

std::vector<int> output_steps(int nsteps, int num_output_times) {
  std::vector<int> out;
  if (num_output_times < 2 || nsteps < 1) return out;
  const uint64_t n = uint64_t(nsteps), P = uint64_t(num_output_times - 1);
  for (uint64_t i = 1; i <= n; i++) {
    const uint64_t owner = (P * (i + 1) - 1) / n;  // block_partition::owner
    const uint64_t high = (owner + 1) * n / P - 1;  // block_partition::high
    if (high == i) out.push_back(int(i));
  }
  return out;
}

void run(mesh &m, fields &f, double dt, int nsteps, int num_output_times,
         const std::function<void(int)> &output) {
  const bool out = num_output_times > 0 && output;
  if (out) output(0);
  task::compute_solve_diagnostics(m, f.h[curr], f.u[curr], f.h_edge[curr], f.h_vertex[curr],
                                  f.circulation, f.vorticity, f.ke[curr], f.pv_edge[curr],
                                  f.pv_vertex);
  const std::vector<int> at = out ? output_steps(nsteps, num_output_times) : std::vector<int>();
  size_t next_out = 0;
  for (int i = 1; i <= nsteps; i++) {
    advance_timestep(m, f, dt);
    if (next_out < at.size() && at[next_out] == i) output(int(++next_out));
  }
}

}}}  // namespace mpas::hip::sw
