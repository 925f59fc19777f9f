// The shallow-water driver of apps/sw/run.cc over the C++ task layer:
// the fields of apps/sw/state.hh:74-87 and advance_timestep's 23 task calls.
#pragma once
#include "tasks.hh"

namespace mpas { namespace hip { namespace sw {

// apps/sw/state.hh:74 time_versions {prev, curr = prev, next, provis}
enum time_version { prev = 0, curr = prev, next = 1, provis = 2, num_timeversions = 3 };

struct fields {
  field h[num_timeversions], u[num_timeversions], h_edge[num_timeversions],
      h_vertex[num_timeversions], ke[num_timeversions], pv_edge[num_timeversions],
      tracers[num_timeversions];
  field circulation, vorticity, pv_vertex, tend_h, tend_u, tend_tracers, bottomDepth;
  explicit fields(mesh &m);
};

// apps/sw/run.cc:13-70
void advance_timestep(mesh &m, fields &f, double dt);
// apps/sw/run.cc:88-106 without output: initial diagnostics, then nsteps.
void run(mesh &m, fields &f, double dt, int nsteps);

}}}  // namespace mpas::hip::sw
