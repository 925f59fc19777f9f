// The shallow-water driver of apps/sw/run.cc over the C++ task layer:
// the fields of apps/sw/state.hh:74-87 and advance_timestep's 23 task calls.
#pragma once
#include <functional>
#include <vector>

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
// The steps i in [1, nsteps] after which apps/sw/run.cc:100 writes an output
// time: output_flagger::output_step (mpasoflecsi/io/state.hh:22-38) over
// block_partition(nsteps, num_output_times - 1) (util/part.hh:28-53), i.e.
// i is written when it is the last index of the block that owns it.  The
// partition covers 0..nsteps-1 while the loop counts 1..nsteps, so the last
// step is not always written.  One output time writes only time 0 (the
// reference divides by zero there).
std::vector<int> output_steps(int nsteps, int num_output_times);

// apps/sw/run.cc:73-113: with output, output time 0 is written first (before
// the initial diagnostics), then time t after the t-th step of
// output_steps(); `output(t)` writes time t.  Without output
// (num_output_times = 0) only the initial diagnostics and nsteps steps run.
void run(mesh &m, fields &f, double dt, int nsteps, int num_output_times = 0,
         const std::function<void(int)> &output = {});

}}}  // namespace mpas::hip::sw
