// run.h -- one run of the time loop on the GPU with the reference's NetCDF outputs: the
// body of the reference's shallowWater (src/GPU/ShallowWater.cu:182-500) on the C-ABI of
// libthysea.so.
#pragma once
#include <string>

#include "problem.h"

namespace thyhost {

// Creates the context, uploads the state, runs the loop (Okada, snapshots, gauges,
// metrics, steps) and writes <prefix>.nc and <prefix>_ts.nc.  Returns a thy_status code:
// 0 ok, 1 device memory, 2 host memory, others as include/thysea.h; the message of any
// other failure is printed to stderr.  *seconds: wall time of the time loop.
//
// Environment: THYSEA_PRECISION=1 runs the FP64-everywhere path; THYSEA_AUTOTUNE=0 keeps
// the default strip heights; THYSEA_STEPS_PER_SYNC sets the steps per captured replay;
// THYSEA_QUIET=1 drops the per-iteration lines.
int run_problem(const Problem &p, const std::string &version, double *seconds);

}  // namespace thyhost
