// sgs.h -- sequential Gaussian simulation on the GPU (the reference's
// sequential_simulation.h:51-119 with sequential_gaussian_simulation.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "krige_launch.h"
#include "sgs_plan.h"

namespace hpglx {

// Simulates the schedule's cells level by level on `st`, one launch per
// level and one wave per cell: a template search over the current informed
// mask, the FP64 SK (mean 0) or OK solve with variance over the plan's
// clipped table, and the draw z * sqrt(variance) + estimate written back as
// data.  values and mask (device, N) are updated in place.  The plan must be
// built with the covariance clipped to the radius box (the simple-kriging
// plan).
// d_lvm (device, N; simple kriging only, may be null): the cells' own means
// in the combine (sequential_gaussian_simulation.cpp:80-131).  Several
// realizations (one schedule each) run together, step l taking level l of
// every one: values and mask then hold one slice of N per realization, each
// starting from its conditioning data.
int sgs_run(DevPlan& plan, Workspace& ws, int kind, const int dims[3], float* d_values, unsigned char* d_mask,
            const float* d_lvm, const std::vector<SgsSchedule>& sched, hipStream_t st, std::string& err);

}  // namespace hpglx
