// Internals shared by the C-ABI translation units (abi.cpp, server.cpp).
#pragma once

#include <cstddef>
#include <memory>

#include "mts.h"

namespace mts {

// Sets the calling thread's mts_last_error() message and returns code.
int set_error(int code, const char *msg);

// The host half of mts_instance_create: the MIB tables, stage records and
// node classification, built without touching the device, with the exact
// device bytes the instance will hold.
struct InstancePlan;
struct PlanDeleter {
  void operator()(InstancePlan *p) const;
};
using PlanPtr = std::unique_ptr<InstancePlan, PlanDeleter>;

// Line splits of an instance: the lane-group lower bound of every line's
// split and the lanes per line of the general tiles (16, 32 or 64; 0: not
// chosen).
struct Lanes {
  int min_lanes = 0, gen_lanes = 0;
};
// lanes: the split to use (mts_tune_lanes' candidates, or the choice kept in a
// context); null: 16 / 16, or the MTS_LINE_LANES / MTS_GEN_LANES overrides.
int plan_instance(const mts_setup *setup, const mts_physics_desc *physics, int method, double dt,
                  int precision, PlanPtr &out, const Lanes *lanes = nullptr);
// The split mts_tune_lanes chose in ctx for this grid, method and precision
// (false: none).
bool ctx_tuned_lanes(mts_ctx *ctx, const int n[3], int method, int precision, Lanes &out);
size_t plan_device_bytes(const InstancePlan &p);
// The device half: allocates, uploads and returns the instance.
int instance_from_plan(mts_ctx *ctx, PlanPtr plan, mts_instance **out);

}  // namespace mts
