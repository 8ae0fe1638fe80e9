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

int plan_instance(const mts_setup *setup, const mts_physics_desc *physics, int method, double dt,
                  int precision, PlanPtr &out);
size_t plan_device_bytes(const InstancePlan &p);
// The device half: allocates, uploads and returns the instance.
int instance_from_plan(mts_ctx *ctx, PlanPtr plan, mts_instance **out);

}  // namespace mts
