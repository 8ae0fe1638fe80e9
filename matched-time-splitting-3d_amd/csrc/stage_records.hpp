// Stage records of the MIB blocks (see device_types.hpp, kRec*).
#pragma once

#include <string>
#include <vector>

#include "device_types.hpp"
#include "host_setup.hpp"
#include "host_table.hpp"

namespace mts {

// Records (kRecLen doubles per block of direction d) for the stage kind
// `flags` (kRhs* | kFlag*; EpiF and Seq do not change them).  theta_dt, dt:
// the instance's; c_f1, c_f2: the source factors of a TS z stage.
std::string build_stage_records(const SetupResult &s, const DeviceTableHost &t, int d, int flags,
                                double theta_dt, double dt, double c_f1, double c_f2,
                                std::vector<double> &rec);

}  // namespace mts
