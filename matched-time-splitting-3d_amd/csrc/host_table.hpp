// Builds the device-resident sparse MIB table from the host setup.
#pragma once

#include <string>
#include <vector>

#include "device_types.hpp"
#include "host_setup.hpp"

namespace mts {

struct DeviceTableHost {
  std::vector<int> blk_ptr[3];      // line id -> block range (n_lines + 1)
  std::vector<BlockHdr> blk_hdr[3];
  std::vector<double> blk_val[3];
  std::vector<double> cr_val[3];    // kCrLen per crossing
  std::vector<uint8_t> seq[3];      // per line: 1 = solve with sequential Thomas
  std::vector<uint8_t> cls0[3];     // per line: 1 = first node outside
  int n_seq[3] = {0, 0, 0};
  std::vector<int> spec;            // kSpecLen per special node (Douglas stage 1)
  std::vector<double> tab[3];       // kTabCount * n[a] per axis
  PhysConst phys;
  int n_blocks[3] = {0, 0, 0};
};

// theta_dt: the implicit coefficient (coef * dt) of the method's LHS
// (Douglas and LOD-IE: dt; LOD-CN and TS: 0.5 dt).
// Returns an empty string, or why the table cannot represent the setup.
std::string build_device_table(const SetupResult &s, const BetaModel &beta, double theta_dt,
                               bool with_special, DeviceTableHost &t);

// Operator weights of node (ix, iy, iz) along axis d, the reference's
// Operator_weights_{x,y,z} (temporal/douglas_adi.cc:1444-1539).
void operator_weights(const SetupResult &s, const BetaModel &beta, int d, int ix, int iy,
                      int iz, double vd[3]);

}  // namespace mts
