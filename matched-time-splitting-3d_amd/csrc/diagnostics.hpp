// End-of-run interface diagnostics (the reference's Error_Fp and Error_Jump).
#pragma once

#include <vector>

#include "host_setup.hpp"

namespace mts {

// Per crossing of direction d (line order): the relative errors of its two
// fictitious points (the weights applied to the exact solution and exact
// jumps at time t, against the other side's exact value) and the signed
// relative error of its GFM jump u_dir against the exact [beta u_d].
void crossing_errors(const SetupResult &s, const BetaModel &beta, int d, double t,
                     std::vector<double> &errl, std::vector<double> &errr,
                     std::vector<double> &jerr);

// Lmax and L2 over all crossings: fp[] of the fictitious points (2 per
// crossing), jump[] of the jumps.
void interface_diagnostics(const SetupResult &s, const BetaModel &beta, double t, double fp[2],
                           double jump[2]);

}  // namespace mts
