// Where the reference would have stopped a run that this build completes
// (see ref_checks.cpp).
#pragma once

#include <string>

#include "host_setup.hpp"

namespace mts {

enum : int {
  kStopAuxOutside = 0,  // an auxiliary point outside the domain (set-up exit)
  kStopNoSameSide = 1,  // no three same-side nodes on both neighbouring lines (set-up exit)
  kStopAuxOnNode = 2,   // an auxiliary point on a grid node or NaN (reference: index unset)
  kStopFrameZero = 3,   // the tangent frame component is zero (exit at the first step)
  kStopKinds = 4,
};

struct RefStops {
  int count[kStopKinds] = {0, 0, 0, 0};  // crossings per kind
  int first_kind = -1;                  // the first exit (not kStopAuxOnNode); -1: none
  int first_dir = -1, first_index = -1; // crossing (direction, line-order index)
  std::string message;                  // the reference's message for it
};

void reference_stops(const SetupResult &s, RefStops &out);

}  // namespace mts
