// gx_internal.hpp -- host-side definitions shared by the library's sources.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/galax_amd.h"
#include "gx_potential.hpp"

struct gx_pot {
  int device = 0;
  gx::DPot dev{};                 // pointers into `block`
  void* block = nullptr;  // one device allocation: comps, tracks, knots, aux
  int ncomp = 0;
  bool has_track = false;
  bool has_bar = false;
  bool basic = true;  // only origin-centred Kepler/Hernquist/MN/NFW/Plummer
  int sig = 0;        // a composition with a specialised evaluator, or 0
  int kind0 = 0;      // kind of the first component
};

// Table signatures with a compile-time evaluator: the Milky Way model; the
// Milky Way model plus one Plummer sphere on a track (the mock-stream field
// with the progenitor's own potential); the Milky Way model plus one group
// of perturbers rotating rigidly together (a bar and co-rotating Plummer
// subhaloes).
enum : int {
  kSigNone = 0,
  kSigMilkyWay = 1,
  kSigMilkyWayTracked = 2,
  kSigMilkyWayRotating = 3,
  kSigMilkyWayPieces = 4  // the mock-stream field on a dense-solution track
};

namespace gx {

inline int check(hipError_t e) { return e == hipSuccess ? GX_OK : GX_ELAUNCH; }

// Launch-shape defaults, overridable through gx_set_launch_shape and
// gx_set_launch_fraction.
struct LaunchShape {
  int block = 256;
  int blocks_per_cu = 4;
  int parts = 1;  // a persistent launch takes 1/parts of the device
  long long head_lanes = 4096;  // lanes that start on the queue head
  int chunk = 1;  // tickets a lane takes from the work queue at a time
};
LaunchShape& launch_shape();

int num_cus(int device);

}  // namespace gx
