// Host-side scene state of one simulation (plain C++, shared by capi.hip and scene.cpp): the progress
// of every water source and the last pose of every obstacle and drain, keyed by item id, plus the
// grid quantities the emitters need and the seed stream of the emission records.
#pragma once
#include <stdint.h>
#include <map>
#include <string>

#include "../../include/suiri.h"

// progress of one source: fractional particles carried to the next substep, litres emitted, time to
// the next drop or burst, and whether a burst source has fired once
struct SlSourceState {
  double carry = 0, liters = 0, wait = 0;
  bool fired = false;
};
// pose of an obstacle or drain when the objects were last packed: position and rotation (row-major,
// world = R * local)
struct SlPose {
  double pos[3];
  double R[9];
};

struct sl_scene {
  // grid of the simulation (50-sim.js:33-45): cell size, origin (min corner), particle volume, tank
  double h = 0, origin[3] = {0, 0, 0}, vp = 0, tank[3] = {0, 0, 0};
  std::map<uint32_t, SlSourceState> sources;
  std::map<uint32_t, SlPose> poses;
  uint32_t seed_state = 1;
  std::string err;
};

// grid sizing of a configuration, 50-sim.js:33-45 (capi.hip); 0 or an SL_E* code
int sl_grid_of(const sl_config &c, int n[3], double &h, double origin[3]);
// scene state of a fresh simulation of `cfg`
int sl_scene_init(sl_scene &sc, const sl_config &cfg);
