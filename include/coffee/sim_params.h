// Simulation parameters: the same fields and defaults as Coffee::sim_params
// (src/sim_params.h:15-124), and parse_config reading a config.toml
// (src/sim_params.cpp:7-146) with the same key names and value types.
#ifndef COFFEE_SIM_PARAMS_H
#define COFFEE_SIM_PARAMS_H

#include <cmath>
#include <cstdint>
#include <string>

#include "coffee/typedefs.h"

namespace Coffee {

struct sim_params {
  Scalar dt = 0.01;
  uint64_t max_steps = 10000;
  int data_interval = 100;
  int snapshot_interval = 1000;
  bool periodic_boundary[3] = {false};
  int downsample = 1;
  bool slice_x = false;
  Scalar slice_x_pos = 0.0;
  bool slice_y = false;
  Scalar slice_y_pos = 0.0;
  bool slice_z = false;
  Scalar slice_z_pos = 0.0;
  bool slice_xy = false;
  int slice_interval = 100;
  int downsample_2d = 1;
  int N[3] = {1};
  int guard[3] = {0};
  Scalar lower[3] = {0.0};
  Scalar size[3] = {1.0};
  int nodes[3] = {1};
  int shift_ghost = 0;
  int kn[3] = {1, 1, 0};
  Scalar a = 0.99;
  bool calc_current = true;
  bool clean_ep = true;
  bool check_egb = true;
  int pml[3] = {60, 60, 60};
  int pmllen = 10;
  Scalar sigpml = 0.1;
  bool divB_clean = true;
  Scalar ch2 = 1.0;
  Scalar tau = 2.0;
  Scalar KOeps = 0.50;
  bool KO_geometry = true;
  bool use_edotb_damping = true;
  Scalar damp_gamma = 0.2;
  int problem = 0;
  Scalar radius = 10.0;
  Scalar omega = 0.001;
  Scalar b0 = 1e3;
  Scalar p1 = 0.0, p2 = 0.0, p3 = 1.0;
  Scalar q11 = 0.0, q12 = 0.0, q13 = 0.0, q22 = 0.0, q23 = 0.0;
  Scalar q_offset_x = 0.0, q_offset_y = 0.0, q_offset_z = 0.0;
  Scalar tp_start = 0.0, tp_end = 0.5, rpert1 = 5.0, rpert2 = 10.0, thp1 = 0.3, thp2 = 0.4;
  Scalar dw0 = 0.0, nT = 1.0, shear = 0.0;
  Scalar ph1 = 3.0 / 16.0, ph2 = 5.0 / 16.0, dph = 1.0 / 64.0;
  Scalar theta0 = 0.2, drpert = 0.1;
  Scalar tp_start1 = 1.0, tp_end1 = 1.5, rpert11 = 5.0, rpert21 = 10.0, thp11 = 0.3,
         thp21 = 0.4;
  Scalar dw1 = 0.0, nT1 = 1.0, shear1 = 0.0;
  Scalar ph11 = 3.0 * M_PI / 16.0, ph21 = 5.0 * M_PI / 16.0, dph1 = M_PI / 64.0;
  int pert_type = 0;
  int skymap_Nth = 256;
  int skymap_Nph = 512;
};

// Reads a toml config file into sim_params (keys not present keep defaults).
sim_params parse_config(const std::string& filename);

}  // namespace Coffee

#include "coffee/sim_params_impl.hpp"

#endif
