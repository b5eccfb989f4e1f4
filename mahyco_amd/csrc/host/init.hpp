// Initial state of a problem on the host (the drop-in for Vnr::init and the
// simulate() prologue, vnr_remap/Init.cc), before upload to the device.
#pragma once

#include <string>
#include <vector>

#include "../../../include/mahyco.h"

namespace mhy {

struct Boundary {
  int left = 0, right = 0, bottom = 0, top = 0;  // velocity conditions
  double vleft[2] = {0.0, 1.0}, vright[2] = {0.0, 1.0}, vbottom[2] = {1.0, 0.0}, vtop[2] = {1.0, 0.0};
  int periodic_right = 0;
  int flux_left = 0, flux_right = 0, flux_bottom = 0, flux_top = 0;
};

struct HostInit {
  long nx = 0, ny = 0;
  int nbmat = 0;
  // cell fields (n0 of the reference)
  std::vector<double> rho, p, c, e, vol, M;
  std::vector<double> rho_k[MHY_NBMATMAX], p_k[MHY_NBMATMAX], c_k[MHY_NBMATMAX], e_k[MHY_NBMATMAX];
  std::vector<double> f_k[MHY_NBMATMAX], mu_k[MHY_NBMATMAX], M_k[MHY_NBMATMAX];
  std::vector<int> mixte, pure;
  // node fields
  std::vector<double> ux, uy;
  double deltat_init = 0.0;
  Boundary bc;
};

// Builds the initial state; returns MHY_OK or an error code, with a
// message in err.
int buildInitialState(const mhy_config& o, HostInit& s, std::string& err);

// What an Eucclhyd problem adds: the cell velocity (and the output copy the
// reference writes), the polygon centroid, volume and perimeter of the
// set-up cells.
struct EucInit {
  std::vector<double> vx, vy, outvx, outvy, xc, yc, ve, perim;
};

// Initial state of an Eucclhyd problem (eucclhyd_remap/Init.cc): fills the
// shared fields of s (densities, energies, fractions, flags, boundary
// conditions, initial time step) and u.
int buildEucInitialState(const mhy_config& o, HostInit& s, EucInit& u, std::string& err);

// Outer face normals of the Euler cells (Init.cc:534-545) must be exactly
// +-1 along their axis for the kernels' flux signs; checks every cell, with
// the cell centre of the scheme (VNR: mean of the nodes; Eucclhyd: polygon
// centroid).
bool outerNormalsAxisExact(const mhy_config& o, std::string& err);

}  // namespace mhy
