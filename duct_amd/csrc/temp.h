// Mixed-boundary (temperature) solve: the right-hand side correction for the wall data and
// the edge reconstruction of solvediag_temp (solvediag_temp.f90:2-352); the four eigenbasis
// products are the unfolded projection passes with separate y and z bases.
#pragma once
#include "common.h"

namespace duct {

// Per-direction vectors of the walls, 4 my + 4 mz doubles:
//   y: [0 my) D2y(:, 1), then D2y(:, my), Dy(1, :), Dy(my, :), my each
//   z: from 4 my: D2z(:, 1), D2z(:, mz), Dz(1, :), Dz(mz, :), mz each
struct TempArgs {
    const d2v* rhs;        // complex fields (my, mz, nmodes), y fastest
    d2v* out;              // k_temp_rhs: corrected interior; k_temp_edges: solution (edges, corners)
    const double* bcy;     // [mode][2][mz] wall data over z (y = -1, y = +1)
    const double* bcz;     // [mode][2][my] wall data over y (z = -1, z = +1)
    const double* vec;     // the wall vectors above
    int my, mz, nmodes, mode0;
    int by1, by2, bz1, bz2;   // 0 = Dirichlet, 1 = Neumann
};
// out interior = rhs interior + bcontv + bconth (mod_term_FTy / mod_term_FTz), Re and Im alike
hipError_t launch_temp_rhs(const TempArgs& a, hipStream_t st);
// edge_value_y / edge_value_z on out's edges from out's interior; corners copied from rhs;
// Im zero at k = 0
hipError_t launch_temp_edges(const TempArgs& a, hipStream_t st);

}  // namespace duct
