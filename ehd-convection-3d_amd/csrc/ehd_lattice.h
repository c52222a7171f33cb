// ehd_lattice.h -- D3Q27 lattice tables shared by host and device code.
//
// Direction numbering and vectors follow the pull offsets of gpu_stream
// (LBM.cu:1225-1250): f1[d](x) = f2[d](x - c_d).  Weights LBM.h:73-76.
// Opposites are the pairs (d, d+1) for odd d (LBM.cu:1105-1131).
#pragma once

#include <stdint.h>

namespace ehd {

constexpr int Q = 27;

constexpr int CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
constexpr int CY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
constexpr int CZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};

// weight class: 0 rest (8/27), 1 axis (2/27), 2 edge (1/54), 3 corner (1/216)
constexpr int WCLS[27] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3};

__host__ __device__ constexpr int opp(int d) { return d == 0 ? 0 : ((d & 1) ? d + 1 : d - 1); }

// Signed axis terms (1=x, 2=y, 3=z) of c.u in the equilibria, in the order the
// reference adds them (flow LBM.cu:545-660, charge LBM.cu:560-660).
constexpr int EQF[27][3] = {
    {0, 0, 0},
    {1, 0, 0}, {-1, 0, 0}, {2, 0, 0}, {-2, 0, 0}, {3, 0, 0}, {-3, 0, 0},
    {1, 2, 0}, {-2, -1, 0}, {1, 3, 0}, {-1, -3, 0}, {3, 2, 0}, {-2, -3, 0},
    {1, -2, 0}, {2, -1, 0}, {1, -3, 0}, {3, -1, 0}, {2, -3, 0}, {3, -2, 0},
    {1, 2, 3}, {-2, -1, -3}, {1, 2, -3}, {3, -1, -2}, {1, 3, -2}, {2, -1, -3},
    {2, 3, -1}, {1, -2, -3}};
constexpr int EQH[27][3] = {
    {0, 0, 0},
    {1, 0, 0}, {-1, 0, 0}, {2, 0, 0}, {-2, 0, 0}, {3, 0, 0}, {-3, 0, 0},
    {1, 2, 0}, {-2, -1, 0}, {1, 3, 0}, {-1, -3, 0}, {2, 3, 0}, {-2, -3, 0},
    {1, -2, 0}, {2, -1, 0}, {1, -3, 0}, {3, -1, 0}, {2, -3, 0}, {3, -2, 0},
    {1, 2, 3}, {-1, -2, -3}, {1, 2, -3}, {3, -1, -2}, {1, 3, -2}, {2, -1, -3},
    {2, 3, -1}, {1, -2, -3}};

// Momentum sums (LBM.cu:382-387): per axis the 9 directions with c=+1 then
// the 9 with c=-1, each list in ascending order.
constexpr int MOM[3][18] = {
    {1, 7, 9, 13, 15, 19, 21, 23, 26, 2, 8, 10, 14, 16, 20, 22, 24, 25},
    {3, 7, 11, 14, 17, 19, 21, 24, 25, 4, 8, 12, 13, 18, 20, 22, 23, 26},
    {5, 9, 11, 16, 18, 19, 22, 23, 25, 6, 10, 12, 15, 17, 20, 21, 24, 26}};

// Moving-top-wall term of gpu_boundary (LBM.cu:1145-1170): +1 adds, -1
// subtracts the weight-class term 2*rho0*uw/cs^2*w/CFL, 0 plain bounce-back.
constexpr int TOPSGN[27] = {0, 1, -1, 1, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0,
                            1, -1, 1, -1, 1, -1, -1, 1};

__host__ __device__ constexpr int cv(int d, int a) { return a == 0 ? CX[d] : (a == 1 ? CY[d] : CZ[d]); }

}  // namespace ehd
