// D3Q19 discrete velocity set shared by host and device code.
//
// The velocity ordering is the reference's (src/lbm_d3q19.cu:10-21): the
// restart-dump byte layout (population index fastest) depends on it, so the
// order is kept; everything else about the lattice is constexpr here.
#pragma once

#ifndef DUGKS_HD
#if defined(__HIPCC__)
#define DUGKS_HD __host__ __device__
#else
#define DUGKS_HD
#endif
#endif

namespace dugks {

constexpr int Q = 19;

//                          0  1   2  3   4  5   6  7   8  9  10 11  12 13  14 15  16 17  18
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

constexpr double W0 = 1.0 / 3.0;
constexpr double W1 = 1.0 / 18.0;
constexpr double W2 = 1.0 / 36.0;
constexpr double RT = 1.0 / 3.0;

DUGKS_HD constexpr int cv(int a, int ip) { return a == 0 ? CX[ip] : (a == 1 ? CY[ip] : CZ[ip]); }
DUGKS_HD constexpr double wgt(int ip) { return ip == 0 ? W0 : (ip <= 6 ? W1 : W2); }

}  // namespace dugks
