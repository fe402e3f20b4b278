// Shared constants, index maps and helpers for the MI355X GRMHD fluid step.
//
// Block (patch) geometry follows the reference: 16^3 active zones with two
// ghost layers per side, 20^3 stored cells, k fastest (src/fluid/grid.hpp:19-21,
// src/fluid/grid/grid.cpp:51-89).  The face and edge index maps are kept
// identical to the reference's because the reference leaves fluxes outside
// the computed face ranges at zero and feeds them to the EMF average; the
// padded index space is part of the algorithm, not just a layout.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BH_HD __host__ __device__ __forceinline__
#else
#define BH_HD inline
#endif

namespace bh {

using std::isfinite;

constexpr int BLK = 16;
constexpr int GH = 2;
constexpr int CPN = BLK + 2 * GH;                 // 20
constexpr int NCELL = CPN * CPN * CPN;            // 8000 stored cells per block
constexpr int NINT = BLK * BLK * BLK;             // 4096 active zones per block
constexpr int NFACE = (CPN + 1) * CPN * CPN;      // 8400 padded faces per direction
constexpr int NEDGE = CPN * (CPN + 1) * (CPN + 1);  // 8820 padded edges per direction
constexpr int NFC = (BLK + 1) * BLK * BLK;        // 4352 computed faces per direction
constexpr int NGHOST = NCELL - NINT;              // 3904 ghost cells per block

// State vector component order.  U: D, S0, S1, S2, tau, B0, B1, B2 (cons.hpp:14-59)
//                                W: rho, eps, v0, v1, v2, B0, B1, B2 (prim.hpp:14-27)
enum { UD = 0, US0 = 1, US1 = 2, US2 = 3, UTAU = 4, UB0 = 5, UB1 = 6, UB2 = 7 };
enum { WRHO = 0, WEPS = 1, WV0 = 2, WV1 = 3, WV2 = 4, WB0 = 5, WB1 = 6, WB2 = 7 };

BH_HD int lindx(int i, int j, int k) { return (i + GH) * CPN * CPN + (j + GH) * CPN + (k + GH); }
// x-faces / Bfx: [21][20][20]; y-faces / Bfy / EMFz: [20][21][20]; z-faces / Bfz / EMFy: [20][20][21]
BH_HD int ixf(int i, int j, int k) { return (i + GH) * CPN * CPN + (j + GH) * CPN + (k + GH); }
BH_HD int iyf(int i, int j, int k) { return (i + GH) * (CPN + 1) * CPN + (j + GH) * CPN + (k + GH); }
BH_HD int izf(int i, int j, int k) { return (i + GH) * CPN * (CPN + 1) + (j + GH) * (CPN + 1) + (k + GH); }
// EMFx: [20][21][21]
BH_HD int iex(int i, int j, int k) { return (i + GH) * (CPN + 1) * (CPN + 1) + (j + GH) * (CPN + 1) + (k + GH); }

// std::min / std::max semantics: the first argument wins ties and NaN in the
// second argument never wins.  fmin/fmax would differ on NaN.
BH_HD double smin(double a, double b) { return (b < a) ? b : a; }
BH_HD double smax(double a, double b) { return (a < b) ? b : a; }

}  // namespace bh
