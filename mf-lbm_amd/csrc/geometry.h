// Geometry preprocessing pieces shared by the host and the device implementation (geometry_preprocessing_new,
// Geometry_preprocessing.F90:9-512): the list container, the 27-point smoothing stencil and the ISO8 wall-normal
// gradient.  One definition serves both so the two paths evaluate the same expressions in the same order.
#pragma once
#include <cstdint>
#include <vector>

#include "lattice.h"

struct mflbm_geometry {
    std::vector<int32_t> s_ijk, s_nfl, s_nbr;
    std::vector<double> s_w;
    std::vector<int32_t> f_ijk;
    std::vector<double> f_nw, f_theta;
};

namespace geo {

constexpr int GL = 10;   // ghost layers of the extended wall copy (Geometry_preprocessing.F90:54-83)

// 27-point smoothing: offsets and weights by squared length (Geometry_preprocessing.F90:170-226)
constexpr int IEX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
constexpr int IEY[27] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1};
constexpr int IEZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, 1, 1, -1};
constexpr double WE[4] = {8.0 / 27.0, 2.0 / 27.0, 1.0 / 54.0, 1.0 / 216.0};

template <class F>
MFLBM_HD inline double smooth27(const F& S, int i, int j, int k) {
    double acc = 0.0;
    for (int n = 0; n < 27; ++n) {
        const int m = IEX[n] * IEX[n] + IEY[n] * IEY[n] + IEZ[n] * IEZ[n];
        acc = acc + S(i + IEX[n], j + IEY[n], k + IEZ[n]) * WE[m];
    }
    return acc;
}

// ISO8 gradient of the smoothed wall field at (i, j, k) (Geometry_preprocessing.F90:230-368)
template <class F>
MFLBM_HD inline void iso8_gradient(const F& S, int i, int j, int k, double& nwx, double& nwy, double& nwz) {
    const double I1 = 4.0 / 45.0, I2 = 1.0 / 21.0, I3 = 2.0 / 105.0, I4 = 5.0 / 504.0, I5 = 1.0 / 315.0,
                 I6 = 1.0 / 630.0, I7 = 1.0 / 5040.0;
    nwx = I1 * (S(i + 1, j, k) - S(i - 1, j, k)) +
        I2 * (S(i + 1, j + 1, k) - S(i - 1, j - 1, k) + S(i + 1, j - 1, k) - S(i - 1, j + 1, k) +
              S(i + 1, j, k + 1) - S(i - 1, j, k - 1) + S(i + 1, j, k - 1) - S(i - 1, j, k + 1)) +
        I3 * (S(i + 1, j + 1, k + 1) - S(i - 1, j - 1, k - 1) + S(i + 1, j + 1, k - 1) - S(i - 1, j - 1, k + 1) +
              S(i + 1, j - 1, k + 1) - S(i - 1, j + 1, k - 1) + S(i + 1, j - 1, k - 1) - S(i - 1, j + 1, k + 1)) +
        I4 * (S(i + 2, j, k) - S(i - 2, j, k)) +
        I5 * (S(i + 2, j + 1, k) - S(i - 2, j - 1, k) + S(i + 2, j - 1, k) - S(i - 2, j + 1, k) +
              S(i + 2, j, k + 1) - S(i - 2, j, k - 1) + S(i + 2, j, k - 1) - S(i - 2, j, k + 1) +
              S(i + 1, j + 2, k) - S(i - 1, j - 2, k) + S(i + 1, j - 2, k) - S(i - 1, j + 2, k) +
              S(i + 1, j, k + 2) - S(i - 1, j, k - 2) + S(i + 1, j, k - 2) - S(i - 1, j, k + 2)) +
        I6 * (S(i + 2, j + 1, k + 1) - S(i - 2, j - 1, k - 1) + S(i + 2, j + 1, k - 1) - S(i - 2, j - 1, k + 1) +
              S(i + 2, j - 1, k + 1) - S(i - 2, j + 1, k - 1) + S(i + 2, j - 1, k - 1) - S(i - 2, j + 1, k + 1) +
              S(i + 1, j + 2, k + 1) - S(i - 1, j - 2, k - 1) + S(i + 1, j + 2, k - 1) - S(i - 1, j - 2, k + 1) +
              S(i + 1, j - 2, k + 1) - S(i - 1, j + 2, k - 1) + S(i + 1, j - 2, k - 1) - S(i - 1, j + 2, k + 1) +
              S(i + 1, j + 1, k + 2) - S(i - 1, j - 1, k - 2) + S(i + 1, j + 1, k - 2) - S(i - 1, j - 1, k + 2) +
              S(i + 1, j - 1, k + 2) - S(i - 1, j + 1, k - 2) + S(i + 1, j - 1, k - 2) - S(i - 1, j + 1, k + 2)) +
        I7 * (S(i + 2, j + 2, k) - S(i - 2, j - 2, k) + S(i + 2, j - 2, k) - S(i - 2, j + 2, k) +
              S(i + 2, j, k + 2) - S(i - 2, j, k - 2) + S(i + 2, j, k - 2) - S(i - 2, j, k + 2));
    nwy = I1 * (S(i, j + 1, k) - S(i, j - 1, k)) +
        I2 * (S(i + 1, j + 1, k) - S(i - 1, j - 1, k) + S(i - 1, j + 1, k) - S(i + 1, j - 1, k) +
              S(i, j + 1, k + 1) - S(i, j - 1, k - 1) + S(i, j + 1, k - 1) - S(i, j - 1, k + 1)) +
        I3 * (S(i + 1, j + 1, k + 1) - S(i - 1, j - 1, k - 1) + S(i + 1, j + 1, k - 1) - S(i - 1, j - 1, k + 1) +
              S(i - 1, j + 1, k - 1) - S(i + 1, j - 1, k + 1) + S(i - 1, j + 1, k + 1) - S(i + 1, j - 1, k - 1)) +
        I4 * (S(i, j + 2, k) - S(i, j - 2, k)) +
        I5 * (S(i + 2, j + 1, k) - S(i - 2, j - 1, k) + S(i - 2, j + 1, k) - S(i + 2, j - 1, k) +
              S(i, j + 2, k + 1) - S(i, j - 2, k - 1) + S(i, j + 2, k - 1) - S(i, j - 2, k + 1) +
              S(i + 1, j + 2, k) - S(i - 1, j - 2, k) + S(i - 1, j + 2, k) - S(i + 1, j - 2, k) +
              S(i, j + 1, k + 2) - S(i, j - 1, k - 2) + S(i, j + 1, k - 2) - S(i, j - 1, k + 2)) +
        I6 * (S(i + 2, j + 1, k + 1) - S(i - 2, j - 1, k - 1) + S(i + 2, j + 1, k - 1) - S(i - 2, j - 1, k + 1) +
              S(i - 2, j + 1, k + 1) - S(i + 2, j - 1, k - 1) + S(i - 2, j + 1, k - 1) - S(i + 2, j - 1, k + 1) +
              S(i + 1, j + 2, k + 1) - S(i - 1, j - 2, k - 1) + S(i + 1, j + 2, k - 1) - S(i - 1, j - 2, k + 1) +
              S(i - 1, j + 2, k + 1) - S(i + 1, j - 2, k - 1) + S(i - 1, j + 2, k - 1) - S(i + 1, j - 2, k + 1) +
              S(i + 1, j + 1, k + 2) - S(i - 1, j - 1, k - 2) + S(i + 1, j + 1, k - 2) - S(i - 1, j - 1, k + 2) +
              S(i - 1, j + 1, k + 2) - S(i + 1, j - 1, k - 2) + S(i - 1, j + 1, k - 2) - S(i + 1, j - 1, k + 2)) +
        I7 * (S(i + 2, j + 2, k) - S(i - 2, j - 2, k) + S(i - 2, j + 2, k) - S(i + 2, j - 2, k) +
              S(i, j + 2, k + 2) - S(i, j - 2, k - 2) + S(i, j + 2, k - 2) - S(i, j - 2, k + 2));
    nwz = I1 * (S(i, j, k + 1) - S(i, j, k - 1)) +
        I2 * (S(i, j + 1, k + 1) - S(i, j - 1, k - 1) + S(i, j - 1, k + 1) - S(i, j + 1, k - 1) +
              S(i + 1, j, k + 1) - S(i - 1, j, k - 1) + S(i - 1, j, k + 1) - S(i + 1, j, k - 1)) +
        I3 * (S(i + 1, j + 1, k + 1) - S(i - 1, j - 1, k - 1) + S(i + 1, j - 1, k + 1) - S(i - 1, j + 1, k - 1) +
              S(i - 1, j + 1, k + 1) - S(i + 1, j - 1, k - 1) + S(i - 1, j - 1, k + 1) - S(i + 1, j + 1, k - 1)) +
        I4 * (S(i, j, k + 2) - S(i, j, k - 2)) +
        I5 * (S(i, j + 1, k + 2) - S(i, j - 1, k - 2) + S(i, j - 1, k + 2) - S(i, j + 1, k - 2) +
              S(i + 2, j, k + 1) - S(i - 2, j, k - 1) + S(i - 2, j, k + 1) - S(i + 2, j, k - 1) +
              S(i, j + 2, k + 1) - S(i, j - 2, k - 1) + S(i, j - 2, k + 1) - S(i, j + 2, k - 1) +
              S(i + 1, j, k + 2) - S(i - 1, j, k - 2) + S(i - 1, j, k + 2) - S(i + 1, j, k - 2)) +
        I6 * (S(i + 2, j + 1, k + 1) - S(i - 2, j - 1, k - 1) + S(i + 2, j - 1, k + 1) - S(i - 2, j + 1, k - 1) +
              S(i - 2, j + 1, k + 1) - S(i + 2, j - 1, k - 1) + S(i - 2, j - 1, k + 1) - S(i + 2, j + 1, k - 1) +
              S(i + 1, j + 2, k + 1) - S(i - 1, j - 2, k - 1) + S(i + 1, j - 2, k + 1) - S(i - 1, j + 2, k - 1) +
              S(i - 1, j + 2, k + 1) - S(i + 1, j - 2, k - 1) + S(i - 1, j - 2, k + 1) - S(i + 1, j + 2, k - 1) +
              S(i + 1, j + 1, k + 2) - S(i - 1, j - 1, k - 2) + S(i + 1, j - 1, k + 2) - S(i - 1, j + 1, k - 2) +
              S(i - 1, j + 1, k + 2) - S(i + 1, j - 1, k - 2) + S(i - 1, j - 1, k + 2) - S(i + 1, j + 1, k - 2)) +
        I7 * (S(i, j + 2, k + 2) - S(i, j - 2, k - 2) + S(i, j - 2, k + 2) - S(i, j + 2, k - 2) +
              S(i + 2, j, k + 2) - S(i - 2, j, k - 2) + S(i - 2, j, k + 2) - S(i + 2, j, k - 2));
}

}  // namespace geo
