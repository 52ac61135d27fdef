// D3Q19 lattice and model constants shared by the host set-up and the gfx950 kernels.
// Values and direction numbering are the reference's (multiphase_3D/0.src/Module.F90:111-132, 116-125,
// 225-228); they are part of the data contract (slot q of every population array means direction q).
#pragma once

#if defined(__HIPCC__)
#define MFLBM_HD __host__ __device__
#else
#define MFLBM_HD
#endif

namespace lat {

// direction vectors e_q and opposite directions (Module.F90:111-114)
constexpr int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

// lattice weights (Module.F90:126-131)
constexpr double W0 = 1.0 / 3.0;
constexpr double W1 = 1.0 / 18.0;
constexpr double W2 = 1.0 / 36.0;
constexpr double W[19] = {W0, W1, W1, W1, W1, W1, W1, W2, W2, W2, W2, W2, W2, W2, W2, W2, W2, W2, W2};

constexpr double PI = 3.14159265358979323846;
constexpr double EPS = 1.110223025e-16;  // the reference's eps (Module.F90:11), not 2^-53

// MRT transform coefficients and equilibria (Module.F90:116-125, preprocessor.h mrt = 2)
constexpr double MRT_C1 = 1.0 / 19.0;
constexpr double MRT_C2 = 1.0 / 2394.0;
constexpr double MRT_C3 = 1.0 / 252.0;
constexpr double MRT_C4 = 1.0 / 72.0;
constexpr double MRT_E2_C1 = 0.0;
constexpr double MRT_E2_C2 = -475.0 / 63.0;
constexpr double MRT_OMEGA_XX = 0.0;
// relaxation rates of the mrt == 2 set (Kernel_multiphase.F90:134-140); s_nu follows the local viscosity
constexpr double S_E = 1.19, S_E2 = 1.4, S_PI = 1.4, S_Q = 1.2, S_T = 1.98;
constexpr double TWO3 = 0.666666666666666667;
constexpr double ONE3 = 0.333333333333333333;
constexpr double ONE6 = 0.166666666666666667;

// recolouring weight of the diagonal directions (Module.F90:225)
// 1/sqrt(2)/36; written as its correctly rounded double so device and host agree without a sqrt.
constexpr double RK_W2 = 0x1.41cfe93ff5199p-6;

// isotropic gradient weights (Module.F90:227-228)
constexpr double ISO4_1 = 1.0 / 6.0;
constexpr double ISO4_2 = 1.0 / 12.0;

}  // namespace lat
