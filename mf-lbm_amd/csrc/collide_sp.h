// Per-node single-phase D3Q19 MRT collision with a body force along z, shared by the even and odd
// lattice-update kernels of the single-phase path.  In FP64 the arithmetic is the reference's, operation for
// operation (singlephase_3D/0.src/Kernel.F90:56-165 odd step, :256-365 even step; both parities run the
// same collision), so results are bit-identical when compiled without FMA contraction.
#pragma once
#include "lattice.h"

// relaxation rates (initialization_basic, singlephase_3D/0.src/Initialization.F90:87-112) and force
template <class T>
struct SPCollideParamsT {
    T s_e, s_e2, s_q, s_nu, s_pi, s_t, force_z;
};
using SPCollideParams = SPCollideParamsT<double>;

// the same rates rounded for the FP32 collision
inline SPCollideParamsT<float> to_float(const SPCollideParams& P) {
    return SPCollideParamsT<float>{(float)P.s_e, (float)P.s_e2, (float)P.s_q, (float)P.s_nu,
                                   (float)P.s_pi, (float)P.s_t, (float)P.force_z};
}

// MRT rate sets of the single-phase solver.  omega = 1/(3 nu + 1/2) sets the shear rate s_nu.
// preset 1 ("optimised for bounce-back"): s_e = s_e2 = s_pi = omega, s_q = s_t = 8(2 - omega)/(8 - omega).
// Any other preset runs the single-relaxation-time set (every rate omega): the reference's branch meant for
// the "original" set tests preset == 1 a second time and is unreachable (Initialization.F90:96-103).
inline SPCollideParams sp_collide_params(double la_nu, int preset, double force_z) {
    const double omega = 1.0 / (3.0 * la_nu + 0.5);
    SPCollideParams P{};
    P.s_nu = omega;
    if (preset == 1) {
        P.s_e = omega;
        P.s_e2 = omega;
        P.s_pi = omega;
        P.s_q = 8.0 * (2.0 - omega) / (8.0 - omega);
        P.s_t = P.s_q;
    } else {
        P.s_e = omega;
        P.s_e2 = omega;
        P.s_pi = omega;
        P.s_q = omega;
        P.s_t = omega;
    }
    P.force_z = force_z;
    return P;
}

// Rest equilibrium (rho = 1, u = 0) of this MRT model: the moments (rho, e, e2) = (1, -11, MRT_E2_C1), all
// others 0, mapped back to populations by the transform below.  With MRT_E2_C1 = 0 (the e2 equilibrium has no
// density term) it is (4/21, 13/126, 1/63) rather than the lattice weights (1/3, 1/18, 1/36).  The equilibria
// are linear in the density, u is the momentum itself (not divided by the density) and REST carries density 1
// and no momentum, so collide(REST + d) = REST + collide(d) in exact arithmetic: the deviations d obey the
// same collision code, with den the density deviation.  The FP32 storage mode keeps d.
constexpr double REST0 = lat::MRT_C1 + 330.0 * lat::MRT_C2 + 12.0 * lat::MRT_C3 * lat::MRT_E2_C1;
constexpr double REST1 = lat::MRT_C1 + 121.0 * lat::MRT_C2 - 4.0 * lat::MRT_C3 * lat::MRT_E2_C1;
constexpr double REST2 = lat::MRT_C1 - 88.0 * lat::MRT_C2 + lat::MRT_C3 * lat::MRT_E2_C1;
MFLBM_HD constexpr double rest_weight(int q) { return q == 0 ? REST0 : (q < 7 ? REST1 : REST2); }

// The MRT collision with a body force (fx, fy, fz), shared by the single-phase kernels (force (0, 0, force_z))
// and the FP32 two-phase kernel (the CSF force; P.force_z is not read here).
// in:  a[q] = incoming populations (q = 0..18); out: a[q] = outgoing post-collision populations.
// T = double is the reference arithmetic; T = float the FP32 variant (every constant rounded to float).
template <class T>
MFLBM_HD inline void mrt_relax(T* __restrict__ a, const SPCollideParamsT<T>& P, const T fx, const T fy, const T fz) {
    using namespace lat;
    const T s_e = P.s_e, s_e2 = P.s_e2, s_q = P.s_q, s_nu = P.s_nu, s_pi = P.s_pi, s_t = P.s_t;
    T ft0 = a[0], ft1 = a[1], ft2 = a[2], ft3 = a[3], ft4 = a[4], ft5 = a[5], ft6 = a[6], ft7 = a[7];
    T ft8 = a[8], ft9 = a[9], ft10 = a[10], ft11 = a[11], ft12 = a[12], ft13 = a[13], ft14 = a[14];
    T ft15 = a[15], ft16 = a[16], ft17 = a[17], ft18 = a[18];

    const T den = ft0 + ft1 + ft2 + ft3 + ft4 + ft5 + ft6 + ft7 + ft8 + ft9 + ft10 + ft11 + ft12 + ft13 + ft14 +
                       ft15 + ft16 + ft17 + ft18;
    const T ux = ft1 - ft2 + ft7 - ft8 + ft9 - ft10 + ft11 - ft12 + ft13 - ft14 + T(0.5) * fx;
    const T uy = ft3 - ft4 + ft7 + ft8 - ft9 - ft10 + ft15 - ft16 + ft17 - ft18 + T(0.5) * fy;
    const T uz = ft5 - ft6 + ft11 + ft12 - ft13 - ft14 + ft15 + ft16 - ft17 - ft18 + T(0.5) * fz;
    const T u2 = ux * ux + uy * uy + uz * uz;

    T sum1 = ft1 + ft2 + ft3 + ft4 + ft5 + ft6;
    T sum2 = ft7 + ft8 + ft9 + ft10 + ft11 + ft12 + ft13 + ft14 + ft15 + ft16 + ft17 + ft18;
    T sum3 = ft7 - ft8 + ft9 - ft10 + ft11 - ft12 + ft13 - ft14;
    T sum4 = ft7 + ft8 - ft9 - ft10 + ft15 - ft16 + ft17 - ft18;
    T sum5 = ft11 + ft12 - ft13 - ft14 + ft15 + ft16 - ft17 - ft18;
    T sum6 = T(2.0) * (ft1 + ft2) - ft3 - ft4 - ft5 - ft6;
    T sum7 = ft7 + ft8 + ft9 + ft10 + ft11 + ft12 + ft13 + ft14 - T(2.0) * (ft15 + ft16 + ft17 + ft18);
    T sum8 = ft3 + ft4 - ft5 - ft6;
    T sum9 = ft7 + ft8 + ft9 + ft10 - ft11 - ft12 - ft13 - ft14;

    T m_rho = den;
    T m_e = -T(30.0) * ft0 - T(11.0) * sum1 + T(8.0) * sum2;
    T m_e2 = T(12.0) * ft0 - T(4.0) * sum1 + sum2;
    T m_jx = ft1 - ft2 + sum3;
    T m_qx = -T(4.0) * (ft1 - ft2) + sum3;
    T m_jy = ft3 - ft4 + sum4;
    T m_qy = -T(4.0) * (ft3 - ft4) + sum4;
    T m_jz = ft5 - ft6 + sum5;
    T m_qz = -T(4.0) * (ft5 - ft6) + sum5;
    T m_3pxx = sum6 + sum7;
    T m_3pixx = -T(2.0) * sum6 + sum7;
    T m_pww = sum8 + sum9;
    T m_piww = -T(2.0) * sum8 + sum9;
    T m_pxy = ft7 - ft8 - ft9 + ft10;
    T m_pyz = ft15 - ft16 - ft17 + ft18;
    T m_pzx = ft11 - ft12 - ft13 + ft14;
    T m_tx = ft7 - ft8 + ft9 - ft10 - ft11 + ft12 - ft13 + ft14;
    T m_ty = -ft7 - ft8 + ft9 + ft10 + ft15 - ft16 + ft17 - ft18;
    T m_tz = ft11 + ft12 - ft13 - ft14 - ft15 - ft16 + ft17 + ft18;

    m_e = m_e - s_e * (m_e - (-T(11.0) * den + T(19.0) * u2)) + (T(38.0) - T(19.0) * s_e) * (fx * ux + fy * uy + fz * uz);
    m_e2 = m_e2 - s_e2 * (m_e2 - (T(MRT_E2_C1) * den + T(MRT_E2_C2) * u2)) + (-T(11.0) + T(5.5) * s_e2) * (fx * ux + fy * uy + fz * uz);
    m_jx = m_jx + fx;
    m_qx = m_qx - s_q * (m_qx - (-T(TWO3) * ux)) + (-T(TWO3) + T(ONE3) * s_q) * fx;
    m_jy = m_jy + fy;
    m_qy = m_qy - s_q * (m_qy - (-T(TWO3) * uy)) + (-T(TWO3) + T(ONE3) * s_q) * fy;
    m_jz = m_jz + fz;
    m_qz = m_qz - s_q * (m_qz - (-T(TWO3) * uz)) + (-T(TWO3) + T(ONE3) * s_q) * fz;
    m_3pxx = m_3pxx - s_nu * (m_3pxx - (T(3.0) * ux * ux - u2)) + (T(2.0) - s_nu) * (T(2.0) * fx * ux - fy * uy - fz * uz);
    m_3pixx = m_3pixx - s_pi * (m_3pixx - T(MRT_OMEGA_XX) * (T(3.0) * ux * ux - u2)) +
              (T(1.0) - T(0.5) * s_pi) * (-T(2.0) * fx * ux + fy * uy + fz * uz);
    m_pww = m_pww - s_nu * (m_pww - (uy * uy - uz * uz)) + (T(2.0) - s_nu) * (fy * uy - fz * uz);
    m_piww = m_piww - s_pi * (m_piww - T(MRT_OMEGA_XX) * (uy * uy - uz * uz)) + (T(1.0) - T(0.5) * s_pi) * (-fy * uy + fz * uz);
    m_pxy = m_pxy - s_nu * (m_pxy - (ux * uy)) + (T(1.0) - T(0.5) * s_nu) * (fx * uy + fy * ux);
    m_pyz = m_pyz - s_nu * (m_pyz - (uy * uz)) + (T(1.0) - T(0.5) * s_nu) * (fy * uz + fz * uy);
    m_pzx = m_pzx - s_nu * (m_pzx - (ux * uz)) + (T(1.0) - T(0.5) * s_nu) * (fx * uz + fz * ux);
    m_tx = m_tx - s_t * (m_tx);
    m_ty = m_ty - s_t * (m_ty);
    m_tz = m_tz - s_t * (m_tz);

    m_rho = T(MRT_C1) * m_rho;
    m_e = T(MRT_C2) * m_e;
    m_e2 = T(MRT_C3) * m_e2;
    m_jx = T(0.1) * m_jx;
    m_qx = T(0.025) * m_qx;
    m_jy = T(0.1) * m_jy;
    m_qy = T(0.025) * m_qy;
    m_jz = T(0.1) * m_jz;
    m_qz = T(0.025) * m_qz;
    m_3pxx = T(2.0) * T(MRT_C4) * m_3pxx;
    m_3pixx = T(MRT_C4) * m_3pixx;
    m_pww = T(6.0) * T(MRT_C4) * m_pww;
    m_piww = T(3.0) * T(MRT_C4) * m_piww;
    m_pxy = T(0.25) * m_pxy;
    m_pyz = T(0.25) * m_pyz;
    m_pzx = T(0.25) * m_pzx;
    m_tx = T(0.125) * m_tx;
    m_ty = T(0.125) * m_ty;
    m_tz = T(0.125) * m_tz;
    sum1 = m_rho - T(11.0) * m_e - T(4.0) * m_e2;
    sum2 = T(2.0) * m_3pxx - T(4.0) * m_3pixx;
    sum3 = m_pww - T(2.0) * m_piww;
    sum4 = m_rho + T(8.0) * m_e + m_e2;
    sum5 = m_jx + m_qx;
    sum6 = m_jy + m_qy;
    sum7 = m_jz + m_qz;
    sum8 = m_3pxx + m_3pixx;
    sum9 = m_pww + m_piww;

    a[0] = m_rho - T(30.0) * m_e + T(12.0) * m_e2;
    a[1] = sum1 + m_jx - T(4.0) * m_qx + sum2;
    a[2] = sum1 - m_jx + T(4.0) * m_qx + sum2;
    a[3] = sum1 + m_jy - T(4.0) * m_qy - T(0.5) * sum2 + sum3;
    a[4] = sum1 - m_jy + T(4.0) * m_qy - T(0.5) * sum2 + sum3;
    a[5] = sum1 + m_jz - T(4.0) * m_qz - T(0.5) * sum2 - sum3;
    a[6] = sum1 - m_jz + T(4.0) * m_qz - T(0.5) * sum2 - sum3;
    a[7] = sum4 + sum5 + sum6 + sum8 + sum9 + m_pxy + m_tx - m_ty;
    a[8] = sum4 - sum5 + sum6 + sum8 + sum9 - m_pxy - m_tx - m_ty;
    a[9] = sum4 + sum5 - sum6 + sum8 + sum9 - m_pxy + m_tx + m_ty;
    a[10] = sum4 - sum5 - sum6 + sum8 + sum9 + m_pxy - m_tx + m_ty;
    a[11] = sum4 + sum5 + sum7 + sum8 - sum9 + m_pzx - m_tx + m_tz;
    a[12] = sum4 - sum5 + sum7 + sum8 - sum9 - m_pzx + m_tx + m_tz;
    a[13] = sum4 + sum5 - sum7 + sum8 - sum9 - m_pzx - m_tx - m_tz;
    a[14] = sum4 - sum5 - sum7 + sum8 - sum9 + m_pzx + m_tx - m_tz;
    a[15] = sum4 + sum6 + sum7 - sum8 * T(2.0) + m_pyz + m_ty - m_tz;
    a[16] = sum4 - sum6 + sum7 - sum8 * T(2.0) - m_pyz - m_ty - m_tz;
    a[17] = sum4 + sum6 - sum7 - sum8 * T(2.0) - m_pyz + m_ty + m_tz;
    a[18] = sum4 - sum6 - sum7 - sum8 * T(2.0) + m_pyz - m_ty + m_tz;
}

template <class T>
MFLBM_HD inline void collide_single(T* __restrict__ a, const SPCollideParamsT<T>& P) {
    mrt_relax<T>(a, P, T(0.0), T(0.0), P.force_z);
}
