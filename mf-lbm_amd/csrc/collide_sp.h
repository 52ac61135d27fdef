// Per-node single-phase D3Q19 MRT collision with a body force along z (FP64), shared by the even and odd
// lattice-update kernels of the single-phase path.  The arithmetic is the reference's, operation for
// operation (singlephase_3D/0.src/Kernel.F90:56-165 odd step, :256-365 even step; both parities run the
// same collision), so results are bit-identical when compiled without FMA contraction.
#pragma once
#include "lattice.h"

// relaxation rates (initialization_basic, singlephase_3D/0.src/Initialization.F90:87-112) and force
struct SPCollideParams {
    double s_e, s_e2, s_q, s_nu, s_pi, s_t, force_z;
};

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

// in:  a[q] = incoming populations (q = 0..18); out: a[q] = outgoing post-collision populations
MFLBM_HD inline void collide_single(double* __restrict__ a, const SPCollideParams& P) {
    using namespace lat;
    const double fx = 0.0, fy = 0.0, fz = P.force_z;
    const double s_e = P.s_e, s_e2 = P.s_e2, s_q = P.s_q, s_nu = P.s_nu, s_pi = P.s_pi, s_t = P.s_t;
    double ft0 = a[0], ft1 = a[1], ft2 = a[2], ft3 = a[3], ft4 = a[4], ft5 = a[5], ft6 = a[6], ft7 = a[7];
    double ft8 = a[8], ft9 = a[9], ft10 = a[10], ft11 = a[11], ft12 = a[12], ft13 = a[13], ft14 = a[14];
    double ft15 = a[15], ft16 = a[16], ft17 = a[17], ft18 = a[18];

    const double den = ft0 + ft1 + ft2 + ft3 + ft4 + ft5 + ft6 + ft7 + ft8 + ft9 + ft10 + ft11 + ft12 + ft13 + ft14 +
                       ft15 + ft16 + ft17 + ft18;
    const double ux = ft1 - ft2 + ft7 - ft8 + ft9 - ft10 + ft11 - ft12 + ft13 - ft14 + 0.5 * fx;
    const double uy = ft3 - ft4 + ft7 + ft8 - ft9 - ft10 + ft15 - ft16 + ft17 - ft18 + 0.5 * fy;
    const double uz = ft5 - ft6 + ft11 + ft12 - ft13 - ft14 + ft15 + ft16 - ft17 - ft18 + 0.5 * fz;
    const double u2 = ux * ux + uy * uy + uz * uz;

    double sum1 = ft1 + ft2 + ft3 + ft4 + ft5 + ft6;
    double sum2 = ft7 + ft8 + ft9 + ft10 + ft11 + ft12 + ft13 + ft14 + ft15 + ft16 + ft17 + ft18;
    double sum3 = ft7 - ft8 + ft9 - ft10 + ft11 - ft12 + ft13 - ft14;
    double sum4 = ft7 + ft8 - ft9 - ft10 + ft15 - ft16 + ft17 - ft18;
    double sum5 = ft11 + ft12 - ft13 - ft14 + ft15 + ft16 - ft17 - ft18;
    double sum6 = 2.0 * (ft1 + ft2) - ft3 - ft4 - ft5 - ft6;
    double sum7 = ft7 + ft8 + ft9 + ft10 + ft11 + ft12 + ft13 + ft14 - 2.0 * (ft15 + ft16 + ft17 + ft18);
    double sum8 = ft3 + ft4 - ft5 - ft6;
    double sum9 = ft7 + ft8 + ft9 + ft10 - ft11 - ft12 - ft13 - ft14;

    double m_rho = den;
    double m_e = -30.0 * ft0 - 11.0 * sum1 + 8.0 * sum2;
    double m_e2 = 12.0 * ft0 - 4.0 * sum1 + sum2;
    double m_jx = ft1 - ft2 + sum3;
    double m_qx = -4.0 * (ft1 - ft2) + sum3;
    double m_jy = ft3 - ft4 + sum4;
    double m_qy = -4.0 * (ft3 - ft4) + sum4;
    double m_jz = ft5 - ft6 + sum5;
    double m_qz = -4.0 * (ft5 - ft6) + sum5;
    double m_3pxx = sum6 + sum7;
    double m_3pixx = -2.0 * sum6 + sum7;
    double m_pww = sum8 + sum9;
    double m_piww = -2.0 * sum8 + sum9;
    double m_pxy = ft7 - ft8 - ft9 + ft10;
    double m_pyz = ft15 - ft16 - ft17 + ft18;
    double m_pzx = ft11 - ft12 - ft13 + ft14;
    double m_tx = ft7 - ft8 + ft9 - ft10 - ft11 + ft12 - ft13 + ft14;
    double m_ty = -ft7 - ft8 + ft9 + ft10 + ft15 - ft16 + ft17 - ft18;
    double m_tz = ft11 + ft12 - ft13 - ft14 - ft15 - ft16 + ft17 + ft18;

    m_e = m_e - s_e * (m_e - (-11.0 * den + 19.0 * u2)) + (38.0 - 19.0 * s_e) * (fx * ux + fy * uy + fz * uz);
    m_e2 = m_e2 - s_e2 * (m_e2 - (MRT_E2_C1 * den + MRT_E2_C2 * u2)) + (-11.0 + 5.5 * s_e2) * (fx * ux + fy * uy + fz * uz);
    m_jx = m_jx + fx;
    m_qx = m_qx - s_q * (m_qx - (-TWO3 * ux)) + (-TWO3 + ONE3 * s_q) * fx;
    m_jy = m_jy + fy;
    m_qy = m_qy - s_q * (m_qy - (-TWO3 * uy)) + (-TWO3 + ONE3 * s_q) * fy;
    m_jz = m_jz + fz;
    m_qz = m_qz - s_q * (m_qz - (-TWO3 * uz)) + (-TWO3 + ONE3 * s_q) * fz;
    m_3pxx = m_3pxx - s_nu * (m_3pxx - (3.0 * ux * ux - u2)) + (2.0 - s_nu) * (2.0 * fx * ux - fy * uy - fz * uz);
    m_3pixx = m_3pixx - s_pi * (m_3pixx - MRT_OMEGA_XX * (3.0 * ux * ux - u2)) +
              (1.0 - 0.5 * s_pi) * (-2.0 * fx * ux + fy * uy + fz * uz);
    m_pww = m_pww - s_nu * (m_pww - (uy * uy - uz * uz)) + (2.0 - s_nu) * (fy * uy - fz * uz);
    m_piww = m_piww - s_pi * (m_piww - MRT_OMEGA_XX * (uy * uy - uz * uz)) + (1.0 - 0.5 * s_pi) * (-fy * uy + fz * uz);
    m_pxy = m_pxy - s_nu * (m_pxy - (ux * uy)) + (1.0 - 0.5 * s_nu) * (fx * uy + fy * ux);
    m_pyz = m_pyz - s_nu * (m_pyz - (uy * uz)) + (1.0 - 0.5 * s_nu) * (fy * uz + fz * uy);
    m_pzx = m_pzx - s_nu * (m_pzx - (ux * uz)) + (1.0 - 0.5 * s_nu) * (fx * uz + fz * ux);
    m_tx = m_tx - s_t * (m_tx);
    m_ty = m_ty - s_t * (m_ty);
    m_tz = m_tz - s_t * (m_tz);

    m_rho = MRT_C1 * m_rho;
    m_e = MRT_C2 * m_e;
    m_e2 = MRT_C3 * m_e2;
    m_jx = 0.1 * m_jx;
    m_qx = 0.025 * m_qx;
    m_jy = 0.1 * m_jy;
    m_qy = 0.025 * m_qy;
    m_jz = 0.1 * m_jz;
    m_qz = 0.025 * m_qz;
    m_3pxx = 2.0 * MRT_C4 * m_3pxx;
    m_3pixx = MRT_C4 * m_3pixx;
    m_pww = 6.0 * MRT_C4 * m_pww;
    m_piww = 3.0 * MRT_C4 * m_piww;
    m_pxy = 0.25 * m_pxy;
    m_pyz = 0.25 * m_pyz;
    m_pzx = 0.25 * m_pzx;
    m_tx = 0.125 * m_tx;
    m_ty = 0.125 * m_ty;
    m_tz = 0.125 * m_tz;
    sum1 = m_rho - 11.0 * m_e - 4.0 * m_e2;
    sum2 = 2.0 * m_3pxx - 4.0 * m_3pixx;
    sum3 = m_pww - 2.0 * m_piww;
    sum4 = m_rho + 8.0 * m_e + m_e2;
    sum5 = m_jx + m_qx;
    sum6 = m_jy + m_qy;
    sum7 = m_jz + m_qz;
    sum8 = m_3pxx + m_3pixx;
    sum9 = m_pww + m_piww;

    a[0] = m_rho - 30.0 * m_e + 12.0 * m_e2;
    a[1] = sum1 + m_jx - 4.0 * m_qx + sum2;
    a[2] = sum1 - m_jx + 4.0 * m_qx + sum2;
    a[3] = sum1 + m_jy - 4.0 * m_qy - 0.5 * sum2 + sum3;
    a[4] = sum1 - m_jy + 4.0 * m_qy - 0.5 * sum2 + sum3;
    a[5] = sum1 + m_jz - 4.0 * m_qz - 0.5 * sum2 - sum3;
    a[6] = sum1 - m_jz + 4.0 * m_qz - 0.5 * sum2 - sum3;
    a[7] = sum4 + sum5 + sum6 + sum8 + sum9 + m_pxy + m_tx - m_ty;
    a[8] = sum4 - sum5 + sum6 + sum8 + sum9 - m_pxy - m_tx - m_ty;
    a[9] = sum4 + sum5 - sum6 + sum8 + sum9 - m_pxy + m_tx + m_ty;
    a[10] = sum4 - sum5 - sum6 + sum8 + sum9 + m_pxy - m_tx + m_ty;
    a[11] = sum4 + sum5 + sum7 + sum8 - sum9 + m_pzx - m_tx + m_tz;
    a[12] = sum4 - sum5 + sum7 + sum8 - sum9 - m_pzx + m_tx + m_tz;
    a[13] = sum4 + sum5 - sum7 + sum8 - sum9 - m_pzx - m_tx - m_tz;
    a[14] = sum4 - sum5 - sum7 + sum8 - sum9 + m_pzx + m_tx - m_tz;
    a[15] = sum4 + sum6 + sum7 - sum8 * 2.0 + m_pyz + m_ty - m_tz;
    a[16] = sum4 - sum6 + sum7 - sum8 * 2.0 - m_pyz - m_ty - m_tz;
    a[17] = sum4 + sum6 - sum7 - sum8 * 2.0 - m_pyz + m_ty + m_tz;
    a[18] = sum4 - sum6 - sum7 - sum8 * 2.0 + m_pyz - m_ty + m_tz;
}
