// Per-node two-colour MRT collision with CSF force and R-K recolouring (FP64), shared by the even and
// odd lattice-update kernels.  The arithmetic is the reference's, operation for operation, so results are
// bit-identical when compiled without FMA contraction (Kernel_multiphase.F90:86-315 odd step,
// :450-678 even step; both parities run the same physics).
#pragma once
#include "lattice.h"

struct CollideParams {
    double gamma, beta, force_z, la_nui1, la_nui2;
};

// in:  a[q] = incoming fluid-1 populations, b[q] = incoming fluid-2 populations (q = 0..18),
//      cnx/cny/cnz unit interface normal, tf = 0.5*gamma*curv*c_norm (the CSF magnitude, evaluated by the
//      colour-gradient chain in the reference's order, Kernel_multiphase.F90:116).
// out: a[q], b[q] overwritten with the outgoing recoloured populations; returns phi.
MFLBM_HD inline double collide_node_tf(double* __restrict__ a, double* __restrict__ b, double cnx, double cny,
                                       double cnz, double tf, const CollideParams& P) {
    using namespace lat;
    double ft0 = a[0] + b[0], ft1 = a[1] + b[1], ft2 = a[2] + b[2], ft3 = a[3] + b[3], ft4 = a[4] + b[4];
    double ft5 = a[5] + b[5], ft6 = a[6] + b[6], ft7 = a[7] + b[7], ft8 = a[8] + b[8], ft9 = a[9] + b[9];
    double ft10 = a[10] + b[10], ft11 = a[11] + b[11], ft12 = a[12] + b[12], ft13 = a[13] + b[13];
    double ft14 = a[14] + b[14], ft15 = a[15] + b[15], ft16 = a[16] + b[16], ft17 = a[17] + b[17];
    double ft18 = a[18] + b[18];
    const double rho1 = a[0] + a[1] + a[2] + a[3] + a[4] + a[5] + a[6] + a[7] + a[8] + a[9] + a[10] + a[11] + a[12] +
                        a[13] + a[14] + a[15] + a[16] + a[17] + a[18];
    const double rho2 = b[0] + b[1] + b[2] + b[3] + b[4] + b[5] + b[6] + b[7] + b[8] + b[9] + b[10] + b[11] + b[12] +
                        b[13] + b[14] + b[15] + b[16] + b[17] + b[18];
    const double phi = (rho1 - rho2) / (rho1 + rho2);

    double tmp = tf;
    const double fx = tmp * cnx;
    const double fy = tmp * cny;
    const double fz = tmp * cnz + P.force_z;

    const double omega = 1.0 / (6.0 / ((1.0 + phi) * P.la_nui1 + (1.0 - phi) * P.la_nui2) + 0.5);
    const double s_nu = omega;
    const double s_e = S_E, s_e2 = S_E2, s_pi = S_PI, s_q = S_Q, s_t = S_T;

    const double den = rho1 + rho2;
    const double ux1 = ft1 - ft2 + ft7 - ft8 + ft9 - ft10 + ft11 - ft12 + ft13 - ft14 + 0.5 * fx;
    const double uy1 = ft3 - ft4 + ft7 + ft8 - ft9 - ft10 + ft15 - ft16 + ft17 - ft18 + 0.5 * fy;
    const double uz1 = ft5 - ft6 + ft11 + ft12 - ft13 - ft14 + ft15 + ft16 - ft17 - ft18 + 0.5 * fz;
    const double u2 = ux1 * ux1 + uy1 * uy1 + uz1 * uz1;

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

    m_e = m_e - s_e * (m_e - (-11.0 * den + 19.0 * u2)) + (38.0 - 19.0 * s_e) * (fx * ux1 + fy * uy1 + fz * uz1);
    m_e2 = m_e2 - s_e2 * (m_e2 - (MRT_E2_C1 * den + MRT_E2_C2 * u2)) + (-11.0 + 5.5 * s_e2) * (fx * ux1 + fy * uy1 + fz * uz1);
    m_jx = m_jx + fx;
    m_qx = m_qx - s_q * (m_qx - (-TWO3 * ux1)) + (-TWO3 + ONE3 * s_q) * fx;
    m_jy = m_jy + fy;
    m_qy = m_qy - s_q * (m_qy - (-TWO3 * uy1)) + (-TWO3 + ONE3 * s_q) * fy;
    m_jz = m_jz + fz;
    m_qz = m_qz - s_q * (m_qz - (-TWO3 * uz1)) + (-TWO3 + ONE3 * s_q) * fz;
    m_3pxx = m_3pxx - s_nu * (m_3pxx - (3.0 * ux1 * ux1 - u2)) + (2.0 - s_nu) * (2.0 * fx * ux1 - fy * uy1 - fz * uz1);
    m_3pixx = m_3pixx - s_pi * (m_3pixx - MRT_OMEGA_XX * (3.0 * ux1 * ux1 - u2)) +
              (1.0 - 0.5 * s_pi) * (-2.0 * fx * ux1 + fy * uy1 + fz * uz1);
    m_pww = m_pww - s_nu * (m_pww - (uy1 * uy1 - uz1 * uz1)) + (2.0 - s_nu) * (fy * uy1 - fz * uz1);
    m_piww = m_piww - s_pi * (m_piww - MRT_OMEGA_XX * (uy1 * uy1 - uz1 * uz1)) + (1.0 - 0.5 * s_pi) * (-fy * uy1 + fz * uz1);
    m_pxy = m_pxy - s_nu * (m_pxy - (ux1 * uy1)) + (1.0 - 0.5 * s_nu) * (fx * uy1 + fy * ux1);
    m_pyz = m_pyz - s_nu * (m_pyz - (uy1 * uz1)) + (1.0 - 0.5 * s_nu) * (fy * uz1 + fz * uy1);
    m_pzx = m_pzx - s_nu * (m_pzx - (ux1 * uz1)) + (1.0 - 0.5 * s_nu) * (fx * uz1 + fz * ux1);
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

    ft0 = m_rho - 30.0 * m_e + 12.0 * m_e2;
    ft1 = sum1 + m_jx - 4.0 * m_qx + sum2;
    ft2 = sum1 - m_jx + 4.0 * m_qx + sum2;
    ft3 = sum1 + m_jy - 4.0 * m_qy - 0.5 * sum2 + sum3;
    ft4 = sum1 - m_jy + 4.0 * m_qy - 0.5 * sum2 + sum3;
    ft5 = sum1 + m_jz - 4.0 * m_qz - 0.5 * sum2 - sum3;
    ft6 = sum1 - m_jz + 4.0 * m_qz - 0.5 * sum2 - sum3;
    ft7 = sum4 + sum5 + sum6 + sum8 + sum9 + m_pxy + m_tx - m_ty;
    ft8 = sum4 - sum5 + sum6 + sum8 + sum9 - m_pxy - m_tx - m_ty;
    ft9 = sum4 + sum5 - sum6 + sum8 + sum9 - m_pxy + m_tx + m_ty;
    ft10 = sum4 - sum5 - sum6 + sum8 + sum9 + m_pxy - m_tx + m_ty;
    ft11 = sum4 + sum5 + sum7 + sum8 - sum9 + m_pzx - m_tx + m_tz;
    ft12 = sum4 - sum5 + sum7 + sum8 - sum9 - m_pzx + m_tx + m_tz;
    ft13 = sum4 + sum5 - sum7 + sum8 - sum9 - m_pzx - m_tx - m_tz;
    ft14 = sum4 - sum5 - sum7 + sum8 - sum9 + m_pzx + m_tx - m_tz;
    ft15 = sum4 + sum6 + sum7 - sum8 * 2.0 + m_pyz + m_ty - m_tz;
    ft16 = sum4 - sum6 + sum7 - sum8 * 2.0 - m_pyz - m_ty - m_tz;
    ft17 = sum4 + sum6 - sum7 - sum8 * 2.0 - m_pyz + m_ty + m_tz;
    ft18 = sum4 - sum6 - sum7 - sum8 * 2.0 + m_pyz - m_ty + m_tz;

    // recolouring (Latva-Kokko / Rothman), Kernel_multiphase.F90:272-315
    const double tmp1 = rho1 / den;
    a[0] = tmp1 * ft0;
    b[0] = ft0 * (1.0 - tmp1);
    tmp = rho1 * rho2 * P.beta / den;
    a[1] = tmp1 * ft1 + W1 * tmp * (cnx);
    a[2] = tmp1 * ft2 + W1 * tmp * (-cnx);
    a[3] = tmp1 * ft3 + W1 * tmp * (cny);
    a[4] = tmp1 * ft4 + W1 * tmp * (-cny);
    a[5] = tmp1 * ft5 + W1 * tmp * (cnz);
    a[6] = tmp1 * ft6 + W1 * tmp * (-cnz);
    a[7] = tmp1 * ft7 + RK_W2 * tmp * (cnx + cny);
    a[8] = tmp1 * ft8 + RK_W2 * tmp * (-cnx + cny);
    a[9] = tmp1 * ft9 + RK_W2 * tmp * (cnx - cny);
    a[10] = tmp1 * ft10 + RK_W2 * tmp * (-cnx - cny);
    a[11] = tmp1 * ft11 + RK_W2 * tmp * (cnx + cnz);
    a[12] = tmp1 * ft12 + RK_W2 * tmp * (-cnx + cnz);
    a[13] = tmp1 * ft13 + RK_W2 * tmp * (cnx - cnz);
    a[14] = tmp1 * ft14 + RK_W2 * tmp * (-cnx - cnz);
    a[15] = tmp1 * ft15 + RK_W2 * tmp * (cny + cnz);
    a[16] = tmp1 * ft16 + RK_W2 * tmp * (-cny + cnz);
    a[17] = tmp1 * ft17 + RK_W2 * tmp * (cny - cnz);
    a[18] = tmp1 * ft18 + RK_W2 * tmp * (-cny - cnz);
    b[1] = ft1 - a[1];
    b[2] = ft2 - a[2];
    b[3] = ft3 - a[3];
    b[4] = ft4 - a[4];
    b[5] = ft5 - a[5];
    b[6] = ft6 - a[6];
    b[7] = ft7 - a[7];
    b[8] = ft8 - a[8];
    b[9] = ft9 - a[9];
    b[10] = ft10 - a[10];
    b[11] = ft11 - a[11];
    b[12] = ft12 - a[12];
    b[13] = ft13 - a[13];
    b[14] = ft14 - a[14];
    b[15] = ft15 - a[15];
    b[16] = ft16 - a[16];
    b[17] = ft17 - a[17];
    b[18] = ft18 - a[18];
    return phi;
}

// ------------------------------------------------------------------------------------------------------
// FP32 storage of the two-colour state: per direction q the total deviation s_q = f_q + g_q - REST_q from the
// MRT model's rest equilibrium (collide_sp.h) and the colour difference a_q = f_q - g_q.  Streaming, bounce-back
// and the periodic exchange move f_q and g_q together, hence s_q and a_q too.  The collision acts on the total
// population only (its equilibria are linear in the density and u is the momentum itself), so it runs on s_q
// unchanged (mrt_relax); the recolouring f' = (rho1/rho) ft' + W tmp (e.cn), g' = ft' - f' becomes
// a' = f' - g' = phi ft' + 2 W tmp (e.cn) with ft' = REST + s', phi = (rho1 - rho2)/rho, tmp = rho1 rho2 beta/rho.
// rho = 1 + sum s, rho1 - rho2 = sum a.  No reference counterpart (the FP64 path above is the reference
// arithmetic and the oracle for this one).
// ------------------------------------------------------------------------------------------------------
#include "collide_sp.h"

template <class C>
struct CollideParamsDev {
    C gamma, beta, force_z, la_nui1, la_nui2;
};

// in: s[q], a[q] of the incoming populations; cn unit normal, tf = 0.5*gamma*curv*c_norm.  out: s[q], a[q] of
// the outgoing recoloured populations; returns phi.  C is the arithmetic: float for the FP32 mode, double for
// the mixed mode (FP32 storage, FP64 collision; the kernel rounds s and a to FP32 once on store).
template <class C>
MFLBM_HD inline C collide_node_dev(C* __restrict__ s, C* __restrict__ a, C cnx, C cny, C cnz, C tf,
                                   const CollideParamsDev<C>& P) {
    C dr = s[0], dl = a[0];
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        dr += s[q];
        dl += a[q];
    }
    const C den = C(1.0) + dr;
    const C rho1 = C(0.5) * (den + dl), rho2 = C(0.5) * (den - dl);
    const C phi = dl / den;
    const C fx = tf * cnx, fy = tf * cny, fz = tf * cnz + P.force_z;
    const C omega = C(1.0) / (C(6.0) / ((C(1.0) + phi) * P.la_nui1 + (C(1.0) - phi) * P.la_nui2) + C(0.5));
    const SPCollideParamsT<C> R{C(lat::S_E), C(lat::S_E2), C(lat::S_Q), omega, C(lat::S_PI), C(lat::S_T), C(0.0)};
    mrt_relax<C>(s, R, fx, fy, fz);
    const C t2 = C(2.0) * (rho1 * rho2 * P.beta / den);
    const C k1 = C(lat::W1) * t2, k2 = C(lat::RK_W2) * t2;
    const C r0 = C(REST0), r1 = C(REST1), r2 = C(REST2);
    a[0] = phi * (r0 + s[0]);
    a[1] = phi * (r1 + s[1]) + k1 * cnx;
    a[2] = phi * (r1 + s[2]) - k1 * cnx;
    a[3] = phi * (r1 + s[3]) + k1 * cny;
    a[4] = phi * (r1 + s[4]) - k1 * cny;
    a[5] = phi * (r1 + s[5]) + k1 * cnz;
    a[6] = phi * (r1 + s[6]) - k1 * cnz;
    a[7] = phi * (r2 + s[7]) + k2 * (cnx + cny);
    a[8] = phi * (r2 + s[8]) + k2 * (-cnx + cny);
    a[9] = phi * (r2 + s[9]) + k2 * (cnx - cny);
    a[10] = phi * (r2 + s[10]) + k2 * (-cnx - cny);
    a[11] = phi * (r2 + s[11]) + k2 * (cnx + cnz);
    a[12] = phi * (r2 + s[12]) + k2 * (-cnx + cnz);
    a[13] = phi * (r2 + s[13]) + k2 * (cnx - cnz);
    a[14] = phi * (r2 + s[14]) + k2 * (-cnx - cnz);
    a[15] = phi * (r2 + s[15]) + k2 * (cny + cnz);
    a[16] = phi * (r2 + s[16]) + k2 * (-cny + cnz);
    a[17] = phi * (r2 + s[17]) + k2 * (cny - cnz);
    a[18] = phi * (r2 + s[18]) + k2 * (-cny - cnz);
    return phi;
}
