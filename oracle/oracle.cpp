// CPU restatement of the reference's two-phase time step (checker only; see oracle.h).
//
// Arrays keep the reference's shapes: populations/curv/u/v/w/rho over (0:nx+1,0:ny+1,0:nz+1), phi over
// (-3:nx+4,...), cn_x/cn_y/cn_z/c_norm and walls over (-1:nx+2,...), w_in and the convective history over
// (0:nx+1,0:ny+1[,0:18]).  Loops run in the reference's order; OpenMP splits only loops whose iterations are
// independent (the AA update writes each slot from exactly one node), so results do not depend on the
// thread count.  Reductions run serially in the reference's loop order.
#include "oracle.h"

#include <omp.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace {

const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
const int OPC[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};   // Module.F90:114
const double W_EQU[19] = {1.0 / 3.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};   // Module.F90:126-127
const double W_EQU_1 = 1.0 / 18.0, W_EQU_2 = 1.0 / 36.0;
const double MRT_COEF1 = 1.0 / 19.0, MRT_COEF2 = 1.0 / 2394.0, MRT_COEF3 = 1.0 / 252.0, MRT_COEF4 = 1.0 / 72.0;
const double MRT_E2_COEF1 = 0.0, MRT_E2_COEF2 = -475.0 / 63.0, MRT_OMEGA_XX = 0.0;   // Module.F90:116-122
const double ISO4_1 = 1.0 / 6.0, ISO4_2 = 1.0 / 12.0;                                 // Module.F90:227

// a 3-D array over (lo:nx+hi...) in Fortran order
struct Arr3 {
    int lo, sx, sy;
    std::vector<double> v;
    void init(int nx, int ny, int nz, int h) {
        lo = 1 - h;
        sx = nx + 2 * h;
        sy = ny + 2 * h;
        v.assign((size_t)sx * sy * (nz + 2 * h), 0.0);
    }
    double& operator()(int i, int j, int k) { return v[(size_t)(i - lo) + (size_t)(j - lo) * sx + (size_t)(k - lo) * sx * sy]; }
    // flat index of (i, j, k) and the strides, for loops that walk rows with a running index
    long at(int i, int j, int k) const { return (long)(i - lo) + (long)(j - lo) * sx + (long)(k - lo) * sx * sy; }
    long pxy() const { return (long)sx * sy; }
};

}  // namespace

struct oracle_ctx {
    oracle_params p;
    int nx, ny, nz;
    Arr3 f[19], g[19], phi, cn_x, cn_y, cn_z, c_norm, curv, u, v, w, rho, phi_old;
    std::vector<int8_t> walls;     // (-1:nx+2, ...)
    std::vector<double> w_in, phi_convec, f_convec, g_convec;   // (0:nx+1, 0:ny+1[, 0:18])
    // boundary-node lists (Module.F90:90-103)
    std::vector<int> s_ijk, s_nfl, s_nbr;
    std::vector<double> s_w;
    std::vector<int> f_ijk;
    std::vector<double> f_nw, f_theta;

    int8_t& W(int i, int j, int k) {
        return walls[(size_t)(i + 1) + (size_t)(j + 1) * (nx + 4) + (size_t)(k + 1) * (nx + 4) * (ny + 4)];
    }
    double& WIN(int i, int j) { return w_in[(size_t)i + (size_t)j * (nx + 2)]; }
    double& PC(int i, int j) { return phi_convec[(size_t)i + (size_t)j * (nx + 2)]; }
    double& FC(int i, int j, int q) { return f_convec[(size_t)i + (size_t)j * (nx + 2) + (size_t)q * (nx + 2) * (ny + 2)]; }
    double& GC(int i, int j, int q) { return g_convec[(size_t)i + (size_t)j * (nx + 2) + (size_t)q * (nx + 2) * (ny + 2)]; }
};

namespace {

// ---------------------------------------------------------------------------------------------------------
// collision + recolouring of one node (Kernel_multiphase.F90:86-315; identical arithmetic at :450-678)
// g1[q], g2[q]: incoming populations of fluid 1 and 2, replaced by the outgoing ones; returns phi
// ---------------------------------------------------------------------------------------------------------
double collide(oracle_ctx* c, double* g1, double* g2, double cnx, double cny, double cnz, double curvv, double cnorm) {
    const oracle_params& P = c->p;
    double ft[19];
    for (int q = 0; q < 19; ++q) ft[q] = g1[q] + g2[q];
    double rho1 = g1[0];
    for (int q = 1; q < 19; ++q) rho1 = rho1 + g1[q];
    double rho2 = g2[0];
    for (int q = 1; q < 19; ++q) rho2 = rho2 + g2[q];
    const double phi = (rho1 - rho2) / (rho1 + rho2);

    double tmp = 0.5 * P.gamma * curvv * cnorm;
    const double fx = tmp * cnx, fy = tmp * cny, fz = tmp * cnz + P.force_z;
    const double omega = 1.0 / (6.0 / ((1.0 + phi) * P.la_nui1 + (1.0 - phi) * P.la_nui2) + 0.5);
    const double s_nu = omega, s_e = 1.19, s_e2 = 1.4, s_pi = 1.4, s_q = 1.2, s_t = 1.98;   // mrt == 2

    const double den = rho1 + rho2;
    const double ux1 = ft[1] - ft[2] + ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14] + 0.5 * fx;
    const double uy1 = ft[3] - ft[4] + ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18] + 0.5 * fy;
    const double uz1 = ft[5] - ft[6] + ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18] + 0.5 * fz;
    const double u2 = ux1 * ux1 + uy1 * uy1 + uz1 * uz1;

    double sum1 = ft[1] + ft[2] + ft[3] + ft[4] + ft[5] + ft[6];
    double sum2 = ft[7] + ft[8] + ft[9] + ft[10] + ft[11] + ft[12] + ft[13] + ft[14] + ft[15] + ft[16] + ft[17] + ft[18];
    double sum3 = ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14];
    double sum4 = ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18];
    double sum5 = ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18];
    double sum6 = 2.0 * (ft[1] + ft[2]) - ft[3] - ft[4] - ft[5] - ft[6];
    double sum7 = ft[7] + ft[8] + ft[9] + ft[10] + ft[11] + ft[12] + ft[13] + ft[14] - 2.0 * (ft[15] + ft[16] + ft[17] + ft[18]);
    double sum8 = ft[3] + ft[4] - ft[5] - ft[6];
    double sum9 = ft[7] + ft[8] + ft[9] + ft[10] - ft[11] - ft[12] - ft[13] - ft[14];

    double m_rho = den;
    double m_e = -(30.0 * ft[0]) - 11.0 * sum1 + 8.0 * sum2;
    double m_e2 = 12.0 * ft[0] - 4.0 * sum1 + sum2;
    double m_jx = ft[1] - ft[2] + sum3;
    double m_qx = -(4.0 * (ft[1] - ft[2])) + sum3;
    double m_jy = ft[3] - ft[4] + sum4;
    double m_qy = -(4.0 * (ft[3] - ft[4])) + sum4;
    double m_jz = ft[5] - ft[6] + sum5;
    double m_qz = -(4.0 * (ft[5] - ft[6])) + sum5;
    double m_3pxx = sum6 + sum7;
    double m_3pixx = -(2.0 * sum6) + sum7;
    double m_pww = sum8 + sum9;
    double m_piww = -(2.0 * sum8) + sum9;
    double m_pxy = ft[7] - ft[8] - ft[9] + ft[10];
    double m_pyz = ft[15] - ft[16] - ft[17] + ft[18];
    double m_pzx = ft[11] - ft[12] - ft[13] + ft[14];
    double m_tx = ft[7] - ft[8] + ft[9] - ft[10] - ft[11] + ft[12] - ft[13] + ft[14];
    double m_ty = -ft[7] - ft[8] + ft[9] + ft[10] + ft[15] - ft[16] + ft[17] - ft[18];
    double m_tz = ft[11] + ft[12] - ft[13] - ft[14] - ft[15] - ft[16] + ft[17] + ft[18];

    const double fu = fx * ux1 + fy * uy1 + fz * uz1;
    const double c23 = 0.666666666666666667, c13 = 0.333333333333333333;
    m_e = m_e - s_e * (m_e - (-(11.0 * den) + 19.0 * u2)) + (38.0 - 19.0 * s_e) * fu;
    m_e2 = m_e2 - s_e2 * (m_e2 - (MRT_E2_COEF1 * den + MRT_E2_COEF2 * u2)) + (-11.0 + 5.5 * s_e2) * fu;
    m_jx = m_jx + fx;
    m_qx = m_qx - s_q * (m_qx - (-(c23 * ux1))) + (-c23 + c13 * s_q) * fx;
    m_jy = m_jy + fy;
    m_qy = m_qy - s_q * (m_qy - (-(c23 * uy1))) + (-c23 + c13 * s_q) * fy;
    m_jz = m_jz + fz;
    m_qz = m_qz - s_q * (m_qz - (-(c23 * uz1))) + (-c23 + c13 * s_q) * fz;
    m_3pxx = m_3pxx - s_nu * (m_3pxx - (3.0 * ux1 * ux1 - u2)) + (2.0 - s_nu) * (2.0 * fx * ux1 - fy * uy1 - fz * uz1);
    m_3pixx = m_3pixx - s_pi * (m_3pixx - MRT_OMEGA_XX * (3.0 * ux1 * ux1 - u2)) +
              (1.0 - 0.5 * s_pi) * (-(2.0 * fx * ux1) + fy * uy1 + fz * uz1);
    m_pww = m_pww - s_nu * (m_pww - (uy1 * uy1 - uz1 * uz1)) + (2.0 - s_nu) * (fy * uy1 - fz * uz1);
    m_piww = m_piww - s_pi * (m_piww - MRT_OMEGA_XX * (uy1 * uy1 - uz1 * uz1)) + (1.0 - 0.5 * s_pi) * (-(fy * uy1) + fz * uz1);
    m_pxy = m_pxy - s_nu * (m_pxy - (ux1 * uy1)) + (1.0 - 0.5 * s_nu) * (fx * uy1 + fy * ux1);
    m_pyz = m_pyz - s_nu * (m_pyz - (uy1 * uz1)) + (1.0 - 0.5 * s_nu) * (fy * uz1 + fz * uy1);
    m_pzx = m_pzx - s_nu * (m_pzx - (ux1 * uz1)) + (1.0 - 0.5 * s_nu) * (fx * uz1 + fz * ux1);
    m_tx = m_tx - s_t * m_tx;
    m_ty = m_ty - s_t * m_ty;
    m_tz = m_tz - s_t * m_tz;

    m_rho = MRT_COEF1 * m_rho;
    m_e = MRT_COEF2 * m_e;
    m_e2 = MRT_COEF3 * m_e2;
    m_jx = 0.1 * m_jx;
    m_qx = 0.025 * m_qx;
    m_jy = 0.1 * m_jy;
    m_qy = 0.025 * m_qy;
    m_jz = 0.1 * m_jz;
    m_qz = 0.025 * m_qz;
    m_3pxx = (2.0 * MRT_COEF4) * m_3pxx;
    m_3pixx = MRT_COEF4 * m_3pixx;
    m_pww = (6.0 * MRT_COEF4) * m_pww;
    m_piww = (3.0 * MRT_COEF4) * m_piww;
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

    double o[19];
    o[0] = m_rho - 30.0 * m_e + 12.0 * m_e2;
    o[1] = sum1 + m_jx - 4.0 * m_qx + sum2;
    o[2] = sum1 - m_jx + 4.0 * m_qx + sum2;
    o[3] = sum1 + m_jy - 4.0 * m_qy - 0.5 * sum2 + sum3;
    o[4] = sum1 - m_jy + 4.0 * m_qy - 0.5 * sum2 + sum3;
    o[5] = sum1 + m_jz - 4.0 * m_qz - 0.5 * sum2 - sum3;
    o[6] = sum1 - m_jz + 4.0 * m_qz - 0.5 * sum2 - sum3;
    o[7] = sum4 + sum5 + sum6 + sum8 + sum9 + m_pxy + m_tx - m_ty;
    o[8] = sum4 - sum5 + sum6 + sum8 + sum9 - m_pxy - m_tx - m_ty;
    o[9] = sum4 + sum5 - sum6 + sum8 + sum9 - m_pxy + m_tx + m_ty;
    o[10] = sum4 - sum5 - sum6 + sum8 + sum9 + m_pxy - m_tx + m_ty;
    o[11] = sum4 + sum5 + sum7 + sum8 - sum9 + m_pzx - m_tx + m_tz;
    o[12] = sum4 - sum5 + sum7 + sum8 - sum9 - m_pzx + m_tx + m_tz;
    o[13] = sum4 + sum5 - sum7 + sum8 - sum9 - m_pzx - m_tx - m_tz;
    o[14] = sum4 - sum5 - sum7 + sum8 - sum9 + m_pzx + m_tx - m_tz;
    o[15] = sum4 + sum6 + sum7 - sum8 * 2.0 + m_pyz + m_ty - m_tz;
    o[16] = sum4 - sum6 + sum7 - sum8 * 2.0 - m_pyz - m_ty - m_tz;
    o[17] = sum4 + sum6 - sum7 - sum8 * 2.0 - m_pyz + m_ty + m_tz;
    o[18] = sum4 - sum6 - sum7 - sum8 * 2.0 + m_pyz - m_ty + m_tz;

    // R-K recolouring: directional perturbation along the interface normal (Kernel_multiphase.F90:272-315)
    const double tmp1 = rho1 / den;
    g1[0] = tmp1 * o[0];
    g2[0] = o[0] * (1.0 - tmp1);
    tmp = rho1 * rho2 * P.beta / den;
    const double rk2 = 1.0 / std::sqrt(2.0) / 36.0;   // RK_weight2, Module.F90:225
    const double dirc[19] = {0.0,         cnx,        -cnx,       cny,        -cny,        cnz,        -cnz,
                             cnx + cny,   -cnx + cny, cnx - cny,  -cnx - cny, cnx + cnz,   -cnx + cnz, cnx - cnz,
                             -cnx - cnz,  cny + cnz,  -cny + cnz, cny - cnz,  -cny - cnz};
    for (int q = 1; q < 19; ++q) {
        const double wq = q <= 6 ? W_EQU_1 : rk2;
        g1[q] = tmp1 * o[q] + wq * tmp * dirc[q];
        g2[q] = o[q] - g1[q];
    }
    return phi;
}

// kernel_even_color (Kernel_multiphase.F90:371-725) / kernel_odd_color (:6-362) over the whole interior.
// The 38 population arrays share one shape (halo 1), so one flat index per node addresses all of them; the
// normal arrays (halo 2), curv (halo 1) and phi (halo 4) have their own.
template <bool ODD>
void lattice_update_t(oracle_ctx* c) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    double* F[19];
    double* G[19];
    long off[19];
    const long sx = c->f[0].sx, sxy = c->f[0].pxy();
    for (int q = 0; q < 19; ++q) {
        F[q] = c->f[q].v.data();
        G[q] = c->g[q].v.data();
        off[q] = EX[q] + EY[q] * sx + EZ[q] * sxy;
    }
    double* const cx = c->cn_x.v.data();
    double* const cy = c->cn_y.v.data();
    double* const cz = c->cn_z.v.data();
    double* const cn = c->c_norm.v.data();
    double* const cu = c->curv.v.data();
    double* const ph = c->phi.v.data();
    // dynamic: nodes near the interface cost more (subnormal populations of the receding colour), so the
    // layers do not cost the same
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j) {
            const long n0 = c->f[0].at(1, j, k), m0 = c->cn_x.at(1, j, k), u0 = c->curv.at(1, j, k), p0 = c->phi.at(1, j, k);
            const int8_t* wrow = &c->W(1, j, k);
            for (int i = 0; i < nx; ++i) {
                if (wrow[i] != 0) continue;
                const long n = n0 + i;
                double g1[19], g2[19];
                g1[0] = F[0][n];
                g2[0] = G[0][n];
                for (int q = 1; q < 19; ++q) {
                    if (ODD) {   // pull from the upwind neighbour
                        g1[q] = F[q][n - off[q]];
                        g2[q] = G[q][n - off[q]];
                    } else {     // own slots, stored swapped after the previous odd step
                        g1[q] = F[OPC[q]][n];
                        g2[q] = G[OPC[q]][n];
                    }
                }
                ph[p0 + i] = collide(c, g1, g2, cx[m0 + i], cy[m0 + i], cz[m0 + i], cu[u0 + i], cn[m0 + i]);
                F[0][n] = g1[0];
                G[0][n] = g2[0];
                for (int q = 1; q < 19; ++q) {
                    if (ODD) {   // push into the downwind neighbour's opposite slot
                        F[OPC[q]][n + off[q]] = g1[q];
                        G[OPC[q]][n + off[q]] = g2[q];
                    } else {
                        F[q][n] = g1[q];
                        G[q][n] = g2[q];
                    }
                }
            }
        }
}
void lattice_update(oracle_ctx* c, bool odd) {
    if (odd) lattice_update_t<true>(c);
    else lattice_update_t<false>(c);
}

// Periodic exchange of one rank with itself (Mpi.F90).  Pull type after even steps (:101-346): the
// halo receives the boundary layer's outgoing populations; push type after odd steps (:354-598): the
// populations pushed into the halo go back to the opposite boundary layer, faces first then x-parallel
// edges.  Buffers are packed before any unpack, as with MPI.
void exchange_pdf(oracle_ctx* c, bool odd) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const bool my = c->p.jper == 1, mz = c->p.kper == 1;
    const int zP[5] = {5, 11, 12, 15, 16}, zM[5] = {6, 14, 13, 18, 17};
    const int yP[5] = {3, 7, 8, 15, 17}, yM[5] = {4, 10, 9, 16, 18};
    for (int col = 0; col < 2; ++col) {
        Arr3* A = col == 0 ? c->f : c->g;
        if (!odd) {
            if (mz)
                for (int j = 1; j <= ny; ++j)
                    for (int i = 1; i <= nx; ++i)
                        for (int m = 0; m < 5; ++m) {
                            A[zP[m]](i, j, 0) = A[zP[m]](i, j, nz);
                            A[zM[m]](i, j, nz + 1) = A[zM[m]](i, j, 1);
                        }
            if (my)
                for (int k = 1; k <= nz; ++k)
                    for (int i = 1; i <= nx; ++i)
                        for (int m = 0; m < 5; ++m) {
                            A[yP[m]](i, 0, k) = A[yP[m]](i, ny, k);
                            A[yM[m]](i, ny + 1, k) = A[yM[m]](i, 1, k);
                        }
            if (my && mz)
                for (int i = 1; i <= nx; ++i) {
                    A[18](i, ny + 1, nz + 1) = A[18](i, 1, 1);
                    A[16](i, ny + 1, 0) = A[16](i, 1, nz);
                    A[17](i, 0, nz + 1) = A[17](i, ny, 1);
                    A[15](i, 0, 0) = A[15](i, ny, nz);
                }
        } else {
            if (mz)
                for (int j = 1; j <= ny; ++j)
                    for (int i = 1; i <= nx; ++i)
                        for (int m = 0; m < 5; ++m) {
                            A[zM[m]](i, j, 1) = A[zM[m]](i, j, nz + 1);
                            A[zP[m]](i, j, nz) = A[zP[m]](i, j, 0);
                        }
            if (my)
                for (int k = 1; k <= nz; ++k)
                    for (int i = 1; i <= nx; ++i)
                        for (int m = 0; m < 5; ++m) {
                            A[yM[m]](i, 1, k) = A[yM[m]](i, ny + 1, k);
                            A[yP[m]](i, ny, k) = A[yP[m]](i, 0, k);
                        }
            if (my && mz)
                for (int i = 1; i <= nx; ++i) {
                    A[15](i, ny, nz) = A[15](i, 0, 0);
                    A[17](i, ny, 1) = A[17](i, 0, nz + 1);
                    A[16](i, 1, nz) = A[16](i, ny + 1, 0);
                    A[18](i, 1, 1) = A[18](i, ny + 1, nz + 1);
                }
        }
    }
}

// phi halo exchange, 4 layers (MPI_phi_async_pack / MPI_phi_sync_update, Mpi.F90:608-867)
void exchange_phi(oracle_ctx* c) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const bool my = c->p.jper == 1, mz = c->p.kper == 1;
    if (mz)
        for (int l = 1; l <= 4; ++l)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    c->phi(i, j, l - 4) = c->phi(i, j, nz + l - 4);
                    c->phi(i, j, l + nz) = c->phi(i, j, l);
                }
    if (my)
        for (int k = 1; k <= nz; ++k)
            for (int l = 1; l <= 4; ++l)
                for (int i = 1; i <= nx; ++i) {
                    c->phi(i, l - 4, k) = c->phi(i, ny + l - 4, k);
                    c->phi(i, l + ny, k) = c->phi(i, l, k);
                }
    if (my && mz)
        for (int kk = 1; kk <= 4; ++kk)
            for (int jj = 1; jj <= 4; ++jj)
                for (int i = 1; i <= nx; ++i) {
                    c->phi(i, jj - 4, kk - 4) = c->phi(i, ny + jj - 4, nz + kk - 4);
                    c->phi(i, jj + ny, kk - 4) = c->phi(i, jj, nz + kk - 4);
                    c->phi(i, jj + ny, kk + nz) = c->phi(i, jj, kk);
                    c->phi(i, jj - 4, kk + nz) = c->phi(i, ny + jj - 4, kk);
                }
}

// inlet phase-field halo (Boundary_multiphase_inlet.F90:26-29)
void inlet_phi(oracle_ctx* c, int i, int j, int wi) {
    c->phi(i, j, 0) = c->p.phi_inlet * (double)(1 - wi) + c->phi(i, j, 0) * (double)wi;
    c->phi(i, j, -1) = c->phi(i, j, 0);
    c->phi(i, j, -2) = c->phi(i, j, 0);
    c->phi(i, j, -3) = c->phi(i, j, 0);
}

// inlet_bounce_back_velocity_BC_before_odd / _after_odd (Boundary_multiphase_inlet.F90:6-102)
void inlet_velocity(oracle_ctx* c, bool after) {
    const oracle_params& P = c->p;
    for (int j = 1; j <= c->ny; ++j)
        for (int i = 1; i <= c->nx; ++i) {
            const int wi = c->W(i, j, 1);
            const double a = (double)(1 - wi), b = (double)wi;
            inlet_phi(c, i, j, wi);
            double tmp2 = c->WIN(i, j) * P.relaxation;
            const double tmp1 = tmp2 * P.sa_inject;
            tmp2 = tmp2 - tmp1;
            for (int col = 0; col < 2; ++col) {
                Arr3* A = col == 0 ? c->f : c->g;
                const double s = col == 0 ? tmp1 : tmp2;
                if (!after) {
                    A[5](i, j, 0) = (A[6](i, j, 1) + 6.0 * W_EQU_1 * s) * a + A[5](i, j, 0) * b;
                    A[11](i - 1, j, 0) = (A[14](i, j, 1) + 6.0 * W_EQU_2 * s) * a + A[11](i - 1, j, 0) * b;
                    A[12](i + 1, j, 0) = (A[13](i, j, 1) + 6.0 * W_EQU_2 * s) * a + A[12](i + 1, j, 0) * b;
                    A[15](i, j - 1, 0) = (A[18](i, j, 1) + 6.0 * W_EQU_2 * s) * a + A[15](i, j - 1, 0) * b;
                    A[16](i, j + 1, 0) = (A[17](i, j, 1) + 6.0 * W_EQU_2 * s) * a + A[16](i, j + 1, 0) * b;
                } else {
                    A[6](i, j, 1) = (A[5](i, j, 0) + 6.0 * W_EQU_1 * s) * a + A[6](i, j, 1) * b;
                    A[13](i, j, 1) = (A[12](i + 1, j, 0) + 6.0 * W_EQU_2 * s) * a + A[13](i, j, 1) * b;
                    A[14](i, j, 1) = (A[11](i - 1, j, 0) + 6.0 * W_EQU_2 * s) * a + A[14](i, j, 1) * b;
                    A[17](i, j, 1) = (A[16](i, j + 1, 0) + 6.0 * W_EQU_2 * s) * a + A[17](i, j, 1) * b;
                    A[18](i, j, 1) = (A[15](i, j - 1, 0) + 6.0 * W_EQU_2 * s) * a + A[18](i, j, 1) * b;
                }
            }
        }
}

// inlet_Zou_He_pressure_BC_before_odd / _after_odd (Boundary_multiphase_inlet.F90:115-294)
void inlet_pressure(oracle_ctx* c, bool after) {
    const oracle_params& P = c->p;
    const double c3 = 0.333333333333333333, c6 = 0.166666666666666667;
    for (int j = 1; j <= c->ny; ++j)
        for (int i = 1; i <= c->nx; ++i) {
            const int wi = c->W(i, j, 1);
            const double a = (double)(1 - wi), b = (double)wi;
            inlet_phi(c, i, j, wi);
            double r2 = P.rho_in;
            const double r1 = P.rho_in * P.sa_inject;
            r2 = r2 - r1;
            for (int col = 0; col < 2; ++col) {
                Arr3* A = col == 0 ? c->f : c->g;
                const double rr = col == 0 ? r1 : r2;
                if (!after) {
                    const double t = (rr - (A[0](i, j, 1) + A[1](i - 1, j, 1) + A[2](i + 1, j, 1) + A[3](i, j - 1, 1) +
                                            A[4](i, j + 1, 1) + A[7](i - 1, j - 1, 1) + A[8](i + 1, j - 1, 1) +
                                            A[9](i - 1, j + 1, 1) + A[10](i + 1, j + 1, 1) +
                                            2.0 * (A[6](i, j, 2) + A[14](i + 1, j, 2) + A[13](i - 1, j, 2) +
                                                   A[18](i, j + 1, 2) + A[17](i, j - 1, 2)))) * P.relaxation;
                    const double tnx = 0.5 * (A[1](i - 1, j, 1) + A[7](i - 1, j - 1, 1) + A[9](i - 1, j + 1, 1) -
                                              (A[2](i + 1, j, 1) + A[8](i + 1, j - 1, 1) + A[10](i + 1, j + 1, 1)));
                    const double tny = 0.5 * (A[3](i, j - 1, 1) + A[7](i - 1, j - 1, 1) + A[8](i + 1, j - 1, 1) -
                                              (A[4](i, j + 1, 1) + A[10](i + 1, j + 1, 1) + A[9](i - 1, j + 1, 1)));
                    A[5](i, j, 0) = (A[6](i, j, 2) + c3 * t) * a + A[5](i, j, 0) * b;
                    A[11](i - 1, j, 0) = (A[14](i + 1, j, 2) + c6 * t - tnx) * a + A[11](i - 1, j, 0) * b;
                    A[12](i + 1, j, 0) = (A[13](i - 1, j, 2) + c6 * t + tnx) * a + A[12](i + 1, j, 0) * b;
                    A[15](i, j - 1, 0) = (A[18](i, j + 1, 2) + c6 * t - tny) * a + A[15](i, j - 1, 0) * b;
                    A[16](i, j + 1, 0) = (A[17](i, j - 1, 2) + c6 * t + tny) * a + A[16](i, j + 1, 0) * b;
                } else {
                    const double t = (rr - (A[0](i, j, 1) + A[2](i, j, 1) + A[1](i, j, 1) + A[4](i, j, 1) + A[3](i, j, 1) +
                                            A[8](i, j, 1) + A[7](i, j, 1) + A[10](i, j, 1) + A[9](i, j, 1) +
                                            2.0 * (A[5](i, j, 1) + A[11](i, j, 1) + A[12](i, j, 1) + A[15](i, j, 1) +
                                                   A[16](i, j, 1)))) * P.relaxation;
                    const double tnx = 0.5 * (A[2](i, j, 1) + A[8](i, j, 1) + A[10](i, j, 1) - (A[1](i, j, 1) + A[7](i, j, 1) + A[9](i, j, 1)));
                    const double tny = 0.5 * (A[4](i, j, 1) + A[9](i, j, 1) + A[10](i, j, 1) - (A[3](i, j, 1) + A[8](i, j, 1) + A[7](i, j, 1)));
                    A[6](i, j, 1) = (A[5](i, j, 1) + c3 * t) * a + A[6](i, j, 1) * b;
                    A[13](i, j, 1) = (A[12](i, j, 1) + c6 * t + tnx) * a + A[13](i, j, 1) * b;
                    A[14](i, j, 1) = (A[11](i, j, 1) + c6 * t - tnx) * a + A[14](i, j, 1) * b;
                    A[17](i, j, 1) = (A[16](i, j, 1) + c6 * t + tny) * a + A[17](i, j, 1) * b;
                    A[18](i, j, 1) = (A[15](i, j, 1) + c6 * t - tny) * a + A[18](i, j, 1) * b;
                }
            }
        }
}

// outlet_convective_BC_before_odd / _after_odd (Boundary_multiphase_outlet.F90:7-127)
void outlet_convective(oracle_ctx* c, bool after) {
    const int nz = c->nz;
    const double u_convec = c->p.uin_avg;
    const double temp = 1.0 / (1.0 + u_convec);
    for (int j = 1; j <= c->ny; ++j)
        for (int i = 1; i <= c->nx; ++i) {
            const int wi = c->W(i, j, nz);
            const double a = (double)(1 - wi), b = (double)wi;
            c->phi(i, j, nz + 1) = ((c->PC(i, j) + u_convec * c->phi(i, j, nz)) * temp) * a + c->phi(i, j, nz + 1) * b;
            c->PC(i, j) = c->phi(i, j, nz + 1);
            c->phi(i, j, nz + 2) = c->phi(i, j, nz + 1);
            c->phi(i, j, nz + 3) = c->phi(i, j, nz + 1);
            c->phi(i, j, nz + 4) = c->phi(i, j, nz + 1);
            for (int col = 0; col < 2; ++col) {
                Arr3* A = col == 0 ? c->f : c->g;
                auto H = [&](int q) -> double& { return col == 0 ? c->FC(i, j, q) : c->GC(i, j, q); };
                if (!after) {
                    A[6](i, j, nz + 1) = ((H(6) + u_convec * A[6](i, j, nz)) * temp) * a + A[6](i, j, nz + 1) * b;
                    A[13](i - 1, j, nz + 1) = ((H(13) + u_convec * A[13](i - 1, j, nz)) * temp) * a + A[13](i - 1, j, nz + 1) * b;
                    A[14](i + 1, j, nz + 1) = ((H(14) + u_convec * A[14](i + 1, j, nz)) * temp) * a + A[14](i + 1, j, nz + 1) * b;
                    A[17](i, j - 1, nz + 1) = ((H(17) + u_convec * A[17](i, j - 1, nz)) * temp) * a + A[17](i, j - 1, nz + 1) * b;
                    A[18](i, j + 1, nz + 1) = ((H(18) + u_convec * A[18](i, j + 1, nz)) * temp) * a + A[18](i, j + 1, nz + 1) * b;
                } else {
                    A[5](i, j, nz) = ((H(6) + u_convec * A[5](i, j, nz - 1)) * temp) * a + A[5](i, j, nz) * b;
                    A[11](i, j, nz) = ((H(14) + u_convec * A[11](i, j, nz - 1)) * temp) * a + A[11](i, j, nz) * b;
                    A[12](i, j, nz) = ((H(13) + u_convec * A[12](i, j, nz - 1)) * temp) * a + A[12](i, j, nz) * b;
                    A[15](i, j, nz) = ((H(18) + u_convec * A[15](i, j, nz - 1)) * temp) * a + A[15](i, j, nz) * b;
                    A[16](i, j, nz) = ((H(17) + u_convec * A[16](i, j, nz - 1)) * temp) * a + A[16](i, j, nz) * b;
                }
            }
            for (int col = 0; col < 2; ++col) {
                Arr3* A = col == 0 ? c->f : c->g;
                auto H = [&](int q) -> double& { return col == 0 ? c->FC(i, j, q) : c->GC(i, j, q); };
                if (!after) {
                    H(6) = A[6](i, j, nz + 1);
                    H(13) = A[13](i - 1, j, nz + 1);
                    H(14) = A[14](i + 1, j, nz + 1);
                    H(17) = A[17](i, j - 1, nz + 1);
                    H(18) = A[18](i, j + 1, nz + 1);
                } else {
                    H(6) = A[5](i, j, nz);
                    H(14) = A[11](i, j, nz);
                    H(13) = A[12](i, j, nz);
                    H(18) = A[15](i, j, nz);
                    H(17) = A[16](i, j, nz);
                }
            }
        }
}

// outlet_Zou_He_pressure_BC_before_odd / _after_odd (Boundary_multiphase_outlet.F90:139-311)
void outlet_pressure(oracle_ctx* c, bool after) {
    const int nz = c->nz;
    const double c3 = 0.333333333333333333, c6 = 0.166666666666666667;
    const double rho_out = c->p.rho_out;
    for (int j = 1; j <= c->ny; ++j)
        for (int i = 1; i <= c->nx; ++i) {
            const int wi = c->W(i, j, nz);
            const double a = 1.0 - (double)wi, b = (double)wi;
            for (int l = 1; l <= 4; ++l) c->phi(i, j, nz + l) = c->phi(i, j, nz);
            Arr3* f = c->f;
            Arr3* g = c->g;
            double t1, tnx, tny;
            if (!after) {
                t1 = (f[0](i, j, nz) + f[1](i - 1, j, nz) + f[2](i + 1, j, nz) + f[3](i, j - 1, nz) + f[4](i, j + 1, nz) +
                      f[7](i - 1, j - 1, nz) + f[8](i + 1, j - 1, nz) + f[9](i - 1, j + 1, nz) + f[10](i + 1, j + 1, nz) +
                      2.0 * (f[5](i, j, nz - 1) + f[11](i - 1, j, nz - 1) + f[12](i + 1, j, nz - 1) + f[15](i, j - 1, nz - 1) +
                             f[16](i, j + 1, nz - 1)) +
                      g[0](i, j, nz) + g[1](i - 1, j, nz) + g[2](i + 1, j, nz) + g[3](i, j - 1, nz) + g[4](i, j + 1, nz) +
                      g[7](i - 1, j - 1, nz) + g[8](i + 1, j - 1, nz) + g[9](i - 1, j + 1, nz) + g[10](i + 1, j + 1, nz) +
                      2.0 * (g[5](i, j, nz - 1) + g[11](i - 1, j, nz - 1) + g[12](i + 1, j, nz - 1) + g[15](i, j - 1, nz - 1) +
                             g[16](i, j + 1, nz - 1))) - rho_out;
            } else {
                t1 = (f[0](i, j, nz) + f[2](i, j, nz) + f[1](i, j, nz) + f[4](i, j, nz) + f[3](i, j, nz) + f[8](i, j, nz) +
                      f[7](i, j, nz) + f[10](i, j, nz) + f[9](i, j, nz) +
                      2.0 * (f[6](i, j, nz) + f[14](i, j, nz) + f[13](i, j, nz) + f[18](i, j, nz) + f[17](i, j, nz)) +
                      g[0](i, j, nz) + g[2](i, j, nz) + g[1](i, j, nz) + g[4](i, j, nz) + g[3](i, j, nz) + g[8](i, j, nz) +
                      g[7](i, j, nz) + g[10](i, j, nz) + g[9](i, j, nz) +
                      2.0 * (g[6](i, j, nz) + g[14](i, j, nz) + g[13](i, j, nz) + g[18](i, j, nz) + g[17](i, j, nz))) - rho_out;
            }
            const double t2 = t1 * 0.5 * (1.0 - c->phi(i, j, nz));
            t1 = t1 - t2;
            for (int col = 0; col < 2; ++col) {
                Arr3* A = col == 0 ? f : g;
                const double s = col == 0 ? t1 : t2;
                if (!after) {
                    tnx = 0.5 * (A[1](i - 1, j, nz) + A[7](i - 1, j - 1, nz) + A[9](i - 1, j + 1, nz) -
                                 (A[2](i + 1, j, nz) + A[8](i + 1, j - 1, nz) + A[10](i + 1, j + 1, nz)));
                    tny = 0.5 * (A[3](i, j - 1, nz) + A[7](i - 1, j - 1, nz) + A[8](i + 1, j - 1, nz) -
                                 (A[4](i, j + 1, nz) + A[10](i + 1, j + 1, nz) + A[9](i - 1, j + 1, nz)));
                    A[6](i, j, nz + 1) = (A[5](i, j, nz - 1) - c3 * s) * a + A[6](i, j, nz + 1) * b;
                    A[13](i - 1, j, nz + 1) = (A[12](i + 1, j, nz - 1) - c6 * s - tnx) * a + A[13](i - 1, j, nz + 1) * b;
                    A[14](i + 1, j, nz + 1) = (A[11](i - 1, j, nz - 1) - c6 * s + tnx) * a + A[14](i + 1, j, nz + 1) * b;
                    A[17](i, j - 1, nz + 1) = (A[16](i, j + 1, nz - 1) - c6 * s - tny) * a + A[17](i, j - 1, nz + 1) * b;
                    A[18](i, j + 1, nz + 1) = (A[15](i, j - 1, nz - 1) - c6 * s + tny) * a + A[18](i, j + 1, nz + 1) * b;
                } else {
                    tnx = 0.5 * (A[2](i, j, nz) + A[8](i, j, nz) + A[10](i, j, nz) - (A[1](i, j, nz) + A[7](i, j, nz) + A[9](i, j, nz)));
                    tny = 0.5 * (A[4](i, j, nz) + A[10](i, j, nz) + A[9](i, j, nz) - (A[3](i, j, nz) + A[7](i, j, nz) + A[8](i, j, nz)));
                    A[5](i, j, nz) = (A[6](i, j, nz) - c3 * s) * a + A[5](i, j, nz) * b;
                    A[11](i, j, nz) = (A[14](i, j, nz) - c6 * s + tnx) * a + A[11](i, j, nz) * b;
                    A[12](i, j, nz) = (A[13](i, j, nz) - c6 * s - tnx) * a + A[12](i, j, nz) * b;
                    A[15](i, j, nz) = (A[18](i, j, nz) - c6 * s + tny) * a + A[15](i, j, nz) * b;
                    A[16](i, j, nz) = (A[17](i, j, nz) - c6 * s - tny) * a + A[16](i, j, nz) * b;
                }
            }
        }
}

// porous_plate_BC_before_odd / _after_odd (Boundary_multiphase_other.F90:8-184)
void porous_plate(oracle_ctx* c, bool after) {
    const int z = c->p.z_porous_plate;
    if (z < 1 || z > c->nz) return;
    Arr3* B = c->p.porous_plate_cmd == 1 ? c->f : c->g;   // reflected colour
    Arr3* T = c->p.porous_plate_cmd == 1 ? c->g : c->f;   // transmitted colour
    for (int j = 1; j <= c->ny; ++j)
        for (int i = 1; i <= c->nx; ++i) {
            if (!after) {
                B[6](i, j, z) = B[5](i, j, z - 1);
                B[13](i - 1, j, z) = B[12](i, j, z - 1);
                B[14](i + 1, j, z) = B[11](i, j, z - 1);
                B[17](i, j - 1, z) = B[16](i, j, z - 1);
                B[18](i, j + 1, z) = B[15](i, j, z - 1);
                B[5](i, j, z) = B[6](i, j, z + 1);
                B[12](i + 1, j, z) = B[13](i, j, z + 1);
                B[11](i - 1, j, z) = B[14](i, j, z + 1);
                B[16](i, j + 1, z) = B[17](i, j, z + 1);
                B[15](i, j - 1, z) = B[18](i, j, z + 1);
                T[6](i, j, z) = T[6](i, j, z + 1);
                T[13](i, j, z) = T[13](i, j, z + 1);
                T[14](i, j, z) = T[14](i, j, z + 1);
                T[17](i, j, z) = T[17](i, j, z + 1);
                T[18](i, j, z) = T[18](i, j, z + 1);
                T[5](i, j, z) = T[5](i, j, z - 1);
                T[12](i, j, z) = T[12](i, j, z - 1);
                T[11](i, j, z) = T[11](i, j, z - 1);
                T[16](i, j, z) = T[16](i, j, z - 1);
                T[15](i, j, z) = T[15](i, j, z - 1);
            } else {
                B[5](i, j, z - 1) = B[6](i, j, z);
                B[11](i, j, z - 1) = B[14](i + 1, j, z);
                B[12](i, j, z - 1) = B[13](i - 1, j, z);
                B[15](i, j, z - 1) = B[18](i, j + 1, z);
                B[16](i, j, z - 1) = B[17](i, j - 1, z);
                B[6](i, j, z + 1) = B[5](i, j, z);
                B[14](i, j, z + 1) = B[11](i - 1, j, z);
                B[13](i, j, z + 1) = B[12](i + 1, j, z);
                B[18](i, j, z + 1) = B[15](i, j - 1, z);
                B[17](i, j, z + 1) = B[16](i, j + 1, z);
                T[5](i, j, z - 1) = T[5](i, j, z);
                T[11](i, j, z - 1) = T[11](i, j, z);
                T[12](i, j, z - 1) = T[12](i, j, z);
                T[15](i, j, z - 1) = T[15](i, j, z);
                T[16](i, j, z - 1) = T[16](i, j, z);
                T[6](i, j, z + 1) = T[6](i, j, z);
                T[14](i, j, z + 1) = T[14](i, j, z);
                T[13](i, j, z + 1) = T[13](i, j, z);
                T[18](i, j, z + 1) = T[18](i, j, z);
                T[17](i, j, z + 1) = T[17](i, j, z);
            }
        }
}

// ISO4 derivative stencils with the reference's term order (Phase_gradient.F90:40-62, 122-192)
// ISO4 first derivatives at the element a points to (x stride 1, row stride px, plane stride pxy), the
// reference's term order (Phase_gradient.F90:40-62)
double ddx(const double* a, long px, long pxy) {
    return ISO4_1 * (a[1] - a[-1]) +
           ISO4_2 * (a[1 + px] - a[-1 - px] + a[1 - px] - a[-1 + px] + a[1 + pxy] - a[-1 - pxy] + a[1 - pxy] - a[-1 + pxy]);
}
double ddy(const double* a, long px, long pxy) {
    return ISO4_1 * (a[px] - a[-px]) +
           ISO4_2 * (a[1 + px] - a[-1 - px] + a[-1 + px] - a[1 - px] + a[px + pxy] - a[-px - pxy] + a[px - pxy] - a[-px + pxy]);
}
double ddz(const double* a, long px, long pxy) {
    return ISO4_1 * (a[pxy] - a[-pxy]) +
           ISO4_2 * (a[1 + pxy] - a[-1 - pxy] + a[-1 + pxy] - a[1 - pxy] + a[px + pxy] - a[-px - pxy] + a[-px + pxy] - a[px - pxy]);
}

// color_gradient (Phase_gradient.F90:5-204) with alter_color_gradient_solid_surface (:210-265)
void color_gradient(oracle_ctx* c) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const int ns = (int)c->s_nfl.size(), nf = (int)c->f_theta.size();
    // phi at solid boundary nodes: weighted mean over fluid neighbours (:12-30)
#pragma omp parallel for schedule(static)
    for (int t = 0; t < ns; ++t) {
        const int i = c->s_ijk[3 * t], j = c->s_ijk[3 * t + 1], k = c->s_ijk[3 * t + 2];
        double acc = 0.0;
        for (int m = 0; m < c->s_nfl[t]; ++m) {
            const int q = c->s_nbr[18 * t + m];
            acc = acc + c->phi(i + EX[q], j + EY[q], k + EZ[q]) * W_EQU[q];
        }
        c->phi(i, j, k) = acc / c->s_w[t];
    }
    // normalised colour gradient over [-1, n+2]^3 (:33-79)
    {
        const long ppx = c->phi.sx, ppxy = c->phi.pxy();
#pragma omp parallel for schedule(static)
        for (int k = -1; k <= nz + 2; ++k)
            for (int j = -1; j <= ny + 2; ++j) {
                const double* prow = &c->phi(-1, j, k);
                const int8_t* wrow = &c->W(-1, j, k);
                double* xr = &c->cn_x(-1, j, k);
                double* yr = &c->cn_y(-1, j, k);
                double* zr = &c->cn_z(-1, j, k);
                double* nr = &c->c_norm(-1, j, k);
                for (int i = 0; i < nx + 4; ++i) {
                    double cx = ddx(prow + i, ppx, ppxy), cy = ddy(prow + i, ppx, ppxy), cz = ddz(prow + i, ppx, ppxy);
                    double cn = std::sqrt(cx * cx + cy * cy + cz * cz);
                    if (cn < 1e-6 || wrow[i] == 1) {
                        cx = cy = cz = cn = 0.0;
                    } else {
                        cx = cx / cn;
                        cy = cy / cn;
                        cz = cz / cn;
                    }
                    xr[i] = cx;
                    yr[i] = cy;
                    zr[i] = cz;
                    nr[i] = cn;
                }
            }
    }
    // wetting: rotate the normal at fluid boundary nodes to the contact angle (:210-265)
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nf; ++t) {
        const int i = c->f_ijk[3 * t], j = c->f_ijk[3 * t + 1], k = c->f_ijk[3 * t + 2];
        if (!(c->c_norm(i, j, k) > 1e-6)) continue;
        const double th = c->f_theta[t];
        const double nwx = c->f_nw[3 * t], nwy = c->f_nw[3 * t + 1], nwz = c->f_nw[3 * t + 2];
        const double x0 = c->cn_x(i, j, k), y0 = c->cn_y(i, j, k), z0 = c->cn_z(i, j, k);
        const double tcos = std::cos(th);
        const double t1 = nwx * x0 + nwy * y0 + nwz * z0;
        const double t2 = 1.0 / std::sqrt(1.0 - t1 * t1);
        const double coe1 = std::sin(th) * t1 * t2;
        const double coe2 = std::sin(th) * t2;
        const double xp = (tcos - coe1) * nwx + coe2 * x0, yp = (tcos - coe1) * nwy + coe2 * y0, zp = (tcos - coe1) * nwz + coe2 * z0;
        const double xm = (tcos + coe1) * nwx - coe2 * x0, ym = (tcos + coe1) * nwy - coe2 * y0, zm = (tcos + coe1) * nwz - coe2 * z0;
        const double dP = (xp - x0) * (xp - x0) + (yp - y0) * (yp - y0) + (zp - z0) * (zp - z0);
        const double dM = (xm - x0) * (xm - x0) + (ym - y0) * (ym - y0) + (zm - z0) * (zm - z0);
        if (dP <= dM) {
            c->cn_x(i, j, k) = xp;
            c->cn_y(i, j, k) = yp;
            c->cn_z(i, j, k) = zp;
        } else {
            c->cn_x(i, j, k) = xm;
            c->cn_y(i, j, k) = ym;
            c->cn_z(i, j, k) = zm;
        }
    }
    // normal at solid boundary nodes within one halo layer (:84-110)
#pragma omp parallel for schedule(static)
    for (int t = 0; t < ns; ++t) {
        const int i = c->s_ijk[3 * t], j = c->s_ijk[3 * t + 1], k = c->s_ijk[3 * t + 2];
        if (!(i >= 0 && i <= nx + 1 && j >= 0 && j <= ny + 1 && k >= 0 && k <= nz + 1)) continue;
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int m = 0; m < c->s_nfl[t]; ++m) {
            const int q = c->s_nbr[18 * t + m];
            ax = ax + c->cn_x(i + EX[q], j + EY[q], k + EZ[q]) * W_EQU[q];
            ay = ay + c->cn_y(i + EX[q], j + EY[q], k + EZ[q]) * W_EQU[q];
            az = az + c->cn_z(i + EX[q], j + EY[q], k + EZ[q]) * W_EQU[q];
        }
        c->cn_x(i, j, k) = ax / c->s_w[t];
        c->cn_y(i, j, k) = ay / c->s_w[t];
        c->cn_z(i, j, k) = az / c->s_w[t];
    }
    // curvature of the interface (:112-201)
    {
        const long npx = c->cn_x.sx, npxy = c->cn_x.pxy();
#pragma omp parallel for schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j) {
                const double* X = &c->cn_x(1, j, k);
                const double* Y = &c->cn_y(1, j, k);
                const double* Z = &c->cn_z(1, j, k);
                double* cr = &c->curv(1, j, k);
                for (int i = 0; i < nx; ++i) {
                    const double kxx = ddx(X + i, npx, npxy), kyy = ddy(Y + i, npx, npxy), kzz = ddz(Z + i, npx, npxy);
                    const double kxy = ddy(X + i, npx, npxy), kxz = ddz(X + i, npx, npxy);
                    const double kyx = ddx(Y + i, npx, npxy), kyz = ddz(Y + i, npx, npxy);
                    const double kzx = ddx(Z + i, npx, npxy), kzy = ddy(Z + i, npx, npxy);
                    const double x = X[i], y = Y[i], z = Z[i];
                    cr[i] = (x * x - 1.0) * kxx + (y * y - 1.0) * kyy + (z * z - 1.0) * kzz + x * y * (kxy + kyx) +
                            x * z * (kxz + kzx) + y * z * (kzy + kyz);
                }
            }
    }
}

// main_iteration_kernel (Main_multiphase.F90:341-486) for one rank without neighbours
void step(oracle_ctx* c, int ntime) {
    const oracle_params& P = c->p;
    const bool odd = (ntime % 2) != 0;
    lattice_update(c, odd);
    if (P.jper == 1 || P.kper == 1) {
        exchange_pdf(c, odd);
        exchange_phi(c);
    }
    if (P.kper == 0 && P.wall_z_min == 0 && P.wall_z_max == 0) {
        if (P.inlet_bc == 1) inlet_velocity(c, odd);
        else if (P.inlet_bc == 2) inlet_pressure(c, odd);
        if (P.outlet_bc == 1) outlet_convective(c, odd);
        else if (P.outlet_bc == 2) outlet_pressure(c, odd);
    }
    if (P.porous_plate_cmd != 0) porous_plate(c, odd);
    color_gradient(c);
}

void copy_in(Arr3& A, int h, const double* src) { std::memcpy(A.v.data(), src, A.v.size() * sizeof(double)); (void)h; }
void copy_out(Arr3& A, double* dst) { std::memcpy(dst, A.v.data(), A.v.size() * sizeof(double)); }

}  // namespace

extern "C" int oracle_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

extern "C" int oracle_create(const oracle_params* p, oracle_ctx** out) {
    if (!p || !out || p->nx < 3 || p->ny < 3 || p->nz < 3) return 1;
    oracle_ctx* c = new (std::nothrow) oracle_ctx();
    if (!c) return 4;
    c->p = *p;
    c->nx = p->nx;
    c->ny = p->ny;
    c->nz = p->nz;
    for (int q = 0; q < 19; ++q) {
        c->f[q].init(c->nx, c->ny, c->nz, 1);
        c->g[q].init(c->nx, c->ny, c->nz, 1);
    }
    c->phi.init(c->nx, c->ny, c->nz, 4);
    c->phi_old.init(c->nx, c->ny, c->nz, 4);
    for (Arr3* a : {&c->cn_x, &c->cn_y, &c->cn_z, &c->c_norm}) a->init(c->nx, c->ny, c->nz, 2);
    for (Arr3* a : {&c->curv, &c->u, &c->v, &c->w, &c->rho}) a->init(c->nx, c->ny, c->nz, 1);
    c->walls.assign((size_t)(c->nx + 4) * (c->ny + 4) * (c->nz + 4), 1);
    const size_t plane = (size_t)(c->nx + 2) * (c->ny + 2);
    c->w_in.assign(plane, 0.0);
    c->phi_convec.assign(plane, 0.0);
    c->f_convec.assign(plane * 19, 0.0);
    c->g_convec.assign(plane * 19, 0.0);
    *out = c;
    return 0;
}

extern "C" void oracle_destroy(oracle_ctx* c) { delete c; }

extern "C" int oracle_upload_geometry(oracle_ctx* c, const int8_t* walls_h2, int32_t ns, const int32_t* s_ijk,
                                      const int32_t* s_nfl, const int32_t* s_nbr, const double* s_w, int32_t nf,
                                      const int32_t* f_ijk, const double* f_nw, const double* f_theta, const double* w_in) {
    if (!c || !walls_h2) return 1;
    std::memcpy(c->walls.data(), walls_h2, c->walls.size());
    c->s_ijk.assign(s_ijk, s_ijk + 3 * (size_t)ns);
    c->s_nfl.assign(s_nfl, s_nfl + ns);
    c->s_nbr.assign(s_nbr, s_nbr + 18 * (size_t)ns);
    c->s_w.assign(s_w, s_w + ns);
    c->f_ijk.assign(f_ijk, f_ijk + 3 * (size_t)nf);
    c->f_nw.assign(f_nw, f_nw + 3 * (size_t)nf);
    c->f_theta.assign(f_theta, f_theta + nf);
    if (w_in) std::memcpy(c->w_in.data(), w_in, c->w_in.size() * sizeof(double));
    return 0;
}

extern "C" int oracle_upload_state(oracle_ctx* c, const double* const* f, const double* const* g, const double* phi,
                                   const double* f_convec, const double* g_convec, const double* phi_convec) {
    if (!c || !f || !g || !phi) return 1;
    for (int q = 0; q < 19; ++q) {
        copy_in(c->f[q], 1, f[q]);
        copy_in(c->g[q], 1, g[q]);
    }
    copy_in(c->phi, 4, phi);
    if (f_convec) std::memcpy(c->f_convec.data(), f_convec, c->f_convec.size() * sizeof(double));
    if (g_convec) std::memcpy(c->g_convec.data(), g_convec, c->g_convec.size() * sizeof(double));
    if (phi_convec) std::memcpy(c->phi_convec.data(), phi_convec, c->phi_convec.size() * sizeof(double));
    std::fill(c->phi_old.v.begin(), c->phi_old.v.end(), 0.0);   // phi_old reads as zero until the first update
    return 0;
}

extern "C" int oracle_download_state(oracle_ctx* c, double* const* f, double* const* g, double* phi, double* f_convec,
                                     double* g_convec, double* phi_convec) {
    if (!c) return 1;
    for (int q = 0; q < 19; ++q) {
        if (f && f[q]) copy_out(c->f[q], f[q]);
        if (g && g[q]) copy_out(c->g[q], g[q]);
    }
    if (phi) copy_out(c->phi, phi);
    if (f_convec) std::memcpy(f_convec, c->f_convec.data(), c->f_convec.size() * sizeof(double));
    if (g_convec) std::memcpy(g_convec, c->g_convec.data(), c->g_convec.size() * sizeof(double));
    if (phi_convec) std::memcpy(phi_convec, c->phi_convec.data(), c->phi_convec.size() * sizeof(double));
    return 0;
}

extern "C" int oracle_step(oracle_ctx* c, int32_t ntime) {
    if (!c) return 1;
    step(c, ntime);
    return 0;
}

extern "C" int oracle_run(oracle_ctx* c, int32_t ntime0, int32_t nsteps) {
    if (!c) return 1;
    for (int s = 0; s < nsteps; ++s) step(c, ntime0 + s);
    return 0;
}

extern "C" int oracle_color_gradient(oracle_ctx* c) {
    if (!c) return 1;
    color_gradient(c);
    return 0;
}

// compute_macro_vars (Misc.F90:372-430)
extern "C" int oracle_compute_macro(oracle_ctx* c) {
    if (!c) return 1;
    const double gamma = c->p.gamma, force_z = c->p.force_z;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) {
                const int wi = c->W(i, j, k);
                const double a = (double)(1 - wi);
                double ft[19];
                for (int q = 0; q < 19; ++q) ft[q] = c->f[q](i, j, k) + c->g[q](i, j, k);
                double s = ft[0];
                for (int q = 1; q < 19; ++q) s = s + ft[q];
                c->rho(i, j, k) = s * a;
                const double tmp = 0.5 * gamma * c->curv(i, j, k) * c->c_norm(i, j, k);
                const double fx = tmp * c->cn_x(i, j, k), fy = tmp * c->cn_y(i, j, k), fz = tmp * c->cn_z(i, j, k) + force_z;
                c->u(i, j, k) = (ft[1] - ft[2] + ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14] - 0.5 * fx) * a;
                c->v(i, j, k) = (ft[3] - ft[4] + ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18] - 0.5 * fy) * a;
                c->w(i, j, k) = (ft[5] - ft[6] + ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18] - 0.5 * fz) * a;
                c->phi(i, j, k) = 0.0 * (double)wi + c->phi(i, j, k) * a;
            }
    return 0;
}

// monitor sums (Monitor.F90:32-85 monitor, :305-322 steady phase field, :384-423 capillary pressure,
// :485-494 breakthrough, :526-538 cal_saturation), serial in the reference's loop order
extern "C" int oracle_monitor(oracle_ctx* c, double* layers, oracle_monitor_sums* sums, int32_t update_phi_old) {
    if (!c) return 1;
    oracle_compute_macro(c);
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    oracle_monitor_sums s{};
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const int wi = c->W(i, j, k);
                const double a = (double)(1 - wi);
                const double uu = c->u(i, j, k), vv = c->v(i, j, k), ww = c->w(i, j, k), ph = c->phi(i, j, k);
                const double temp = (uu * uu + vv * vv + ww * ww) * a;
                if (s.umax2 < temp) s.umax2 = temp;
                if (ph > 0.999) s.usq1 = s.usq1 + temp;
                else if (ph < -0.999) s.usq2 = s.usq2 + temp;
                {
                    const double d = std::fabs(ph - c->phi_old(i, j, k)) * a;
                    if (s.dphi_max < d) s.dphi_max = d;
                }
                if (wi == 0) {
                    if (ph < -0.99) { s.pre_w = s.pre_w + c->rho(i, j, k); s.n_w++; }
                    if (ph > 0.99) { s.pre_nw = s.pre_nw + c->rho(i, j, k); s.n_nw++; }
                }
                s.vol1_full = s.vol1_full + 0.5 * (1.0 + ph) * a;
                s.vol2_full = s.vol2_full + 0.5 * (1.0 - ph) * a;
            }
    for (int k = 1; k <= nz; ++k) {
        double t1 = 0, t2 = 0, pre = 0, t3 = 0, t4 = 0, t5 = 0, t6 = 0;
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double a = (double)(1 - c->W(i, j, k));
                const double ph = c->phi(i, j, k), r = c->rho(i, j, k), ww = c->w(i, j, k);
                t3 = t3 + 0.5 * (1.0 + ph) * a;
                t4 = t4 + 0.5 * (1.0 - ph) * a;
                t5 = t5 + r * 0.5 * (1.0 + ph) * a;
                t6 = t6 + r * 0.5 * (1.0 - ph) * a;
                t1 = t1 + ww * 0.5 * (1.0 + ph) * a;
                t2 = t2 + ww * 0.5 * (1.0 - ph) * a;
                pre = pre + r * a;
            }
        if (layers) {
            layers[0 * nz + k - 1] = t1;
            layers[1 * nz + k - 1] = t2;
            layers[2 * nz + k - 1] = t3;
            layers[3 * nz + k - 1] = t4;
            layers[4 * nz + k - 1] = t5;
            layers[5 * nz + k - 1] = t6;
            layers[6 * nz + k - 1] = pre;
        }
    }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
            if (c->W(i, j, nz - 1) == 0 && c->phi(i, j, nz - 1) > 0.0) s.outlet_phase1++;
    if (sums) *sums = s;
    if (update_phi_old) {
        c->phi_old.v = c->phi.v;
    }
    return 0;
}

extern "C" int oracle_download_macro(oracle_ctx* c, double* u, double* v, double* w, double* rho, double* phi) {
    if (!c) return 1;
    if (u) copy_out(c->u, u);
    if (v) copy_out(c->v, v);
    if (w) copy_out(c->w, w);
    if (rho) copy_out(c->rho, rho);
    if (phi) copy_out(c->phi, phi);
    return 0;
}

extern "C" int oracle_download_gradient(oracle_ctx* c, double* cn_x, double* cn_y, double* cn_z, double* c_norm, double* curv) {
    if (!c) return 1;
    if (cn_x) copy_out(c->cn_x, cn_x);
    if (cn_y) copy_out(c->cn_y, cn_y);
    if (cn_z) copy_out(c->cn_z, cn_z);
    if (c_norm) copy_out(c->c_norm, c_norm);
    if (curv) copy_out(c->curv, curv);
    return 0;
}
