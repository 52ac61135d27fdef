// CPU restatement of the reference's single-phase time step (checker only; see oracle.h).
//
// singlephase_3D/0.src: kernel_odd / kernel_even (Kernel.F90:5-400), the MPI self-exchange of a periodic
// single rank (Mpi.F90:73-448), the open boundaries (Boundary.F90:5-389), compute_macro_vars
// (Misc.F90:368-423) and the monitor sums (Monitor.F90:20-48).  Arrays keep the reference's shapes:
// populations, u, v, w, rho over (0:nx+1, 0:ny+1, 0:nz+1), walls over (-1:nx+2, ...), w_in and f_convec_bc
// over (0:nx+1, 0:ny+1[, 0:18]).  Loops run in the reference's order; OpenMP splits only the lattice update
// (each slot is written by exactly one node), reductions run serially.
#include <omp.h>

#include <cstring>
#include <new>
#include <vector>

#include "oracle.h"

namespace {

const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
const int OPC[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};   // Module.F90:80
const double MRT_COEF1 = 1.0 / 19.0, MRT_COEF2 = 1.0 / 2394.0, MRT_COEF3 = 1.0 / 252.0, MRT_COEF4 = 1.0 / 72.0;
const double MRT_E2_COEF1 = 0.0, MRT_E2_COEF2 = -475.0 / 63.0, MRT_OMEGA_XX = 0.0;   // Module.F90:82-88
const double W_EQU_1 = 1.0 / 18.0, W_EQU_2 = 1.0 / 36.0;                               // Module.F90:96-98

struct Field {   // (0:nx+1, 0:ny+1, 0:nz+1), x fastest
    int sx = 0, sy = 0;
    std::vector<double> v;
    void init(int nx, int ny, int nz) {
        sx = nx + 2;
        sy = ny + 2;
        v.assign((size_t)sx * sy * (nz + 2), 0.0);
    }
    double& operator()(int i, int j, int k) { return v[(size_t)i + (size_t)j * sx + (size_t)k * sx * sy]; }
};

}  // namespace

struct oracle_sp_ctx {
    oracle_sp_params p;
    int nx, ny, nz;
    double s_e, s_e2, s_q, s_nu, s_pi, s_t;
    Field f[19], u, v, w, rho;
    std::vector<int8_t> walls;                 // (-1:nx+2, ...)
    std::vector<double> w_in, f_convec;        // (0:nx+1, 0:ny+1[, 0:18])
    int8_t& W(int i, int j, int k) {
        return walls[(size_t)(i + 1) + (size_t)(j + 1) * (nx + 4) + (size_t)(k + 1) * (nx + 4) * (ny + 4)];
    }
    double& WIN(int i, int j) { return w_in[(size_t)i + (size_t)j * (nx + 2)]; }
    double& FC(int i, int j, int q) { return f_convec[(size_t)i + (size_t)j * (nx + 2) + (size_t)q * (nx + 2) * (ny + 2)]; }
};

namespace {

// MRT collision of one node (Kernel.F90:56-165), ft[] in, ft[] out
void sp_collide(const oracle_sp_ctx* c, double* ft) {
    const double fx = 0.0, fy = 0.0, fz = c->p.force_z;
    const double s_e = c->s_e, s_e2 = c->s_e2, s_q = c->s_q, s_nu = c->s_nu, s_pi = c->s_pi, s_t = c->s_t;
    const double den = ft[0] + ft[1] + ft[2] + ft[3] + ft[4] + ft[5] + ft[6] + ft[7] + ft[8] + ft[9] + ft[10] + ft[11] +
                       ft[12] + ft[13] + ft[14] + ft[15] + ft[16] + ft[17] + ft[18];
    const double ux = ft[1] - ft[2] + ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14] + 0.5 * fx;
    const double uy = ft[3] - ft[4] + ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18] + 0.5 * fy;
    const double uz = ft[5] - ft[6] + ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18] + 0.5 * fz;
    const double u2 = ux * ux + uy * uy + uz * uz;
    double s1 = ft[1] + ft[2] + ft[3] + ft[4] + ft[5] + ft[6];
    double s2 = ft[7] + ft[8] + ft[9] + ft[10] + ft[11] + ft[12] + ft[13] + ft[14] + ft[15] + ft[16] + ft[17] + ft[18];
    double s3 = ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14];
    double s4 = ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18];
    double s5 = ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18];
    double s6 = 2.0 * (ft[1] + ft[2]) - ft[3] - ft[4] - ft[5] - ft[6];
    double s7 = ft[7] + ft[8] + ft[9] + ft[10] + ft[11] + ft[12] + ft[13] + ft[14] - 2.0 * (ft[15] + ft[16] + ft[17] + ft[18]);
    double s8 = ft[3] + ft[4] - ft[5] - ft[6];
    double s9 = ft[7] + ft[8] + ft[9] + ft[10] - ft[11] - ft[12] - ft[13] - ft[14];
    // moments m0..m18 in the reference's order
    double m[19];
    m[0] = den;
    m[1] = -30.0 * ft[0] - 11.0 * s1 + 8.0 * s2;
    m[2] = 12.0 * ft[0] - 4.0 * s1 + s2;
    m[3] = ft[1] - ft[2] + s3;
    m[4] = -4.0 * (ft[1] - ft[2]) + s3;
    m[5] = ft[3] - ft[4] + s4;
    m[6] = -4.0 * (ft[3] - ft[4]) + s4;
    m[7] = ft[5] - ft[6] + s5;
    m[8] = -4.0 * (ft[5] - ft[6]) + s5;
    m[9] = s6 + s7;
    m[10] = -2.0 * s6 + s7;
    m[11] = s8 + s9;
    m[12] = -2.0 * s8 + s9;
    m[13] = ft[7] - ft[8] - ft[9] + ft[10];
    m[14] = ft[15] - ft[16] - ft[17] + ft[18];
    m[15] = ft[11] - ft[12] - ft[13] + ft[14];
    m[16] = ft[7] - ft[8] + ft[9] - ft[10] - ft[11] + ft[12] - ft[13] + ft[14];
    m[17] = -ft[7] - ft[8] + ft[9] + ft[10] + ft[15] - ft[16] + ft[17] - ft[18];
    m[18] = ft[11] + ft[12] - ft[13] - ft[14] - ft[15] - ft[16] + ft[17] + ft[18];
    const double c23 = 0.666666666666666667, c13 = 0.333333333333333333;
    m[1] = m[1] - s_e * (m[1] - (-11.0 * den + 19.0 * u2)) + (38.0 - 19.0 * s_e) * (fx * ux + fy * uy + fz * uz);
    m[2] = m[2] - s_e2 * (m[2] - (MRT_E2_COEF1 * den + MRT_E2_COEF2 * u2)) + (-11.0 + 5.5 * s_e2) * (fx * ux + fy * uy + fz * uz);
    m[3] = m[3] + fx;
    m[4] = m[4] - s_q * (m[4] - (-c23 * ux)) + (-c23 + c13 * s_q) * fx;
    m[5] = m[5] + fy;
    m[6] = m[6] - s_q * (m[6] - (-c23 * uy)) + (-c23 + c13 * s_q) * fy;
    m[7] = m[7] + fz;
    m[8] = m[8] - s_q * (m[8] - (-c23 * uz)) + (-c23 + c13 * s_q) * fz;
    m[9] = m[9] - s_nu * (m[9] - (3.0 * ux * ux - u2)) + (2.0 - s_nu) * (2.0 * fx * ux - fy * uy - fz * uz);
    m[10] = m[10] - s_pi * (m[10] - MRT_OMEGA_XX * (3.0 * ux * ux - u2)) + (1.0 - 0.5 * s_pi) * (-2.0 * fx * ux + fy * uy + fz * uz);
    m[11] = m[11] - s_nu * (m[11] - (uy * uy - uz * uz)) + (2.0 - s_nu) * (fy * uy - fz * uz);
    m[12] = m[12] - s_pi * (m[12] - MRT_OMEGA_XX * (uy * uy - uz * uz)) + (1.0 - 0.5 * s_pi) * (-fy * uy + fz * uz);
    m[13] = m[13] - s_nu * (m[13] - (ux * uy)) + (1.0 - 0.5 * s_nu) * (fx * uy + fy * ux);
    m[14] = m[14] - s_nu * (m[14] - (uy * uz)) + (1.0 - 0.5 * s_nu) * (fy * uz + fz * uy);
    m[15] = m[15] - s_nu * (m[15] - (ux * uz)) + (1.0 - 0.5 * s_nu) * (fx * uz + fz * ux);
    m[16] = m[16] - s_t * (m[16]);
    m[17] = m[17] - s_t * (m[17]);
    m[18] = m[18] - s_t * (m[18]);
    // back to populations (Kernel.F90:118-165)
    m[0] = MRT_COEF1 * m[0];
    m[1] = MRT_COEF2 * m[1];
    m[2] = MRT_COEF3 * m[2];
    m[3] = 0.1 * m[3];
    m[4] = 0.025 * m[4];
    m[5] = 0.1 * m[5];
    m[6] = 0.025 * m[6];
    m[7] = 0.1 * m[7];
    m[8] = 0.025 * m[8];
    m[9] = 2.0 * MRT_COEF4 * m[9];
    m[10] = MRT_COEF4 * m[10];
    m[11] = 6.0 * MRT_COEF4 * m[11];
    m[12] = 3.0 * MRT_COEF4 * m[12];
    m[13] = 0.25 * m[13];
    m[14] = 0.25 * m[14];
    m[15] = 0.25 * m[15];
    m[16] = 0.125 * m[16];
    m[17] = 0.125 * m[17];
    m[18] = 0.125 * m[18];
    s1 = m[0] - 11.0 * m[1] - 4.0 * m[2];
    s2 = 2.0 * m[9] - 4.0 * m[10];
    s3 = m[11] - 2.0 * m[12];
    s4 = m[0] + 8.0 * m[1] + m[2];
    s5 = m[3] + m[4];
    s6 = m[5] + m[6];
    s7 = m[7] + m[8];
    s8 = m[9] + m[10];
    s9 = m[11] + m[12];
    ft[0] = m[0] - 30.0 * m[1] + 12.0 * m[2];
    ft[1] = s1 + m[3] - 4.0 * m[4] + s2;
    ft[2] = s1 - m[3] + 4.0 * m[4] + s2;
    ft[3] = s1 + m[5] - 4.0 * m[6] - 0.5 * s2 + s3;
    ft[4] = s1 - m[5] + 4.0 * m[6] - 0.5 * s2 + s3;
    ft[5] = s1 + m[7] - 4.0 * m[8] - 0.5 * s2 - s3;
    ft[6] = s1 - m[7] + 4.0 * m[8] - 0.5 * s2 - s3;
    ft[7] = s4 + s5 + s6 + s8 + s9 + m[13] + m[16] - m[17];
    ft[8] = s4 - s5 + s6 + s8 + s9 - m[13] - m[16] - m[17];
    ft[9] = s4 + s5 - s6 + s8 + s9 - m[13] + m[16] + m[17];
    ft[10] = s4 - s5 - s6 + s8 + s9 + m[13] - m[16] + m[17];
    ft[11] = s4 + s5 + s7 + s8 - s9 + m[15] - m[16] + m[18];
    ft[12] = s4 - s5 + s7 + s8 - s9 - m[15] + m[16] + m[18];
    ft[13] = s4 + s5 - s7 + s8 - s9 - m[15] - m[16] - m[18];
    ft[14] = s4 - s5 - s7 + s8 - s9 + m[15] + m[16] - m[18];
    ft[15] = s4 + s6 + s7 - s8 * 2.0 + m[14] + m[17] - m[18];
    ft[16] = s4 - s6 + s7 - s8 * 2.0 - m[14] - m[17] - m[18];
    ft[17] = s4 + s6 - s7 - s8 * 2.0 - m[14] + m[17] + m[18];
    ft[18] = s4 - s6 - s7 - s8 * 2.0 + m[14] - m[17] + m[18];
}

// kernel_odd / kernel_even (singlephase_3D/0.src/Kernel.F90:5-400); the 19 arrays share one shape, so one flat
// index per node addresses all of them
template <bool ODD>
void sp_lattice_update_t(oracle_sp_ctx* c) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    double* F[19];
    long off[19];
    const long sx = c->f[0].sx, sxy = (long)c->f[0].sx * c->f[0].sy;
    for (int q = 0; q < 19; ++q) {
        F[q] = c->f[q].v.data();
        off[q] = EX[q] + EY[q] * sx + EZ[q] * sxy;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j) {
            const long n0 = 1 + (long)j * sx + (long)k * sxy;
            const int8_t* wrow = &c->W(1, j, k);
            for (int i = 0; i < nx; ++i) {
                if (wrow[i] != 0) continue;
                const long n = n0 + i;
                double ft[19];
                ft[0] = F[0][n];
                for (int q = 1; q < 19; ++q) ft[q] = ODD ? F[q][n - off[q]] : F[OPC[q]][n];
                sp_collide(c, ft);
                F[0][n] = ft[0];
                for (int q = 1; q < 19; ++q) {
                    if (ODD) F[OPC[q]][n + off[q]] = ft[q];
                    else F[q][n] = ft[q];
                }
            }
        }
}
void sp_lattice_update(oracle_sp_ctx* c, bool odd) {
    if (odd) sp_lattice_update_t<true>(c);
    else sp_lattice_update_t<false>(c);
}

// periodic exchange of a single rank with itself (Mpi.F90:73-448)
void sp_exchange(oracle_sp_ctx* c, bool odd) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const bool my = c->p.jper == 1, mz = c->p.kper == 1;
    const int zP[5] = {5, 11, 12, 15, 16}, zM[5] = {6, 14, 13, 18, 17};
    const int yP[5] = {3, 7, 8, 15, 17}, yM[5] = {4, 10, 9, 16, 18};
    Field* A = c->f;
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

// velocity inlet by bounce-back (Boundary.F90:5-70)
void sp_inlet_velocity(oracle_sp_ctx* c, bool after) {
    Field* f = c->f;
    for (int j = 1; j <= c->ny; ++j)
        for (int i = 1; i <= c->nx; ++i) {
            const int wi = c->W(i, j, 1);
            const double a = (double)(1 - wi), b = (double)wi;
            const double t = c->WIN(i, j) * c->p.relaxation;
            if (!after) {
                f[5](i, j, 0) = (f[6](i, j, 1) + 6.0 * W_EQU_1 * t) * a + f[5](i, j, 0) * b;
                f[11](i - 1, j, 0) = (f[14](i, j, 1) + 6.0 * W_EQU_2 * t) * a + f[11](i - 1, j, 0) * b;
                f[12](i + 1, j, 0) = (f[13](i, j, 1) + 6.0 * W_EQU_2 * t) * a + f[12](i + 1, j, 0) * b;
                f[15](i, j - 1, 0) = (f[18](i, j, 1) + 6.0 * W_EQU_2 * t) * a + f[15](i, j - 1, 0) * b;
                f[16](i, j + 1, 0) = (f[17](i, j, 1) + 6.0 * W_EQU_2 * t) * a + f[16](i, j + 1, 0) * b;
            } else {
                f[6](i, j, 1) = (f[5](i, j, 0) + 6.0 * W_EQU_1 * t) * a + f[6](i, j, 1) * b;
                f[13](i, j, 1) = (f[12](i + 1, j, 0) + 6.0 * W_EQU_2 * t) * a + f[13](i, j, 1) * b;
                f[14](i, j, 1) = (f[11](i - 1, j, 0) + 6.0 * W_EQU_2 * t) * a + f[14](i, j, 1) * b;
                f[17](i, j, 1) = (f[16](i, j + 1, 0) + 6.0 * W_EQU_2 * t) * a + f[17](i, j, 1) * b;
                f[18](i, j, 1) = (f[15](i, j - 1, 0) + 6.0 * W_EQU_2 * t) * a + f[18](i, j, 1) * b;
            }
        }
}

// Zou-He pressure inlet (Boundary.F90:76-183)
void sp_inlet_pressure(oracle_sp_ctx* c, bool after) {
    Field* f = c->f;
    const double rin = c->p.rho_in, rel = c->p.relaxation, o3 = 0.333333333333333333, o6 = 0.166666666666666667;
    for (int j = 1; j <= c->ny; ++j)
        for (int i = 1; i <= c->nx; ++i) {
            const int wi = c->W(i, j, 1);
            const double a = (double)(1 - wi), b = (double)wi;
            if (!after) {
                const double t = (rin - (f[0](i, j, 1) + f[1](i - 1, j, 1) + f[2](i + 1, j, 1) + f[3](i, j - 1, 1) +
                                         f[4](i, j + 1, 1) + f[7](i - 1, j - 1, 1) + f[8](i + 1, j - 1, 1) +
                                         f[9](i - 1, j + 1, 1) + f[10](i + 1, j + 1, 1) +
                                         2.0 * (f[6](i, j, 2) + f[14](i + 1, j, 2) + f[13](i - 1, j, 2) + f[18](i, j + 1, 2) +
                                                f[17](i, j - 1, 2)))) * rel;
                const double tnx = 0.5 * (f[1](i - 1, j, 1) + f[7](i - 1, j - 1, 1) + f[9](i - 1, j + 1, 1) -
                                          (f[2](i + 1, j, 1) + f[8](i + 1, j - 1, 1) + f[10](i + 1, j + 1, 1)));
                const double tny = 0.5 * (f[3](i, j - 1, 1) + f[7](i - 1, j - 1, 1) + f[8](i + 1, j - 1, 1) -
                                          (f[4](i, j + 1, 1) + f[10](i + 1, j + 1, 1) + f[9](i - 1, j + 1, 1)));
                f[5](i, j, 0) = (f[6](i, j, 2) + o3 * t) * a + f[5](i, j, 0) * b;
                f[11](i - 1, j, 0) = (f[14](i + 1, j, 2) + o6 * t - tnx) * a + f[11](i - 1, j, 0) * b;
                f[12](i + 1, j, 0) = (f[13](i - 1, j, 2) + o6 * t + tnx) * a + f[12](i + 1, j, 0) * b;
                f[15](i, j - 1, 0) = (f[18](i, j + 1, 2) + o6 * t - tny) * a + f[15](i, j - 1, 0) * b;
                f[16](i, j + 1, 0) = (f[17](i, j - 1, 2) + o6 * t + tny) * a + f[16](i, j + 1, 0) * b;
            } else {
                const double t = (rin - (f[0](i, j, 1) + f[2](i, j, 1) + f[1](i, j, 1) + f[4](i, j, 1) + f[3](i, j, 1) +
                                         f[8](i, j, 1) + f[7](i, j, 1) + f[10](i, j, 1) + f[9](i, j, 1) +
                                         2.0 * (f[5](i, j, 1) + f[11](i, j, 1) + f[12](i, j, 1) + f[15](i, j, 1) +
                                                f[16](i, j, 1)))) * rel;
                const double tnx = 0.5 * (f[2](i, j, 1) + f[8](i, j, 1) + f[10](i, j, 1) - (f[1](i, j, 1) + f[7](i, j, 1) + f[9](i, j, 1)));
                const double tny = 0.5 * (f[4](i, j, 1) + f[9](i, j, 1) + f[10](i, j, 1) - (f[3](i, j, 1) + f[8](i, j, 1) + f[7](i, j, 1)));
                f[6](i, j, 1) = (f[5](i, j, 1) + o3 * t) * a + f[6](i, j, 1) * b;
                f[13](i, j, 1) = (f[12](i, j, 1) + o6 * t + tnx) * a + f[13](i, j, 1) * b;
                f[14](i, j, 1) = (f[11](i, j, 1) + o6 * t - tnx) * a + f[14](i, j, 1) * b;
                f[17](i, j, 1) = (f[16](i, j, 1) + o6 * t + tny) * a + f[17](i, j, 1) * b;
                f[18](i, j, 1) = (f[15](i, j, 1) + o6 * t - tny) * a + f[18](i, j, 1) * b;
            }
        }
}

// convective outlet (Boundary.F90:192-267)
void sp_outlet_convective(oracle_sp_ctx* c, bool after) {
    Field* f = c->f;
    const int nz = c->nz;
    const double uc = c->p.uin_avg, temp = 1.0 / (1.0 + uc);
    for (int j = 1; j <= c->ny; ++j)
        for (int i = 1; i <= c->nx; ++i) {
            const int wi = c->W(i, j, nz);
            const double a = (double)(1 - wi), b = (double)wi;
            if (!after) {
                f[6](i, j, nz + 1) = ((c->FC(i, j, 6) + uc * f[6](i, j, nz)) * temp) * a + f[6](i, j, nz + 1) * b;
                f[13](i - 1, j, nz + 1) = ((c->FC(i, j, 13) + uc * f[13](i - 1, j, nz)) * temp) * a + f[13](i - 1, j, nz + 1) * b;
                f[14](i + 1, j, nz + 1) = ((c->FC(i, j, 14) + uc * f[14](i + 1, j, nz)) * temp) * a + f[14](i + 1, j, nz + 1) * b;
                f[17](i, j - 1, nz + 1) = ((c->FC(i, j, 17) + uc * f[17](i, j - 1, nz)) * temp) * a + f[17](i, j - 1, nz + 1) * b;
                f[18](i, j + 1, nz + 1) = ((c->FC(i, j, 18) + uc * f[18](i, j + 1, nz)) * temp) * a + f[18](i, j + 1, nz + 1) * b;
                c->FC(i, j, 6) = f[6](i, j, nz + 1);
                c->FC(i, j, 13) = f[13](i - 1, j, nz + 1);
                c->FC(i, j, 14) = f[14](i + 1, j, nz + 1);
                c->FC(i, j, 17) = f[17](i, j - 1, nz + 1);
                c->FC(i, j, 18) = f[18](i, j + 1, nz + 1);
            } else {
                f[5](i, j, nz) = ((c->FC(i, j, 6) + uc * f[5](i, j, nz - 1)) * temp) * a + f[5](i, j, nz) * b;
                f[11](i, j, nz) = ((c->FC(i, j, 14) + uc * f[11](i, j, nz - 1)) * temp) * a + f[11](i, j, nz) * b;
                f[12](i, j, nz) = ((c->FC(i, j, 13) + uc * f[12](i, j, nz - 1)) * temp) * a + f[12](i, j, nz) * b;
                f[15](i, j, nz) = ((c->FC(i, j, 18) + uc * f[15](i, j, nz - 1)) * temp) * a + f[15](i, j, nz) * b;
                f[16](i, j, nz) = ((c->FC(i, j, 17) + uc * f[16](i, j, nz - 1)) * temp) * a + f[16](i, j, nz) * b;
                c->FC(i, j, 6) = f[5](i, j, nz);
                c->FC(i, j, 14) = f[11](i, j, nz);
                c->FC(i, j, 13) = f[12](i, j, nz);
                c->FC(i, j, 18) = f[15](i, j, nz);
                c->FC(i, j, 17) = f[16](i, j, nz);
            }
        }
}

// Zou-He pressure outlet (Boundary.F90:276-389)
void sp_outlet_pressure(oracle_sp_ctx* c, bool after) {
    Field* f = c->f;
    const int nz = c->nz;
    const double rout = c->p.rho_out, o3 = 0.333333333333333333, o6 = 0.166666666666666667;
    for (int j = 1; j <= c->ny; ++j)
        for (int i = 1; i <= c->nx; ++i) {
            const int wi = c->W(i, j, nz);
            const double a = 1.0 - (double)wi, b = (double)wi;
            if (!after) {
                const double t = (f[0](i, j, nz) + f[1](i - 1, j, nz) + f[2](i + 1, j, nz) + f[3](i, j - 1, nz) +
                                  f[4](i, j + 1, nz) + f[7](i - 1, j - 1, nz) + f[8](i + 1, j - 1, nz) +
                                  f[9](i - 1, j + 1, nz) + f[10](i + 1, j + 1, nz) +
                                  2.0 * (f[5](i, j, nz - 1) + f[11](i - 1, j, nz - 1) + f[12](i + 1, j, nz - 1) +
                                         f[15](i, j - 1, nz - 1) + f[16](i, j + 1, nz - 1))) - rout;
                const double tnx = 0.5 * (f[1](i - 1, j, nz) + f[7](i - 1, j - 1, nz) + f[9](i - 1, j + 1, nz) -
                                          (f[2](i + 1, j, nz) + f[8](i + 1, j - 1, nz) + f[10](i + 1, j + 1, nz)));
                const double tny = 0.5 * (f[3](i, j - 1, nz) + f[7](i - 1, j - 1, nz) + f[8](i + 1, j - 1, nz) -
                                          (f[4](i, j + 1, nz) + f[10](i + 1, j + 1, nz) + f[9](i - 1, j + 1, nz)));
                f[6](i, j, nz + 1) = (f[5](i, j, nz - 1) - o3 * t) * a + f[6](i, j, nz + 1) * b;
                f[13](i - 1, j, nz + 1) = (f[12](i + 1, j, nz - 1) - o6 * t - tnx) * a + f[13](i - 1, j, nz + 1) * b;
                f[14](i + 1, j, nz + 1) = (f[11](i - 1, j, nz - 1) - o6 * t + tnx) * a + f[14](i + 1, j, nz + 1) * b;
                f[17](i, j - 1, nz + 1) = (f[16](i, j + 1, nz - 1) - o6 * t - tny) * a + f[17](i, j - 1, nz + 1) * b;
                f[18](i, j + 1, nz + 1) = (f[15](i, j - 1, nz - 1) - o6 * t + tny) * a + f[18](i, j + 1, nz + 1) * b;
            } else {
                const double t = (f[0](i, j, nz) + f[2](i, j, nz) + f[1](i, j, nz) + f[4](i, j, nz) + f[3](i, j, nz) +
                                  f[8](i, j, nz) + f[7](i, j, nz) + f[10](i, j, nz) + f[9](i, j, nz) +
                                  2.0 * (f[6](i, j, nz) + f[14](i, j, nz) + f[13](i, j, nz) + f[18](i, j, nz) +
                                         f[17](i, j, nz))) - rout;
                const double tnx = 0.5 * (f[2](i, j, nz) + f[8](i, j, nz) + f[10](i, j, nz) - (f[1](i, j, nz) + f[7](i, j, nz) + f[9](i, j, nz)));
                const double tny = 0.5 * (f[4](i, j, nz) + f[10](i, j, nz) + f[9](i, j, nz) - (f[3](i, j, nz) + f[7](i, j, nz) + f[8](i, j, nz)));
                f[5](i, j, nz) = (f[6](i, j, nz) - o3 * t) * a + f[5](i, j, nz) * b;
                f[11](i, j, nz) = (f[14](i, j, nz) - o6 * t + tnx) * a + f[11](i, j, nz) * b;
                f[12](i, j, nz) = (f[13](i, j, nz) - o6 * t - tnx) * a + f[12](i, j, nz) * b;
                f[15](i, j, nz) = (f[18](i, j, nz) - o6 * t + tny) * a + f[15](i, j, nz) * b;
                f[16](i, j, nz) = (f[17](i, j, nz) - o6 * t - tny) * a + f[16](i, j, nz) * b;
            }
        }
}

// main_iteration_kernel (Main.F90:291-422)
void sp_step(oracle_sp_ctx* c, int ntime) {
    const oracle_sp_params& P = c->p;
    const bool odd = (ntime % 2) != 0;
    sp_lattice_update(c, odd);
    if (P.jper == 1 || P.kper == 1) sp_exchange(c, odd);
    if (P.kper == 0 && P.wall_z_min == 0 && P.wall_z_max == 0) {
        if (P.inlet_bc == 1) sp_inlet_velocity(c, odd);
        else if (P.inlet_bc == 2) sp_inlet_pressure(c, odd);
        if (P.outlet_bc == 1) sp_outlet_convective(c, odd);
        else if (P.outlet_bc == 2) sp_outlet_pressure(c, odd);
    }
}

}  // namespace

extern "C" int oracle_sp_create(const oracle_sp_params* p, oracle_sp_ctx** out) {
    if (!p || !out || p->nx < 3 || p->ny < 3 || p->nz < 3 || !(p->la_nu > 0.0)) return 1;
    oracle_sp_ctx* c = new (std::nothrow) oracle_sp_ctx();
    if (!c) return 4;
    c->p = *p;
    c->nx = p->nx;
    c->ny = p->ny;
    c->nz = p->nz;
    // MRT rates (Initialization.F90:87-112): preset 1 is the bounce-back optimised set; the branch for the
    // "original" set repeats the test for 1 and is never taken, so every other preset is SRT
    const double omega = 1.0 / (3.0 * p->la_nu + 0.5);
    c->s_nu = omega;
    c->s_e = omega;
    c->s_e2 = omega;
    c->s_pi = omega;
    if (p->mrt_preset == 1) {
        c->s_q = 8.0 * (2.0 - omega) / (8.0 - omega);
        c->s_t = c->s_q;
    } else {
        c->s_q = omega;
        c->s_t = omega;
    }
    for (int q = 0; q < 19; ++q) c->f[q].init(c->nx, c->ny, c->nz);
    for (Field* a : {&c->u, &c->v, &c->w, &c->rho}) a->init(c->nx, c->ny, c->nz);
    c->walls.assign((size_t)(c->nx + 4) * (c->ny + 4) * (c->nz + 4), 1);
    const size_t plane = (size_t)(c->nx + 2) * (c->ny + 2);
    c->w_in.assign(plane, 0.0);
    c->f_convec.assign(plane * 19, 0.0);
    *out = c;
    return 0;
}

extern "C" void oracle_sp_destroy(oracle_sp_ctx* c) { delete c; }

extern "C" int oracle_sp_set_scalars(oracle_sp_ctx* c, double force_z, double rho_in, double uin_avg, double relaxation) {
    if (!c) return 1;
    c->p.force_z = force_z;
    c->p.rho_in = rho_in;
    c->p.uin_avg = uin_avg;
    c->p.relaxation = relaxation;
    return 0;
}

extern "C" int oracle_sp_upload_geometry(oracle_sp_ctx* c, const int8_t* walls_h2, const double* w_in) {
    if (!c || !walls_h2) return 1;
    std::memcpy(c->walls.data(), walls_h2, c->walls.size());
    if (w_in) std::memcpy(c->w_in.data(), w_in, c->w_in.size() * sizeof(double));
    return 0;
}

extern "C" int oracle_sp_upload_state(oracle_sp_ctx* c, const double* const* f, const double* f_convec) {
    if (!c || !f) return 1;
    for (int q = 0; q < 19; ++q) std::memcpy(c->f[q].v.data(), f[q], c->f[q].v.size() * sizeof(double));
    if (f_convec) std::memcpy(c->f_convec.data(), f_convec, c->f_convec.size() * sizeof(double));
    return 0;
}

extern "C" int oracle_sp_download_state(oracle_sp_ctx* c, double* const* f, double* f_convec) {
    if (!c) return 1;
    for (int q = 0; q < 19 && f; ++q)
        if (f[q]) std::memcpy(f[q], c->f[q].v.data(), c->f[q].v.size() * sizeof(double));
    if (f_convec) std::memcpy(f_convec, c->f_convec.data(), c->f_convec.size() * sizeof(double));
    return 0;
}

extern "C" int oracle_sp_step(oracle_sp_ctx* c, int32_t ntime) {
    if (!c) return 1;
    sp_step(c, ntime);
    return 0;
}

extern "C" int oracle_sp_run(oracle_sp_ctx* c, int32_t ntime0, int32_t nsteps) {
    if (!c) return 1;
    for (int s = 0; s < nsteps; ++s) sp_step(c, ntime0 + s);
    return 0;
}

// compute_macro_vars (Misc.F90:368-423)
extern "C" int oracle_sp_compute_macro(oracle_sp_ctx* c) {
    if (!c) return 1;
    const double fx = 0.0, fy = 0.0, fz = c->p.force_z;
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) {
                const int wi = c->W(i, j, k);
                double ft[19];
                for (int q = 0; q < 19; ++q) ft[q] = c->f[q](i, j, k);
                double s = ft[0];
                for (int q = 1; q < 19; ++q) s = s + ft[q];
                c->rho(i, j, k) = s * (double)(1 - wi);
                c->u(i, j, k) = (ft[1] - ft[2] + ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14] - 0.5 * fx) * (double)(1 - wi);
                c->v(i, j, k) = (ft[3] - ft[4] + ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18] - 0.5 * fy) * (double)(1 - wi);
                c->w(i, j, k) = (ft[5] - ft[6] + ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18] - 0.5 * fz) * (double)(1 - wi);
            }
    return 0;
}

// monitor sums (Monitor.F90:20-48), serial in the reference's loop order
extern "C" int oracle_sp_monitor(oracle_sp_ctx* c, double* layers, double* umax2) {
    if (!c) return 1;
    oracle_sp_compute_macro(c);
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    for (int k = 1; k <= nz; ++k) {
        double temp = 0.0, prek = 0.0;
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                temp = temp + c->w(i, j, k);
                prek = prek + c->rho(i, j, k);
            }
        if (layers) {
            layers[k - 1] = temp;
            layers[nz + k - 1] = prek;
        }
    }
    double um = 0.0;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double t = c->u(i, j, k) * c->u(i, j, k) + c->v(i, j, k) * c->v(i, j, k) + c->w(i, j, k) * c->w(i, j, k);
                if (um < t) um = t;
            }
    if (umax2) *umax2 = um;
    return 0;
}

extern "C" int oracle_sp_download_macro(oracle_sp_ctx* c, double* u, double* v, double* w, double* rho) {
    if (!c) return 1;
    Field* src[4] = {&c->u, &c->v, &c->w, &c->rho};
    double* dst[4] = {u, v, w, rho};
    for (int m = 0; m < 4; ++m)
        if (dst[m]) std::memcpy(dst[m], src[m]->v.data(), src[m]->v.size() * sizeof(double));
    return 0;
}
