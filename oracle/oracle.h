/*
 * oracle.h -- CPU restatement of MF-LBM's two-phase time step, used ONLY as the checker of the gfx950
 * path (tests, the smoke check and the CPU-baseline leg of the benchmark).  Nothing in the product links
 * or calls it.
 *
 * It holds the solver state in the reference's own layout (separate arrays per population, Fortran
 * column-major with the reference's halo widths, Init_multiphase.F90:583-700) and restates each routine
 * of the step from the reference source; every function cites the lines it follows.  The entry points
 * mirror the product ABI (include/mflbm.h) so one test harness drives both.
 *
 * The single-phase program (singlephase_3D/0.src) has its own restatement in oracle_sp.cpp behind the
 * oracle_sp_* entry points, mirroring the mflbm_sp_* part of the ABI.
 *
 * Parity pins: the transition from the reference's own initial state is compared against checkpoints of
 * the unmodified reference (fixtures in tests/golden, sha256 of the exact bytes), on periodic, velocity/
 * convective, Zou-He pressure and y/z-periodic cases.
 */
#ifndef MFLBM_ORACLE_H
#define MFLBM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_params {   /* same field order and meaning as mflbm_params */
    int32_t nx, ny, nz;
    int32_t jper, kper;
    int32_t wall_x_min, wall_x_max, wall_y_min, wall_y_max, wall_z_min, wall_z_max;
    int32_t inlet_bc, outlet_bc, porous_plate_cmd, z_porous_plate;
    int32_t n_exclude_inlet, n_exclude_outlet;
    int32_t precision;   /* ignored: the oracle computes in FP64 */
    double la_nui1, la_nui2, gamma, beta, force_z, sa_inject, phi_inlet, rho_in, rho_out, uin_avg, relaxation;
} oracle_params;

typedef struct oracle_monitor_sums {   /* same layout as mflbm_monitor_sums */
    double umax2, usq1, usq2, vol1_full, vol2_full, dphi_max, pre_w, pre_nw;
    int64_t n_w, n_nw, outlet_phase1;
} oracle_monitor_sums;

typedef struct oracle_ctx oracle_ctx;

int oracle_create(const oracle_params* p, oracle_ctx** out);
void oracle_destroy(oracle_ctx* c);
int oracle_set_threads(int n);
int oracle_upload_geometry(oracle_ctx* c, const int8_t* walls_h2, int32_t ns, const int32_t* s_ijk,
                           const int32_t* s_nfl, const int32_t* s_nbr, const double* s_w, int32_t nf,
                           const int32_t* f_ijk, const double* f_nw, const double* f_theta, const double* w_in);
int oracle_upload_state(oracle_ctx* c, const double* const* f, const double* const* g, const double* phi,
                        const double* f_convec, const double* g_convec, const double* phi_convec);
int oracle_download_state(oracle_ctx* c, double* const* f, double* const* g, double* phi, double* f_convec,
                          double* g_convec, double* phi_convec);
int oracle_step(oracle_ctx* c, int32_t ntime);
int oracle_run(oracle_ctx* c, int32_t ntime0, int32_t nsteps);
int oracle_color_gradient(oracle_ctx* c);
int oracle_compute_macro(oracle_ctx* c);
int oracle_monitor(oracle_ctx* c, double* layers, oracle_monitor_sums* sums, int32_t update_phi_old);
int oracle_download_macro(oracle_ctx* c, double* u, double* v, double* w, double* rho, double* phi);
int oracle_download_gradient(oracle_ctx* c, double* cn_x, double* cn_y, double* cn_z, double* c_norm, double* curv);

/* single-phase solver (singlephase_3D/0.src), oracle_sp.cpp; same layout and meaning as mflbm_sp_params */
typedef struct oracle_sp_params {
    int32_t nx, ny, nz;
    int32_t jper, kper;
    int32_t wall_x_min, wall_x_max, wall_y_min, wall_y_max, wall_z_min, wall_z_max;
    int32_t inlet_bc, outlet_bc;
    int32_t n_exclude_inlet, n_exclude_outlet;
    int32_t mrt_preset;
    int32_t precision;   /* ignored: the oracle computes in FP64 whatever the device path stores */
    double la_nu, force_z, rho_in, rho_out, uin_avg, relaxation;
} oracle_sp_params;

typedef struct oracle_sp_ctx oracle_sp_ctx;

int oracle_sp_create(const oracle_sp_params* p, oracle_sp_ctx** out);
void oracle_sp_destroy(oracle_sp_ctx* c);
int oracle_sp_set_scalars(oracle_sp_ctx* c, double force_z, double rho_in, double uin_avg, double relaxation);
int oracle_sp_upload_geometry(oracle_sp_ctx* c, const int8_t* walls_h2, const double* w_in);
int oracle_sp_upload_state(oracle_sp_ctx* c, const double* const* f, const double* f_convec);
int oracle_sp_download_state(oracle_sp_ctx* c, double* const* f, double* f_convec);
int oracle_sp_step(oracle_sp_ctx* c, int32_t ntime);
int oracle_sp_run(oracle_sp_ctx* c, int32_t ntime0, int32_t nsteps);
int oracle_sp_compute_macro(oracle_sp_ctx* c);
int oracle_sp_monitor(oracle_sp_ctx* c, double* layers, double* umax2);
int oracle_sp_download_macro(oracle_sp_ctx* c, double* u, double* v, double* w, double* rho);

#ifdef __cplusplus
}
#endif
#endif
