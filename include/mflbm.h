/*
 * mflbm.h -- C ABI of the MI355X (gfx950) implementation of MF-LBM's per-time-step lattice update.
 *
 * The reference drives its hot path through Fortran subroutines without field arguments that operate
 * on module globals inside one OpenACC data region.  This ABI replaces that boundary with a context
 * handle that owns the device copies of the same state; every array crossing the boundary keeps the
 * reference's host layout (Fortran column-major, x fastest, with the reference's halo widths):
 *
 *   PDF arrays f0..f18, g0..g18, curv, u, v, w, rho : (0:nx+1, 0:ny+1, 0:nz+1)   halo 1
 *   phi                                            : (-3:nx+4, -3:ny+4, -3:nz+4) halo 4
 *   cn_x, cn_y, cn_z, c_norm                       : (-1:nx+2, -1:ny+2, -1:nz+2) halo 2
 *   walls (int8, 1 = solid)                        : (-1:nx+2, -1:ny+2, -1:nz+2) halo 2
 *   w_in, phi_convec_bc                            : (0:nx+1, 0:ny+1)
 *   f_convec_bc, g_convec_bc                       : (0:nx+1, 0:ny+1, 0:18)
 *   (multiphase_3D/0.src/Init_multiphase.F90:583-700, Module.F90:263)
 *
 * Conventions: every entry point returns 0 (MFLBM_OK) on success and a nonzero MFLBM_ERR_* code
 * otherwise; mflbm_last_error() gives the message.  A context is used by one host thread.  Device work
 * is stream ordered; the download and monitor calls synchronise.  The reference aborts the MPI job on
 * fatal errors (e.g. Init_multiphase.F90:316-318); here the caller decides.
 */
#ifndef MFLBM_H
#define MFLBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFLBM_OK 0
#define MFLBM_ERR_ARG 1      /* invalid argument or size */
#define MFLBM_ERR_HIP 2      /* HIP runtime error (message names the call) */
#define MFLBM_ERR_STATE 3    /* call out of order (e.g. step before upload) */
#define MFLBM_ERR_NOMEM 4    /* allocation failed */
#define MFLBM_ERR_NODEV 5    /* no GPU visible */

/* Per-run constants.  Field meanings follow the reference's module variables
 * (multiphase_3D/0.src/Module.F90:4-241) after initialization_basic_multi (Init_multiphase.F90:5-191). */
typedef struct mflbm_params {
    int32_t nx, ny, nz;                 /* interior lattice (nxGlobal, nyGlobal, nzGlobal; one domain) */
    int32_t jper, kper;                 /* periodic y / z (x is never periodic, IO_multiphase.F90:471) */
    int32_t wall_x_min, wall_x_max;     /* domain_wall_status_* : 1 = solid wall plane */
    int32_t wall_y_min, wall_y_max;
    int32_t wall_z_min, wall_z_max;
    int32_t inlet_bc;                   /* 1 velocity bounce-back, 2 Zou-He pressure */
    int32_t outlet_bc;                  /* 1 convective, 2 Zou-He pressure */
    int32_t porous_plate_cmd;           /* 0 off, 1 block fluid 1, 2 block fluid 2 */
    int32_t z_porous_plate;             /* plate position (1-based z) */
    int32_t n_exclude_inlet, n_exclude_outlet;
    int32_t precision;                  /* 0: FP64 populations, the reference arithmetic (bit-identical);
                                         * 1: FP32 storage of f + g - REST and f - g with FP32 collision and
                                         *    recolouring; 2: the same storage collided in FP64.  phi, the
                                         *    gradient chain, boundaries and monitors stay FP64 (1 and 2 have
                                         *    no reference counterpart) */
    double la_nui1, la_nui2;            /* 1/nu of fluid 1 and fluid 2 */
    double gamma;                       /* surface tension */
    double beta;                        /* recolouring parameter */
    double force_z;                     /* body force along z */
    double sa_inject;                   /* fluid-1 fraction of the injected fluid */
    double phi_inlet;                   /* 2*sa_inject-1 */
    double rho_in, rho_out;             /* Zou-He densities */
    double uin_avg;                     /* convective outlet velocity (= mean inlet velocity) */
    double relaxation;                  /* inlet ramp factor, 1 in the reference */
} mflbm_params;

/* Monitor sums of one call (Monitor.F90:32-85 and :512-550), all FP64.
 * layers: 7*nz values, per z layer k (0-based) at [m*nz + k] for m = 0..6:
 *   fl1, fl2 (flux of each fluid), vol1, vol2 (volume), mass1, mass2, pre (sum of rho). */
typedef struct mflbm_monitor_sums {
    double umax2;          /* max over fluid nodes of u^2+v^2+w^2 */
    double usq1, usq2;     /* sum of u^2 over nodes with phi > 0.999 / phi < -0.999 */
    double vol1_full, vol2_full;   /* cal_saturation sums over the whole interior */
    double dphi_max;       /* max |phi - phi_old| over fluid nodes (steady phase-field monitor) */
    double pre_w, pre_nw;  /* sums of rho over fluid nodes with phi < -0.99 / phi > 0.99 */
    int64_t n_w, n_nw;     /* counts of those nodes */
    int64_t outlet_phase1; /* fluid nodes at z = nz-1 with phi > 0 (breakthrough monitor) */
} mflbm_monitor_sums;

typedef struct mflbm_ctx mflbm_ctx;
typedef struct mflbm_geometry mflbm_geometry;

/* ---------------------------------------------------------------- context -------------------------- */

/* Create a context on the current HIP device and allocate all device state.
 * Replaces MemAllocate_multi (Init_multiphase.F90:583-828) and the `!$acc data` region
 * (Main_multiphase.F90:104-115). */
int mflbm_create(const mflbm_params* params, mflbm_ctx** out);
void mflbm_destroy(mflbm_ctx* ctx);
const char* mflbm_last_error(const mflbm_ctx* ctx);
/* Message of the last failed mflbm_create (no context exists then). */
const char* mflbm_create_error(void);

/* Update run scalars between steps (force_z, rho_in, uin_avg, relaxation); the reference changes them
 * from the driver (e.g. IO_multiphase.F90:586 checkpoint header). */
int mflbm_set_scalars(mflbm_ctx* ctx, double force_z, double rho_in, double uin_avg, double relaxation);

/* Upload the geometry: walls with halo 2, the solid-boundary list (Module.F90:90-96: ix,iy,iz,
 * i_fluid_num, neighbor_list(18), la_weight) and the fluid-boundary list (Module.F90:98-103:
 * ix,iy,iz,nwx,nwy,nwz,theta) as structure-of-arrays, and the inlet profile w_in.  Coordinates are the
 * reference's 1-based local indices.  s_ijk/f_ijk hold 3 ints per node, s_nbr 18 ints per node,
 * f_nw 3 doubles per node.  theta is in radians (already converted as IO_multiphase.F90:467-468). */
int mflbm_upload_geometry(mflbm_ctx* ctx, const int8_t* walls_h2,
                          int32_t n_solid, const int32_t* s_ijk, const int32_t* s_nfluid,
                          const int32_t* s_nbr, const double* s_weight,
                          int32_t n_fluid, const int32_t* f_ijk, const double* f_nw, const double* f_theta,
                          const double* w_in);

/* Upload the full solver state (checkpoint content, IO_multiphase.F90:586-631): 19 pointers each for
 * f and g, phi, and the convective-outlet history (may be NULL when outlet_bc != 1). */
int mflbm_upload_state(mflbm_ctx* ctx, const double* const* f, const double* const* g, const double* phi,
                       const double* f_convec, const double* g_convec, const double* phi_convec);
int mflbm_download_state(mflbm_ctx* ctx, double* const* f, double* const* g, double* phi,
                         double* f_convec, double* g_convec, double* phi_convec);
/* The initial state of a new run at rest computed on the device instead of uploaded
 * (initialization_new_multi / _pdf, Init_multiphase.F90:243-470; equals mflbm_init_fields with
 * u = v = w = NULL followed by mflbm_upload_state): phi by initial_fluid_distribution_option 1..5, equilibrium
 * populations at rho = 1, convective-outlet history when outlet_bc == 1. */
int mflbm_init_state_device(mflbm_ctx* ctx, int32_t option, double interface_z0);

/* ---------------------------------------------------------------- hot path ------------------------- */

/* One time step (main_iteration_kernel, Main_multiphase.F90:341-486): even ntime -> in-place collision
 * (kernel_even_color), odd ntime -> pull/push collision (kernel_odd_color); periodic wrap of PDFs and
 * phi (replaces the MPI self-exchange of Mpi.F90); inlet/outlet/porous-plate boundaries; colour
 * gradient chain (color_gradient, Phase_gradient.F90:5-204). */
int mflbm_step(mflbm_ctx* ctx, int32_t ntime);
/* nsteps consecutive steps ntime0 .. ntime0+nsteps-1, launched back to back. */
int mflbm_run(mflbm_ctx* ctx, int32_t ntime0, int32_t nsteps);
/* The colour-gradient chain alone (Phase_gradient.F90:5-204), as the driver calls it after
 * initialisation (Main_multiphase.F90:120). */
int mflbm_color_gradient(mflbm_ctx* ctx);

/* u, v, w, rho from the PDFs (compute_macro_vars, Misc.F90:372-430; meaningful after even steps), also
 * zeroing phi inside solids as the reference does. Stays on the device. */
int mflbm_compute_macro(mflbm_ctx* ctx);
/* Monitor reductions on the device (Monitor.F90); calls compute_macro first.  layers: 7*nz doubles.
 * update_phi_old != 0 copies phi into phi_old afterwards (monitor_multiphase_steady_phasefield).  phi_old
 * is zero after every state upload until such a call (a driver starting a new run calls it once to set
 * phi_old = phi, Init_multiphase.F90:341-351; a continued run starts from zero, as the reference's unset
 * phi_old does). */
int mflbm_monitor(mflbm_ctx* ctx, double* layers, mflbm_monitor_sums* sums, int32_t update_phi_old);

int mflbm_download_macro(mflbm_ctx* ctx, double* u, double* v, double* w, double* rho, double* phi);
int mflbm_download_gradient(mflbm_ctx* ctx, double* cn_x, double* cn_y, double* cn_z, double* c_norm,
                            double* curv);
int mflbm_synchronize(mflbm_ctx* ctx);
/* Mean duration (ms) of the step kernels recorded with HIP events since the last reset, measured on the
 * context's stream: which = 0 collision-streaming kernel, 1 gradient chain, 2 whole step. */
int mflbm_timing(mflbm_ctx* ctx, int32_t which, double* mean_ms, int64_t* count);
int mflbm_timing_enable(mflbm_ctx* ctx, int32_t enable);

/* ---------------------------------------------------------------- host set-up ---------------------- */

/* Local wall array with halo 2 as set_walls builds it for one rank (Misc.F90:137-210, including the
 * periodic wall transport Mpi_misc.F90:337-502).  walls_global: nx*ny*nz int8, x fastest. */
int mflbm_setup_walls(const mflbm_params* params, const int8_t* walls_global, int8_t* walls_h2);

/* Boundary-node classification, wall normals and extrapolation weights
 * (geometry_preprocessing_new, Geometry_preprocessing.F90:9-512). */
int mflbm_geometry_new(const mflbm_params* params, const int8_t* walls_global, double theta,
                       mflbm_geometry** out);
/* The same lists computed on the current HIP device (smoothing, marks, normals and the list extraction
 * as kernels); results are identical to mflbm_geometry_new. */
int mflbm_geometry_new_device(const mflbm_params* params, const int8_t* walls_global, double theta,
                              mflbm_geometry** out);
int mflbm_geometry_counts(const mflbm_geometry* geom, int32_t* n_solid, int32_t* n_fluid);
int mflbm_geometry_export(const mflbm_geometry* geom, int32_t* s_ijk, int32_t* s_nfluid, int32_t* s_nbr,
                          double* s_weight, int32_t* f_ijk, double* f_nw, double* f_theta);
void mflbm_geometry_free(mflbm_geometry* geom);

/* Initial fields of a new run (initialization_new_multi / _pdf, Init_multiphase.F90:243-470):
 * phi by initial_fluid_distribution_option 1..5, equilibrium PDFs from (u, v, w) (NULL = at rest)
 * and rho = 1, convective-outlet history.  f, g: 19 pointers each. */
int mflbm_init_fields(const mflbm_params* params, int32_t option, double interface_z0,
                      const double* u0, const double* v0, const double* w0,
                      double* const* f, double* const* g, double* phi,
                      double* f_convec, double* g_convec, double* phi_convec);

/* Inlet velocity profile w_in (initialization_open_velocity_inlet_BC, Init_multiphase.F90:194-236 and
 * inlet_vel_profile_rectangular, Misc.F90:625-665). */
int mflbm_inlet_profile(const mflbm_params* params, double uin_avg, double la_x, double la_y, double* w_in);

/* ================================================================ single-phase solver =================
 * The reference's single-phase program (singlephase_3D/0.src) runs the same D3Q19 lattice, AA streaming,
 * periodic self-exchange and open boundaries with one population set and no phase field.  Host arrays
 * keep its layout: f0..f18, u, v, w, rho over (0:nx+1, 0:ny+1, 0:nz+1); walls with halo 2; w_in over
 * (0:nx+1, 0:ny+1); f_convec_bc over (0:nx+1, 0:ny+1, 0:18) (Initialization.F90:390-402).  Walls and the
 * inlet profile come from mflbm_setup_walls / mflbm_inlet_profile (the routines are the same in both
 * programs; only the mflbm_params fields nx..wall_z_max are read by them). */

/* Per-run constants after initialization_basic (singlephase_3D/0.src/Initialization.F90:4-152). */
typedef struct mflbm_sp_params {
    int32_t nx, ny, nz;
    int32_t jper, kper;
    int32_t wall_x_min, wall_x_max, wall_y_min, wall_y_max, wall_z_min, wall_z_max;
    int32_t inlet_bc;          /* 1 velocity bounce-back, 2 Zou-He pressure (Boundary.F90) */
    int32_t outlet_bc;         /* 1 convective, 2 Zou-He pressure */
    int32_t n_exclude_inlet, n_exclude_outlet;
    int32_t mrt_preset;        /* MRT_collision_parameter_preset: 1 bounce-back optimised, else SRT */
    int32_t precision;         /* 0: FP64 populations, the reference arithmetic (bit-identical);
                                * 1: FP32 deviations from the rest equilibrium, FP32 collision; 2: the same
                                *    storage collided in FP64.  Boundaries, macroscopic fields and reductions
                                *    stay FP64 (1 and 2 have no reference counterpart) */
    double la_nu;              /* fluid_viscosity */
    double force_z;            /* body force along z (force_z0) */
    double rho_in, rho_out;    /* Zou-He densities (rho_out = 1) */
    double uin_avg;            /* convective outlet velocity (= mean inlet velocity) */
    double relaxation;         /* inlet ramp factor, 1 in the reference */
} mflbm_sp_params;

typedef struct mflbm_sp_ctx mflbm_sp_ctx;

/* Context on the current HIP device (MemAllocate, Initialization.F90:329-460, and the `!$acc data`
 * region, Main.F90:104-110).  A failed create leaves its message in mflbm_create_error(). */
int mflbm_sp_create(const mflbm_sp_params* params, mflbm_sp_ctx** out);
void mflbm_sp_destroy(mflbm_sp_ctx* ctx);
const char* mflbm_sp_last_error(const mflbm_sp_ctx* ctx);
int mflbm_sp_set_scalars(mflbm_sp_ctx* ctx, double force_z, double rho_in, double uin_avg, double relaxation);
/* walls with halo 2 (mflbm_setup_walls) and the inlet profile (may be NULL unless inlet_bc == 1) */
int mflbm_sp_upload_geometry(mflbm_sp_ctx* ctx, const int8_t* walls_h2, const double* w_in);
/* checkpoint content (save_checkpoint, IO.F90:490-547): 19 population pointers and the convective
 * outlet history (may be NULL when outlet_bc != 1) */
int mflbm_sp_upload_state(mflbm_sp_ctx* ctx, const double* const* f, const double* f_convec);
int mflbm_sp_download_state(mflbm_sp_ctx* ctx, double* const* f, double* f_convec);
/* One time step (main_iteration_kernel, Main.F90:291-422): kernel_even / kernel_odd (Kernel.F90:5-400),
 * periodic wrap (Mpi.F90), inlet/outlet boundaries (Boundary.F90). */
int mflbm_sp_step(mflbm_sp_ctx* ctx, int32_t ntime);
int mflbm_sp_run(mflbm_sp_ctx* ctx, int32_t ntime0, int32_t nsteps);
/* u, v, w, rho (compute_macro_vars, Misc.F90:368-423; meaningful after even steps), on the device */
int mflbm_sp_compute_macro(mflbm_sp_ctx* ctx);
/* Monitor reductions (monitor, Monitor.F90:4-172); calls compute_macro first.  layers: 2*nz doubles,
 * fl(k) = sum of w over layer k at [k], pre(k) = sum of rho at [nz + k]; umax2 = max of u^2+v^2+w^2. */
int mflbm_sp_monitor(mflbm_sp_ctx* ctx, double* layers, double* umax2);
int mflbm_sp_download_macro(mflbm_sp_ctx* ctx, double* u, double* v, double* w, double* rho);
int mflbm_sp_synchronize(mflbm_sp_ctx* ctx);
/* Mean duration (ms) per step since the last reset: which = 0 lattice-update kernel, 2 whole step. */
int mflbm_sp_timing(mflbm_sp_ctx* ctx, int32_t which, double* mean_ms, int64_t* count);
int mflbm_sp_timing_enable(mflbm_sp_ctx* ctx, int32_t enable);

/* Initial populations of a new run (initialization_new / _pdf, Initialization.F90:207-307): equilibrium
 * at (rho, u, v, w) over the whole halo-1 array (NULL = rho 1 at rest, the reference's start), and the
 * convective-outlet history f_convec_bc(:, :, q) = f_q(:, :, nz) (may be NULL). */
int mflbm_sp_init_fields(const mflbm_sp_params* params, const double* rho0, const double* u0, const double* v0,
                         const double* w0, double* const* f, double* f_convec);
/* No reference counterpart (the reference starts at the lattice-weight equilibrium): the state of a context set
 * on the device to the MRT collision's equilibrium of (rho, u, v, w) (halo-1 host arrays, NULL = rho 1 / at rest),
 * f_q = (feq_q - w_q rho) + REST_q rho with REST the collision's rest state, and every solid slot facing a fluid
 * node set to that node's opposite population (its first bounce-back); for starting a run from predicted fields
 * without building and uploading 19 population arrays.  Leaves f_convec unchanged. */
int mflbm_sp_set_equilibrium(mflbm_sp_ctx* ctx, const double* rho, const double* u, const double* v, const double* w);

/* ---------------------------------------------------------------------------------------------------------
 * Batched single-phase permeability runs (no reference counterpart: the reference runs one sample per job
 * through singlephase_3D/0.src/Main.F90:124-230).  Every sample is a complete single-phase problem with
 * solid x walls, z periodic, y walls or y periodic, a body force along z and no open boundaries; all samples
 * share the rates and the force.  One lattice-update launch per step covers all samples.  A steady run
 * replays a captured monitor period (ntime_monitor steps + the flow-rate test of Monitor.F90:145-156 per
 * sample) until every sample has converged; a converged sample stops and keeps its state.
 * --------------------------------------------------------------------------------------------------------- */
typedef struct mflbm_spb_params {
    int32_t precision;         /* as mflbm_sp_params.precision: 0 FP64, 1 FP32, 2 FP32 storage with FP64 collision */
    int32_t mrt_preset;        /* as mflbm_sp_params */
    int32_t ntime_monitor;     /* steps per convergence test (even: the flow rate is read after even steps) */
    int32_t tile_rows;         /* compact node order: padded y rows per tile (0: MFLBM_TILE_Y or 32) */
    int32_t block;             /* threads per block of the lattice update: 64, 128, 256 or 512 (0: 256) */
    int32_t rest_state;        /* initial state at rest: 0 the lattice weights (initialization_new), 1 the MRT
                                  collision's rest equilibrium (4/21, 13/126, 1/63), which starts without the
                                  density waves the lattice weights launch at solid links */
    double la_nu;              /* fluid_viscosity */
    double force_z;            /* body force along z */
    double tol;                /* convergence_criteria: |Q - Q_prev| / (|Q| + 1e-14) < tol */
} mflbm_spb_params;
typedef struct mflbm_spb_ctx mflbm_spb_ctx;
int mflbm_spb_create(const mflbm_spb_params* params, mflbm_spb_ctx** out);
void mflbm_spb_destroy(mflbm_spb_ctx* ctx);
const char* mflbm_spb_last_error(const mflbm_spb_ctx* ctx);
/* walls with halo 2 as mflbm_setup_walls returns them for (nx, ny, nz, jper, kper = 1), x walls */
int mflbm_spb_add_sample(mflbm_spb_ctx* ctx, int32_t nx, int32_t ny, int32_t nz, int32_t jper, const int8_t* walls_h2);
/* lays out and uploads every sample; initial state: fluid at rest, rho 1 (initialization_new) */
int mflbm_spb_commit(mflbm_spb_ctx* ctx);
int mflbm_spb_info(mflbm_spb_ctx* ctx, int32_t* nsamples, int64_t* fluid_nodes, int64_t* pop_bytes);
int mflbm_spb_sample_fluid(mflbm_spb_ctx* ctx, int32_t sample, int64_t* fluid_nodes);
/* one sample's populations in the layout of mflbm_sp_upload_state / _download_state */
int mflbm_spb_upload_state(mflbm_spb_ctx* ctx, int32_t sample, const double* const* f);
int mflbm_spb_download_state(mflbm_spb_ctx* ctx, int32_t sample, double* const* f);
/* plain steps of every running sample (kernel_odd / kernel_even by the parity of ntime, periodic wrap) */
int mflbm_spb_run(mflbm_spb_ctx* ctx, int32_t ntime0, int32_t nsteps);
/* up to max_periods monitor periods; returns early when every sample has stopped; nactive = samples left */
int mflbm_spb_run_steady(mflbm_spb_ctx* ctx, int32_t max_periods, int32_t* nactive);
/* per sample: step of convergence (> 0), of a NaN flow rate (< 0) or 0 while running; the last flow rate
 * Q = sum(w) / nz; periods = monitor periods done */
int mflbm_spb_results(mflbm_spb_ctx* ctx, int32_t* conv_step, double* flowrate, int32_t* periods);
/* restarts the monitor: every sample runs again, period count and previous flow rates cleared */
int mflbm_spb_reset_monitor(mflbm_spb_ctx* ctx);
int mflbm_spb_synchronize(mflbm_spb_ctx* ctx);
/* one sample's u, v, w, rho (compute_macro_vars, after an even step; any pointer may be NULL) in the halo-1
 * layout of mflbm_sp_download_macro; the training data of the warm-start network */
int mflbm_spb_download_macro(mflbm_spb_ctx* ctx, int32_t sample, double* u, double* v, double* w, double* rho);

/* ---------------------------------------------------------------------------------------------------------
 * Learned warm start (no reference counterpart: the reference starts every run at rest): kernels of a 3-D
 * convolutional encoder-decoder that maps a wall array to a steady velocity field, trained on fields the
 * solver produces.  Device pointers; tensors channels-last x[n][z][y][x][c] in BF16 with c a multiple of 8;
 * 3x3x3 convolutions with zero padding on the matrix cores (BF16 in, FP32 accumulation).  `stream` is a
 * hipStream_t (NULL = default stream).  Weights: FP32 master w[cout][27][cin]; BF16 rows wf[cout][kpad] for the
 * forward pass and wb[cin][kpadb] (flipped, channels swapped) for the input gradient, kpad = roundup(27 cin, 32),
 * kpadb = roundup(27 cout, 32).
 * --------------------------------------------------------------------------------------------------------- */
int mflbm_nn_weights_to_bf16(const float* w, void* wf, void* wb, int32_t cin, int32_t cout, void* stream);
int mflbm_nn_conv3d_fwd(const void* x, const void* wf, const float* bias, void* y, int32_t n, int32_t d, int32_t h,
                        int32_t w, int32_t cin, int32_t cout, int32_t relu, void* stream);
/* accumulates the weight gradient (FP32, rows of 27 cin) and, when dbias is not NULL, the bias gradient
 * (fewer than 2^31 positions n d h w, else MFLBM_ERR_ARG) */
int mflbm_nn_conv3d_dw(const void* x, const void* dy, float* dw, float* dbias, int32_t n, int32_t d, int32_t h,
                       int32_t w, int32_t cin, int32_t cout, void* stream);
/* the same gradients summed in a fixed order, so that repeated calls give identical bits: the positions are cut
 * into chunks (4096 for the weights, 8192 for the bias), each chunk's sum is stored to its own slice of the
 * caller's FP32 workspace (work_floats >= mflbm_nn_conv3d_dw_work(...), else MFLBM_ERR_ARG), then per element the
 * chunk sums are added in chunk order and that total is added to dw / dbias once */
int64_t mflbm_nn_conv3d_dw_work(int32_t n, int32_t d, int32_t h, int32_t w, int32_t cin, int32_t cout);
int mflbm_nn_conv3d_dw_ordered(const void* x, const void* dy, float* dw, float* dbias, float* work, int64_t work_floats,
                               int32_t n, int32_t d, int32_t h, int32_t w, int32_t cin, int32_t cout, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP graph capture,
 * tight training loops, harnesses that forbid hipMalloc): mflbm_nn_conv3d_dw_ordered with the workspace taken
 * from a fixed array of mflbm_nn_conv3d_dw_scratch_floats() floats (16 MiB) that the library holds on each GPU
 * until the process ends.  It calls no hipMalloc/hipFree; the scratch address of the current device is looked
 * up once per device on the first call there.  MFLBM_ERR_ARG when mflbm_nn_conv3d_dw_work(...) exceeds the array
 * (or for the arguments mflbm_nn_conv3d_dw_ordered refuses).  Results are bit-identical to
 * mflbm_nn_conv3d_dw_ordered.  Every call on one GPU uses that GPU's one array: calls that overlap on one GPU
 * (other host threads, or other streams without ordering between them) race; `stream` must belong to the current
 * device. */
int64_t mflbm_nn_conv3d_dw_scratch_floats(void);
int mflbm_nn_conv3d_dw_ordered_noalloc(const void* x, const void* dy, float* dw, float* dbias, int32_t n, int32_t d,
                                       int32_t h, int32_t w, int32_t cin, int32_t cout, void* stream);
int mflbm_nn_relu_bwd(void* g, const void* y, int64_t n, void* stream);
/* channels 0..2 of the BF16 channels-last input x (d, h, w, cin): fluid indicator (fluid: int8 1 fluid, 0 solid),
 * d / 8 and (d / 8)^2, d the Euclidean distance to the nearest solid node capped at 16 (wrapped along the axes
 * flagged periodic); work: 2 d h w int32 of scratch */
int mflbm_nn_distance_features(const int8_t* fluid, int32_t* work, void* x, int32_t d, int32_t h, int32_t w,
                               int32_t cin, int32_t zper, int32_t yper, int32_t xper, void* stream);
/* 2x average pooling of the fine tensor (bwd = 0) or its adjoint into the fine gradient (bwd = 1); (d, h, w) fine */
int mflbm_nn_pool2(const void* x, void* y, int32_t n, int32_t d, int32_t h, int32_t w, int32_t c, int32_t bwd, void* stream);
/* 2x nearest upsampling into the fine tensor (bwd = 0) or its adjoint into the coarse gradient; (d, h, w) fine */
int mflbm_nn_up2(const void* x, void* y, int32_t n, int32_t d, int32_t h, int32_t w, int32_t c, int32_t bwd, void* stream);
int mflbm_nn_add(const void* a, const void* b, void* y, int64_t n, void* stream);
/* squared error over the first `used` of c channels: g = 2 scale (y - t) (BF16), *loss += sum of squares (FP32) */
int mflbm_nn_mse(const void* y, const void* t, void* g, float* loss, int64_t npos, int32_t c, int32_t used, float scale,
                 void* stream);
int mflbm_nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
                  int32_t step, void* stream);
/* layout check of the matrix-core instruction: d (16 x 16, FP32) = a (16 x 32, BF16 row-major) times b given
 * as 16 columns of 32 (BF16) */
int mflbm_nn_mfma_probe(const void* a, const void* b, float* d, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MFLBM_H */
