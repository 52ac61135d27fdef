! ISO_C_BINDING interface to the C ABI of include/mflbm.h, for the reference's Fortran driver.
!
! The reference's per-step subroutines (kernel_even_color / kernel_odd_color, the boundary routines and
! color_gradient, called from main_iteration_kernel in multiphase_3D/0.src/Main_multiphase.F90:341-486)
! become one call to mflbm_step; the module arrays keep their shapes and are handed over by address.
module mflbm_c_binding
    use, intrinsic :: iso_c_binding
    implicit none

    type, bind(c) :: mflbm_params
        integer(c_int32_t) :: nx, ny, nz
        integer(c_int32_t) :: jper, kper
        integer(c_int32_t) :: wall_x_min, wall_x_max, wall_y_min, wall_y_max, wall_z_min, wall_z_max
        integer(c_int32_t) :: inlet_bc, outlet_bc, porous_plate_cmd, z_porous_plate
        integer(c_int32_t) :: n_exclude_inlet, n_exclude_outlet
        integer(c_int32_t) :: precision   ! 0 FP64 (reference arithmetic), 1 FP32 storage
        real(c_double) :: la_nui1, la_nui2, gamma, beta, force_z, sa_inject, phi_inlet
        real(c_double) :: rho_in, rho_out, uin_avg, relaxation
    end type mflbm_params

    ! single-phase program (singlephase_3D/0.src)
    type, bind(c) :: mflbm_sp_params
        integer(c_int32_t) :: nx, ny, nz
        integer(c_int32_t) :: jper, kper
        integer(c_int32_t) :: wall_x_min, wall_x_max, wall_y_min, wall_y_max, wall_z_min, wall_z_max
        integer(c_int32_t) :: inlet_bc, outlet_bc
        integer(c_int32_t) :: n_exclude_inlet, n_exclude_outlet
        integer(c_int32_t) :: mrt_preset
        integer(c_int32_t) :: precision   ! 0 FP64 (reference arithmetic), 1 FP32 deviation storage
        real(c_double) :: la_nu, force_z, rho_in, rho_out, uin_avg, relaxation
    end type mflbm_sp_params

    ! batched single-phase runs (no reference counterpart; many samples per job)
    type, bind(c) :: mflbm_spb_params
        integer(c_int32_t) :: precision, mrt_preset, ntime_monitor, tile_rows, block, rest_state
        real(c_double) :: la_nu, force_z, tol
    end type mflbm_spb_params

    type, bind(c) :: mflbm_monitor_sums
        real(c_double) :: umax2, usq1, usq2, vol1_full, vol2_full, dphi_max, pre_w, pre_nw
        integer(c_int64_t) :: n_w, n_nw, outlet_phase1
    end type mflbm_monitor_sums

    interface
        integer(c_int) function mflbm_create(params, ctx) bind(c, name="mflbm_create")
            import :: c_int, c_ptr, mflbm_params
            type(mflbm_params), intent(in) :: params
            type(c_ptr), intent(out) :: ctx
        end function
        subroutine mflbm_destroy(ctx) bind(c, name="mflbm_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine
        type(c_ptr) function mflbm_last_error(ctx) bind(c, name="mflbm_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mflbm_set_scalars(ctx, force_z, rho_in, uin_avg, relaxation) &
                bind(c, name="mflbm_set_scalars")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: force_z, rho_in, uin_avg, relaxation
        end function
        integer(c_int) function mflbm_upload_geometry(ctx, walls_h2, n_solid, s_ijk, s_nfluid, s_nbr, s_weight, &
                n_fluid, f_ijk, f_nw, f_theta, w_in) bind(c, name="mflbm_upload_geometry")
            import :: c_int, c_ptr, c_int8_t, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int8_t), intent(in) :: walls_h2(*)
            integer(c_int32_t), value :: n_solid, n_fluid
            integer(c_int32_t), intent(in) :: s_ijk(*), s_nfluid(*), s_nbr(*), f_ijk(*)
            real(c_double), intent(in) :: s_weight(*), f_nw(*), f_theta(*), w_in(*)
        end function
        ! f, g: arrays of 19 addresses (c_loc(f0) ... c_loc(f18))
        integer(c_int) function mflbm_upload_state(ctx, f, g, phi, f_convec, g_convec, phi_convec) &
                bind(c, name="mflbm_upload_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            type(c_ptr), intent(in) :: f(19), g(19)
            real(c_double), intent(in) :: phi(*), f_convec(*), g_convec(*), phi_convec(*)
        end function
        integer(c_int) function mflbm_download_state(ctx, f, g, phi, f_convec, g_convec, phi_convec) &
                bind(c, name="mflbm_download_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            type(c_ptr), intent(in) :: f(19), g(19)
            real(c_double), intent(inout) :: phi(*), f_convec(*), g_convec(*), phi_convec(*)
        end function
        integer(c_int) function mflbm_step(ctx, ntime) bind(c, name="mflbm_step")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: ntime
        end function
        integer(c_int) function mflbm_run(ctx, ntime0, nsteps) bind(c, name="mflbm_run")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: ntime0, nsteps
        end function
        integer(c_int) function mflbm_color_gradient(ctx) bind(c, name="mflbm_color_gradient")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mflbm_compute_macro(ctx) bind(c, name="mflbm_compute_macro")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mflbm_monitor(ctx, layers, sums, update_phi_old) bind(c, name="mflbm_monitor")
            import :: c_int, c_ptr, c_double, c_int32_t, mflbm_monitor_sums
            type(c_ptr), value :: ctx
            real(c_double), intent(inout) :: layers(*)
            type(mflbm_monitor_sums), intent(out) :: sums
            integer(c_int32_t), value :: update_phi_old
        end function
        integer(c_int) function mflbm_download_macro(ctx, u, v, w, rho, phi) bind(c, name="mflbm_download_macro")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(inout) :: u(*), v(*), w(*), rho(*), phi(*)
        end function
        integer(c_int) function mflbm_synchronize(ctx) bind(c, name="mflbm_synchronize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mflbm_setup_walls(params, walls_global, walls_h2) bind(c, name="mflbm_setup_walls")
            import :: c_int, c_int8_t, mflbm_params
            type(mflbm_params), intent(in) :: params
            integer(c_int8_t), intent(in) :: walls_global(*)
            integer(c_int8_t), intent(inout) :: walls_h2(*)
        end function
        ! ---- host set-up (geometry_preprocessing_new, initialization_new_multi, the inlet profile)
        integer(c_int) function mflbm_geometry_new(params, walls_global, theta, geom) bind(c, name="mflbm_geometry_new")
            import :: c_int, c_int8_t, c_double, c_ptr, mflbm_params
            type(mflbm_params), intent(in) :: params
            integer(c_int8_t), intent(in) :: walls_global(*)
            real(c_double), value :: theta
            type(c_ptr), intent(out) :: geom
        end function
        integer(c_int) function mflbm_geometry_counts(geom, n_solid, n_fluid) bind(c, name="mflbm_geometry_counts")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: geom
            integer(c_int32_t), intent(out) :: n_solid, n_fluid
        end function
        integer(c_int) function mflbm_geometry_export(geom, s_ijk, s_nfluid, s_nbr, s_weight, f_ijk, f_nw, f_theta) &
                bind(c, name="mflbm_geometry_export")
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value :: geom
            integer(c_int32_t), intent(inout) :: s_ijk(*), s_nfluid(*), s_nbr(*), f_ijk(*)
            real(c_double), intent(inout) :: s_weight(*), f_nw(*), f_theta(*)
        end function
        subroutine mflbm_geometry_free(geom) bind(c, name="mflbm_geometry_free")
            import :: c_ptr
            type(c_ptr), value :: geom
        end subroutine
        ! u0, v0, w0: c_null_ptr for a run starting at rest; f, g: 19 addresses each
        integer(c_int) function mflbm_init_fields(params, option, interface_z0, u0, v0, w0, f, g, phi, f_convec, &
                g_convec, phi_convec) bind(c, name="mflbm_init_fields")
            import :: c_int, c_ptr, c_int32_t, c_double, mflbm_params
            type(mflbm_params), intent(in) :: params
            integer(c_int32_t), value :: option
            real(c_double), value :: interface_z0
            type(c_ptr), value :: u0, v0, w0
            type(c_ptr), intent(in) :: f(19), g(19)
            real(c_double), intent(inout) :: phi(*), f_convec(*), g_convec(*), phi_convec(*)
        end function
        integer(c_int) function mflbm_inlet_profile(params, uin_avg, la_x, la_y, w_in) bind(c, name="mflbm_inlet_profile")
            import :: c_int, c_double, mflbm_params
            type(mflbm_params), intent(in) :: params
            real(c_double), value :: uin_avg, la_x, la_y
            real(c_double), intent(inout) :: w_in(*)
        end function
        ! ---- single-phase solver: main_iteration_kernel (singlephase_3D/0.src/Main.F90:291-422) -> mflbm_sp_step
        integer(c_int) function mflbm_sp_create(params, ctx) bind(c, name="mflbm_sp_create")
            import :: c_int, c_ptr, mflbm_sp_params
            type(mflbm_sp_params), intent(in) :: params
            type(c_ptr), intent(out) :: ctx
        end function
        subroutine mflbm_sp_destroy(ctx) bind(c, name="mflbm_sp_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine
        integer(c_int) function mflbm_sp_set_scalars(ctx, force_z, rho_in, uin_avg, relaxation) &
                bind(c, name="mflbm_sp_set_scalars")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: force_z, rho_in, uin_avg, relaxation
        end function
        integer(c_int) function mflbm_sp_upload_geometry(ctx, walls_h2, w_in) bind(c, name="mflbm_sp_upload_geometry")
            import :: c_int, c_ptr, c_int8_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int8_t), intent(in) :: walls_h2(*)
            real(c_double), intent(in) :: w_in(*)
        end function
        ! f: 19 addresses (c_loc(f0) ... c_loc(f18))
        integer(c_int) function mflbm_sp_upload_state(ctx, f, f_convec) bind(c, name="mflbm_sp_upload_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            type(c_ptr), intent(in) :: f(19)
            real(c_double), intent(in) :: f_convec(*)
        end function
        integer(c_int) function mflbm_sp_download_state(ctx, f, f_convec) bind(c, name="mflbm_sp_download_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            type(c_ptr), intent(in) :: f(19)
            real(c_double), intent(inout) :: f_convec(*)
        end function
        integer(c_int) function mflbm_sp_step(ctx, ntime) bind(c, name="mflbm_sp_step")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: ntime
        end function
        integer(c_int) function mflbm_sp_run(ctx, ntime0, nsteps) bind(c, name="mflbm_sp_run")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: ntime0, nsteps
        end function
        ! layers: 2*nz doubles, fl(1:nz) then pre(1:nz) of Monitor.F90:20-48
        integer(c_int) function mflbm_sp_monitor(ctx, layers, umax2) bind(c, name="mflbm_sp_monitor")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(inout) :: layers(*)
            real(c_double), intent(out) :: umax2
        end function
        integer(c_int) function mflbm_sp_download_macro(ctx, u, v, w, rho) bind(c, name="mflbm_sp_download_macro")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(inout) :: u(*), v(*), w(*), rho(*)
        end function
        ! the state set on the device to the MRT collision's equilibrium of the fields (halo-1 arrays; any may be
        ! c_null_ptr: rho 1 / at rest), solid slots by bounce-back of their fluid neighbours
        integer(c_int) function mflbm_sp_set_equilibrium(ctx, rho, u, v, w) bind(c, name="mflbm_sp_set_equilibrium")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, rho, u, v, w
        end function
        ! rho0/u0/v0/w0 may be c_null_ptr (rest state at rho = 1); f: 19 addresses; f_convec may be c_null_ptr
        integer(c_int) function mflbm_sp_init_fields(params, rho0, u0, v0, w0, f, f_convec) &
                bind(c, name="mflbm_sp_init_fields")
            import :: c_int, c_ptr, mflbm_sp_params
            type(mflbm_sp_params), intent(in) :: params
            type(c_ptr), value :: rho0, u0, v0, w0, f_convec
            type(c_ptr), intent(in) :: f(19)
        end function
        ! ---- batched single-phase runs
        integer(c_int) function mflbm_spb_create(params, ctx) bind(c, name="mflbm_spb_create")
            import :: c_int, c_ptr, mflbm_spb_params
            type(mflbm_spb_params), intent(in) :: params
            type(c_ptr), intent(out) :: ctx
        end function
        subroutine mflbm_spb_destroy(ctx) bind(c, name="mflbm_spb_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine
        ! walls_h2: the sample's walls with halo 2 (mflbm_setup_walls, kper = 1)
        integer(c_int) function mflbm_spb_add_sample(ctx, nx, ny, nz, jper, walls_h2) bind(c, name="mflbm_spb_add_sample")
            import :: c_int, c_ptr, c_int32_t, c_int8_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nx, ny, nz, jper
            integer(c_int8_t), intent(in) :: walls_h2(*)
        end function
        integer(c_int) function mflbm_spb_commit(ctx) bind(c, name="mflbm_spb_commit")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mflbm_spb_run_steady(ctx, max_periods, nactive) bind(c, name="mflbm_spb_run_steady")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: max_periods
            integer(c_int32_t), intent(out) :: nactive
        end function
        integer(c_int) function mflbm_spb_results(ctx, conv_step, flowrate, periods) bind(c, name="mflbm_spb_results")
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(inout) :: conv_step(*)
            real(c_double), intent(inout) :: flowrate(*)
            integer(c_int32_t), intent(out) :: periods
        end function
        integer(c_int) function mflbm_spb_download_state(ctx, sample, f) bind(c, name="mflbm_spb_download_state")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: sample
            type(c_ptr), intent(in) :: f(19)
        end function
    end interface
end module mflbm_c_binding
