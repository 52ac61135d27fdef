! A Fortran main program that drives the gfx950 hot path through mflbm_c_binding the way the reference's main
! program drives its own (multiphase_3D/0.src/Main_multiphase.F90:20-330): it reads the run's control file
! (the reference's key/value format, IO_multiphase.F90:73-278) and wall file (read_walls / set_walls,
! Misc.F90:18-109, :247-295), derives the run constants (initialization_basic_multi, Init_multiphase.F90:5-191),
! sets the run up with the library's host calls, runs time steps with mflbm_step (main_iteration_kernel,
! Main_multiphase.F90:167), writes a checkpoint in the reference's layout (save_checkpoint,
! IO_multiphase.F90:562-642), reads it back into a new context (initialization_old_multi,
! Init_multiphase.F90:477-557) and continues to the last step, where it writes a second checkpoint.
!
! usage: hotpath_run <control file> <wall file> <first steps> <last step> <first checkpoint> <last checkpoint>
program hotpath_run
    use, intrinsic :: iso_c_binding
    use mflbm_c_binding
    implicit none
    real(c_double), parameter :: pi = 3.14159265358979323846d0
    character(len=512) :: ctl_path, wall_path, ck1, ck2, arg
    integer :: n_first, n_last, ntime, ntime0
    ! control values (defaults of an absent key as the reference's template)
    integer :: nxg = 16, nyg = 16, nzg = 32, iper = 0, jper = 0, kper = 1
    integer :: wx(2) = [1, 1], wy(2) = [1, 1], wz(2) = [0, 0], excl(2) = [0, 0]
    integer :: inlet_bc = 1, outlet_bc = 1, plate_cmd = 0, z_plate = 0, init_option = 1
    real(c_double) :: nu1 = 1d0 / 15d0, nu2 = 1d0 / 15d0, gam = 0.01d0, theta = 90d0, beta = 0.95d0
    real(c_double) :: force0 = 1d-6, sa_inject = 1d0, ca0 = 0d0, z0 = 20d0
    type(mflbm_params) :: p
    real(c_double) :: uin_avg, la_x, la_y, theta_rad
    integer(c_int8_t), allocatable :: wg(:, :, :), wh(:, :, :), sample(:, :, :)
    integer(c_int32_t), allocatable :: s_ijk(:, :), s_nfl(:), s_nbr(:, :), f_ijk(:, :)
    real(c_double), allocatable :: s_w(:), f_nw(:, :), f_theta(:), w_in(:, :)
    real(c_double), allocatable, target :: f(:, :, :, :), g(:, :, :, :), phi(:, :, :), fc(:, :, :), gc(:, :, :), pc(:, :)
    type(c_ptr) :: fp(19), gp(19), geo, ctx
    integer(c_int32_t) :: ns, nfl
    integer :: rc, m

    ctx = c_null_ptr
    call get_command_argument(1, ctl_path)
    call get_command_argument(2, wall_path)
    call get_command_argument(3, arg)
    read(arg, *) n_first
    call get_command_argument(4, arg)
    read(arg, *) n_last
    call get_command_argument(5, ck1)
    call get_command_argument(6, ck2)

    call read_control(trim(ctl_path))
    call derive_params()
    call read_wall_file(trim(wall_path))

    allocate(wh(-1:nxg + 2, -1:nyg + 2, -1:nzg + 2))
    call check(mflbm_setup_walls(p, wg, wh), "mflbm_setup_walls")
    call check(mflbm_geometry_new(p, wg, theta_rad, geo), "mflbm_geometry_new")
    call check(mflbm_geometry_counts(geo, ns, nfl), "mflbm_geometry_counts")
    allocate(s_ijk(3, max(ns, 1)), s_nfl(max(ns, 1)), s_nbr(18, max(ns, 1)), s_w(max(ns, 1)))
    allocate(f_ijk(3, max(nfl, 1)), f_nw(3, max(nfl, 1)), f_theta(max(nfl, 1)))
    call check(mflbm_geometry_export(geo, s_ijk, s_nfl, s_nbr, s_w, f_ijk, f_nw, f_theta), "mflbm_geometry_export")
    call mflbm_geometry_free(geo)
    allocate(w_in(0:nxg + 1, 0:nyg + 1))
    w_in = 0d0
    if (kper == 0 .and. wz(1) == 0 .and. wz(2) == 0 .and. inlet_bc == 1) &
        call check(mflbm_inlet_profile(p, uin_avg, la_x, la_y, w_in), "mflbm_inlet_profile")

    ! the module arrays of the reference, with their halos
    allocate(f(0:nxg + 1, 0:nyg + 1, 0:nzg + 1, 0:18), g(0:nxg + 1, 0:nyg + 1, 0:nzg + 1, 0:18))
    allocate(phi(-3:nxg + 4, -3:nyg + 4, -3:nzg + 4))
    allocate(fc(0:nxg + 1, 0:nyg + 1, 0:18), gc(0:nxg + 1, 0:nyg + 1, 0:18), pc(0:nxg + 1, 0:nyg + 1))
    f = 0d0
    g = 0d0
    phi = 0d0
    fc = 0d0
    gc = 0d0
    pc = 0d0
    do m = 0, 18
        fp(m + 1) = c_loc(f(0, 0, 0, m))
        gp(m + 1) = c_loc(g(0, 0, 0, m))
    end do
    call check(mflbm_init_fields(p, int(init_option, c_int32_t), z0, c_null_ptr, c_null_ptr, c_null_ptr, fp, gp, &
                                 phi, fc, gc, pc), "mflbm_init_fields")

    ! new run: steps 1 .. n_first, then the checkpoint
    call open_run()
    do ntime = 1, n_first
        call check(mflbm_step(ctx, int(ntime, c_int32_t)), "mflbm_step")
    end do
    call check(mflbm_download_state(ctx, fp, gp, phi, fc, gc, pc), "mflbm_download_state")
    call save_checkpoint(trim(ck1), n_first)
    call mflbm_destroy(ctx)

    ! continued run: a new context from the checkpoint, steps ntime0 .. n_last
    f = -1d0
    g = -1d0
    phi = -1d0
    call load_checkpoint(trim(ck1), ntime0)
    call open_run()
    do ntime = ntime0, n_last
        call check(mflbm_step(ctx, int(ntime, c_int32_t)), "mflbm_step")
    end do
    call check(mflbm_download_state(ctx, fp, gp, phi, fc, gc, pc), "mflbm_download_state")
    call save_checkpoint(trim(ck2), n_last)
    call mflbm_destroy(ctx)
    print '(a, i0, a, i0)', 'hotpath_run ok: steps 1..', n_first, ' then restart to ', n_last

contains

    subroutine check(rc_, what)
        integer(c_int), intent(in) :: rc_
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        if (rc_ /= 0) then
            write(*, '(a, a, a, i0)') 'error: ', what, ' returned ', rc_
            if (c_associated(ctx)) then
                call c_f_pointer(mflbm_last_error(ctx), msg, [512])
                n = 0
                do while (n < 512)
                    if (msg(n + 1) == c_null_char) exit
                    n = n + 1
                end do
                print '(512a)', msg(1:n)
            end if
            stop 1
        end if
    end subroutine check

    ! create the context, upload geometry and state, and the colour gradient of the state (Main_multiphase.F90:120)
    subroutine open_run()
        ctx = c_null_ptr
        call check(mflbm_create(p, ctx), "mflbm_create")
        call check(mflbm_upload_geometry(ctx, wh, ns, s_ijk, s_nfl, s_nbr, s_w, nfl, f_ijk, f_nw, f_theta, w_in), &
                   "mflbm_upload_geometry")
        call check(mflbm_upload_state(ctx, fp, gp, phi, fc, gc, pc), "mflbm_upload_state")
        call check(mflbm_color_gradient(ctx), "mflbm_color_gradient")
    end subroutine open_run

    ! first token is the key, the rest is read list-directed (IO_multiphase.F90:73-278)
    subroutine read_control(path)
        character(len=*), intent(in) :: path
        character(len=1024) :: line, key, rest
        integer :: u, ios, sp
        open(newunit=u, file=path, status='old', action='read')
        do
            read(u, '(a)', iostat=ios) line
            if (ios /= 0) exit
            line = adjustl(line)
            if (len_trim(line) == 0 .or. line(1:1) == '#') cycle
            sp = index(line, ' ')
            key = line(1:sp - 1)
            rest = line(sp + 1:)
            select case (trim(key))
            case ('lattice_dimensions'); read(rest, *) nxg, nyg, nzg
            case ('periodic_indicator'); read(rest, *) iper, jper, kper
            case ('domain_wall_status_x'); read(rest, *) wx
            case ('domain_wall_status_y'); read(rest, *) wy
            case ('domain_wall_status_z'); read(rest, *) wz
            case ('excluded_layers'); read(rest, *) excl
            case ('inlet_BC'); read(rest, *) inlet_bc
            case ('outlet_BC'); read(rest, *) outlet_bc
            case ('porous_plate_cmd'); read(rest, *) plate_cmd
            case ('Z_porous_plate'); read(rest, *) z_plate
            case ('initial_fluid_distribution_option'); read(rest, *) init_option
            case ('fluid1_viscosity'); read(rest, *) nu1
            case ('fluid2_viscosity'); read(rest, *) nu2
            case ('surface_tension'); read(rest, *) gam
            case ('theta'); read(rest, *) theta
            case ('RK_beta'); read(rest, *) beta
            case ('body_force_0'); read(rest, *) force0
            case ('saturation_injection'); read(rest, *) sa_inject
            case ('capillary_number'); read(rest, *) ca0
            case ('initial_interface_position'); read(rest, *) z0
            end select
        end do
        close(u)
    end subroutine read_control

    ! the run constants (Init_multiphase.F90:5-236, IO_multiphase.F90:457-552)
    subroutine derive_params()
        logical :: open_z
        p%nx = nxg
        p%ny = nyg
        p%nz = nzg
        p%jper = jper
        p%kper = kper
        p%wall_x_min = wx(1)
        p%wall_x_max = wx(2)
        p%wall_y_min = wy(1)
        p%wall_y_max = wy(2)
        p%wall_z_min = wz(1)
        p%wall_z_max = wz(2)
        p%inlet_bc = inlet_bc
        p%outlet_bc = outlet_bc
        p%porous_plate_cmd = plate_cmd
        p%z_porous_plate = z_plate
        p%n_exclude_inlet = excl(1)
        p%n_exclude_outlet = excl(2)
        p%precision = 0
        p%la_nui1 = 1d0 / nu1
        p%la_nui2 = 1d0 / nu2
        p%gamma = gam
        p%beta = beta
        p%sa_inject = sa_inject
        p%phi_inlet = 2d0 * sa_inject - 1d0
        p%rho_out = 1d0
        p%rho_in = 0d0
        p%force_z = force0
        p%relaxation = 1d0
        theta_rad = (180d0 - theta) * pi / 180.0d0
        la_x = nxg - 1 - 0.5 - 0.5
        la_y = nyg - 1 - 0.5 - 0.5
        uin_avg = 0d0
        open_z = kper == 0 .and. wz(1) == 0 .and. wz(2) == 0
        if (open_z .and. inlet_bc == 1) then
            p%force_z = 0d0
            uin_avg = ca0 * gam / nu1
        else if (open_z .and. inlet_bc == 2) then
            p%rho_in = p%rho_out - (-force0 / 3d0) * nzg
        end if
        p%uin_avg = uin_avg
    end subroutine derive_params

    ! read_walls and the domain walls of set_walls (Misc.F90:86-109, :247-295)
    subroutine read_wall_file(path)
        character(len=*), intent(in) :: path
        integer :: u, nxs, nys, nzs, i, j, k
        open(newunit=u, file=path, status='old', form='unformatted', access='stream')
        read(u) nxs, nys, nzs
        allocate(sample(nxs, nys, nzs), wg(nxg, nyg, nzg))
        read(u) sample
        close(u)
        wg = 0_c_int8_t
        wg(1:nxs, 1:nys, 1:nzs) = sample
        if (wx(2) == 1 .and. wy(2) == 1) then
            do k = 1, nzg
                do j = 1, nyg
                    do i = 1, nxg
                        if (j >= nys .or. i >= nxs) wg(i, j, k) = 1_c_int8_t
                    end do
                end do
            end do
        end if
        if (wz(1) == 1) wg(:, :, 1) = 1_c_int8_t
        if (wz(2) == 1) wg(:, :, nzg) = 1_c_int8_t
        if (wx(1) == 1) wg(1, :, :) = 1_c_int8_t
        if (wx(2) == 1) wg(nxg, :, :) = 1_c_int8_t
        if (wy(1) == 1) wg(:, 1, :) = 1_c_int8_t
        if (wy(2) == 1) wg(:, nyg, :) = 1_c_int8_t
    end subroutine read_wall_file

    ! save_checkpoint (IO_multiphase.F90:562-642): next step, force_z, rho_in, the PDFs, phi, convective history
    subroutine save_checkpoint(path, nt)
        character(len=*), intent(in) :: path
        integer, intent(in) :: nt
        integer :: u, q
        open(newunit=u, file=path, form='unformatted', status='replace', access='stream')
        write(u) int(nt + 1, c_int32_t), p%force_z, p%rho_in
        do q = 0, 18
            write(u) f(:, :, :, q)
        end do
        do q = 0, 18
            write(u) g(:, :, :, q)
        end do
        write(u) phi
        if (outlet_bc == 1) then
            write(u) fc
            write(u) gc
            write(u) pc
        end if
        close(u)
    end subroutine save_checkpoint

    ! initialization_old_multi (Init_multiphase.F90:477-557)
    subroutine load_checkpoint(path, nt0)
        character(len=*), intent(in) :: path
        integer, intent(out) :: nt0
        integer :: u, q
        integer(c_int32_t) :: n32
        open(newunit=u, file=path, form='unformatted', status='old', access='stream')
        read(u) n32, p%force_z, p%rho_in
        nt0 = n32
        do q = 0, 18
            read(u) f(:, :, :, q)
        end do
        do q = 0, 18
            read(u) g(:, :, :, q)
        end do
        read(u) phi
        if (outlet_bc == 1) then
            read(u) fc
            read(u) gc
            read(u) pc
        end if
        close(u)
    end subroutine load_checkpoint
end program hotpath_run
