! Link check of the ISO_C_BINDING interface: the Fortran side builds a parameter block and a wall array
! in its own layout and calls a host set-up entry point of the C ABI; prints the checksum of the result.
program binding_check
    use, intrinsic :: iso_c_binding
    use mflbm_c_binding
    implicit none
    integer, parameter :: nx = 8, ny = 6, nz = 5
    type(mflbm_params) :: p
    integer(c_int8_t) :: wg(nx, ny, nz)
    integer(c_int8_t), allocatable :: wh(:, :, :)
    integer :: rc, i, j, k
    integer(c_int64_t) :: s
    p%nx = nx; p%ny = ny; p%nz = nz
    p%jper = 0; p%kper = 1
    p%wall_x_min = 1; p%wall_x_max = 1; p%wall_y_min = 1; p%wall_y_max = 1
    p%wall_z_min = 0; p%wall_z_max = 0
    wg = 0_c_int8_t
    wg(4, 3, 2) = 1_c_int8_t
    allocate(wh(-1:nx+2, -1:ny+2, -1:nz+2))
    rc = mflbm_setup_walls(p, wg, wh)
    s = 0
    do k = -1, nz + 2
        do j = -1, ny + 2
            do i = -1, nx + 2
                s = s + int(wh(i, j, k), c_int64_t) * (i + 3 * j + 7 * k + 100)
            end do
        end do
    end do
    print '(a, i0, a, i0)', 'rc=', rc, ' checksum=', s
    call sp_check()
contains
    ! single-phase initial populations through the Fortran interface: f_q(i,j,k) at a velocity u = 0.01 i
    subroutine sp_check()
        type(mflbm_sp_params) :: q
        real(c_double), allocatable, target :: f(:, :, :, :), u(:, :, :)
        type(c_ptr) :: fp(19)
        integer :: m, rc2
        real(c_double) :: t
        q%nx = nx; q%ny = ny; q%nz = nz
        q%jper = 0; q%kper = 1
        q%wall_x_min = 1; q%wall_x_max = 1; q%wall_y_min = 1; q%wall_y_max = 1
        q%wall_z_min = 0; q%wall_z_max = 0
        q%inlet_bc = 0; q%outlet_bc = 0; q%mrt_preset = 1; q%precision = 0; q%la_nu = 0.1d0
        allocate(f(0:nx+1, 0:ny+1, 0:nz+1, 0:18), u(0:nx+1, 0:ny+1, 0:nz+1))
        do i = 0, nx + 1
            u(i, :, :) = 0.01d0 * i
        end do
        do m = 0, 18
            fp(m + 1) = c_loc(f(0, 0, 0, m))
        end do
        rc2 = mflbm_sp_init_fields(q, c_null_ptr, c_loc(u), c_null_ptr, c_null_ptr, fp, c_null_ptr)
        t = 0d0
        do m = 0, 18
            t = t + sum(f(:, :, :, m)) * (m + 1)
        end do
        print '(a, i0, a, es24.16)', 'sp_rc=', rc2, ' sp_sum=', t
    end subroutine sp_check
end program binding_check
