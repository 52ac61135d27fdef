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
end program binding_check
