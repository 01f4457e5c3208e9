! Drop-in for the reference module mhd_write_h5 (src/mhd_write_h5.f90): same public
! subroutine, writing the same file layout through libchimera_h5 (HDF5 C API) instead
! of the Fortran HDF5 module.
module mhd_write_h5
    use iso_c_binding
    implicit none
    private
    public write_prims_to_hdf5

    interface
        integer(c_int) function c_h5_write_snapshots(path, nsnap, nx, ny, rho, P, Bx, By, Vx, Vy, &
                                                     times, gamma, M_s, beta) &
                bind(C, name="chimera_h5_write_snapshots")
            import :: c_int, c_double, c_char
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: nsnap, nx, ny
            real(c_double), intent(in) :: rho(*), P(*), Bx(*), By(*), Vx(*), Vy(*), times(*)
            real(c_double), value :: gamma, M_s, beta
        end function
    end interface

contains

    ! src/mhd_write_h5.f90:12-177.  The reference's arrays are (snapshot, i, j); the C
    ! writer takes [snapshot][j][i], i.e. Fortran (i, j, snapshot).
    subroutine write_prims_to_hdf5(filename, rho_dat, P_dat, Bx_dat, By_dat, Vx_dat, Vy_dat, &
                                   time_dat, gamma_val, M_s_val, beta_val)
        character(len=*), intent(in) :: filename
        real(kind=8), dimension(:,:,:), intent(in) :: rho_dat, P_dat, Bx_dat, By_dat, Vx_dat, Vy_dat
        real(kind=8), dimension(:), intent(in) :: time_dat
        real(kind=8), intent(in) :: gamma_val, M_s_val, beta_val
        real(kind=8), allocatable :: buf(:,:,:,:)
        integer :: ns, nx, ny, s, rc

        ns = size(rho_dat, 1)
        nx = size(rho_dat, 2)
        ny = size(rho_dat, 3)
        allocate(buf(nx, ny, max(ns, 1), 6))
        do s = 1, ns
            buf(:, :, s, 1) = rho_dat(s, :, :)
            buf(:, :, s, 2) = P_dat(s, :, :)
            buf(:, :, s, 3) = Bx_dat(s, :, :)
            buf(:, :, s, 4) = By_dat(s, :, :)
            buf(:, :, s, 5) = Vx_dat(s, :, :)
            buf(:, :, s, 6) = Vy_dat(s, :, :)
        end do
        rc = c_h5_write_snapshots(trim(filename)//c_null_char, ns, nx, ny, buf(:, :, :, 1), &
                                  buf(:, :, :, 2), buf(:, :, :, 3), buf(:, :, :, 4), &
                                  buf(:, :, :, 5), buf(:, :, :, 6), time_dat, gamma_val, &
                                  M_s_val, beta_val)
        deallocate(buf)
        if (rc /= 0) then
            print *, "Error opening file"
            stop
        end if
    end subroutine write_prims_to_hdf5

end module mhd_write_h5
