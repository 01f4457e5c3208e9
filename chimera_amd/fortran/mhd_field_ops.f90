! Drop-in for the reference module mhd_field_ops (src/mhd_field_ops.f90): same public
! interface, each routine on the GPU.
module mhd_field_ops
    use iso_c_binding
    use mhd_config
    use chimera_c
    implicit none
    private
    public compute_curl_2d, compute_divB, average_face_to_cell_B

contains

    ! src/mhd_field_ops.f90:15-57
    subroutine compute_curl_2d(Az, dx, nx, ny, bx, by)
        integer, intent(in)  :: nx, ny
        real(8), intent(in)  :: dx, Az(nx, ny)
        real(8), intent(out) :: bx(nx, ny), by(nx, ny)
        call chimera_check(c_compute_curl_2d(Az, dx, nx, ny, bx, by, current_bc()), 'compute_curl_2d')
    end subroutine compute_curl_2d

    ! src/mhd_field_ops.f90:60-117
    subroutine compute_divB(bx, by, dx, nx, ny, divB)
        integer, intent(in)  :: nx, ny
        real(8), intent(in)  :: dx, bx(nx, ny), by(nx, ny)
        real(8), intent(out) :: divB(nx, ny)
        call chimera_check(c_compute_divB(bx, by, dx, nx, ny, divB, current_bc()), 'compute_divB')
    end subroutine compute_divB

    ! src/mhd_field_ops.f90:120-156
    subroutine average_face_to_cell_B(b_x, b_y, nx, ny, Bx, By)
        integer, intent(in)  :: nx, ny
        real(8), intent(in)  :: b_x(nx, ny), b_y(nx, ny)
        real(8), intent(out) :: Bx(nx, ny),  By(nx, ny)
        call chimera_check(c_average_face_to_cell_B(b_x, b_y, nx, ny, Bx, By, current_bc()), &
                           'average_face_to_cell_B')
    end subroutine average_face_to_cell_B

end module mhd_field_ops
