! Drop-in for the reference module mhd_bc (src/mhd_bc.f90): same public interface,
! ghost fill on the GPU through chimera_fill_ghost_cells.
module mhd_bc
    use iso_c_binding
    use mhd_config
    use chimera_c
    implicit none
    private
    public fill_ghost_cells

contains

    ! src/mhd_bc.f90:36-141 -- (N+2) x (N+2) pad of an N x N field (module N).
    subroutine fill_ghost_cells(f, f_pad, f_ambient)
        real(8), intent(in)  :: f(N, N)
        real(8), intent(out) :: f_pad(N+2, N+2)
        real(8), intent(in), optional, target :: f_ambient(N, N)
        type(c_ptr) :: amb
        amb = c_null_ptr
        if (present(f_ambient)) amb = c_loc(f_ambient)
        call chimera_check(c_fill_ghost_cells(f, f_pad, amb, N, N, current_bc()), 'fill_ghost_cells')
    end subroutine fill_ghost_cells

end module mhd_bc
