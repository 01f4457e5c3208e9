! Drop-in for the reference module mhd_change_states (src/mhd_change_states.f90):
! same public interface, each routine on the GPU.
module mhd_change_states
    use iso_c_binding
    use mhd_config
    use chimera_c
    implicit none
    private
    public get_conserved, get_primitive, apply_conserved_floors

contains

    ! src/mhd_change_states.f90:16-58
    subroutine get_conserved(rho, vx, vy, P, Bx, By, gamma, vol, nx, ny, &
                             Mass, Momx, Momy, Energy)
        integer, intent(in) :: nx, ny
        real(8), intent(in) :: rho(nx, ny), vx(nx, ny), vy(nx, ny)
        real(8), intent(in) :: P(nx, ny), Bx(nx, ny), By(nx, ny)
        real(8), intent(in) :: gamma, vol
        real(8), intent(out) :: Mass(nx,ny), Momx(nx,ny), Momy(nx,ny), Energy(nx,ny)
        call chimera_check(c_get_conserved(rho, vx, vy, P, Bx, By, gamma, vol, nx, ny, &
                                           Mass, Momx, Momy, Energy), 'get_conserved')
    end subroutine get_conserved

    ! src/mhd_change_states.f90:61-113
    subroutine get_primitive(Mass, Momx, Momy, Energy, Bx, By, gamma, vol, nx, ny, &
                             rho, vx, vy, P)
        integer, intent(in) :: nx, ny
        real(8), intent(in) :: Mass(nx, ny), Momx(nx, ny), Momy(nx, ny)
        real(8), intent(in) :: Energy(nx, ny), Bx(nx, ny), By(nx, ny)
        real(8), intent(in) :: gamma, vol
        real(8), intent(out) :: rho(nx, ny), vx(nx, ny), vy(nx, ny), P(nx, ny)
        call chimera_check(c_get_primitive(Mass, Momx, Momy, Energy, Bx, By, gamma, vol, nx, ny, &
                                           rho, vx, vy, P), 'get_primitive')
    end subroutine get_primitive

    ! src/mhd_change_states.f90:116-166
    subroutine apply_conserved_floors(Mass, Momx, Momy, Energy, Bx, By, vol, gamma, nx, ny)
        integer, intent(in)    :: nx, ny
        real(8), intent(inout) :: Mass(nx,ny), Momx(nx,ny), Momy(nx,ny)
        real(8), intent(inout) :: Energy(nx,ny)
        real(8), intent(in)    :: Bx(nx,ny), By(nx,ny)
        real(8), intent(in)    :: vol, gamma
        call chimera_check(c_apply_conserved_floors(Mass, Momx, Momy, Energy, Bx, By, vol, gamma, &
                                                    nx, ny), 'apply_conserved_floors')
    end subroutine apply_conserved_floors

end module mhd_change_states
