! Drop-in for the reference module mhd_flux (src/mhd_flux.f90): same public
! interface, each routine on the GPU; the Riemann solver comes from mhd_config.
module mhd_flux
    use iso_c_binding
    use mhd_config
    use chimera_c
    implicit none
    private
    public update_conserved, constrained_transport, compute_fluxes

contains

    ! src/mhd_flux.f90:20-76
    subroutine update_conserved(F, flux_F_X, flux_F_Y, dx, dt, nx, ny)
        integer, intent(in) :: nx, ny
        real(8), intent(in) :: dx, dt
        real(8), intent(in) :: flux_F_X(nx, ny), flux_F_Y(nx, ny)
        real(8), intent(inout) :: F(nx, ny)
        call chimera_check(c_update_conserved(F, flux_F_X, flux_F_Y, dx, dt, nx, ny, current_bc()), &
                           'update_conserved')
    end subroutine update_conserved

    ! src/mhd_flux.f90:79-136
    subroutine constrained_transport(bx, by, flux_By_X, flux_Bx_Y, dx, dt, nx, ny)
        integer, intent(in) :: nx, ny
        real(8), intent(in) :: dx, dt
        real(8), intent(in) :: flux_By_X(nx, ny), flux_Bx_Y(nx, ny)
        real(8), intent(inout) :: bx(nx, ny), by(nx, ny)
        call chimera_check(c_constrained_transport(bx, by, flux_By_X, flux_Bx_Y, dx, dt, nx, ny, &
                                                   current_bc()), 'constrained_transport')
    end subroutine constrained_transport

    ! src/mhd_flux.f90:139-191
    subroutine compute_fluxes(rho_L, rho_R, vx_L, vx_R, vy_L, vy_R, P_L, P_R,   &
                              Bx_L, Bx_R, By_L, By_R, gamma, nx, ny,            &
                              flux_Mass, flux_Momx, flux_Momy, flux_Energy, flux_By)
        integer, intent(in)    :: nx, ny
        real(8), intent(inout) :: rho_L(nx, ny), rho_R(nx, ny)
        real(8), intent(in)    :: vx_L(nx, ny), vx_R(nx, ny)
        real(8), intent(in)    :: vy_L(nx, ny), vy_R(nx, ny)
        real(8), intent(inout) :: P_L(nx, ny), P_R(nx, ny)
        real(8), intent(in)    :: Bx_L(nx, ny), Bx_R(nx, ny)
        real(8), intent(in)    :: By_L(nx, ny), By_R(nx, ny)
        real(8), intent(in)    :: gamma
        real(8), intent(out)   :: flux_Mass(nx, ny), flux_Momx(nx, ny), flux_Momy(nx, ny)
        real(8), intent(out)   :: flux_Energy(nx, ny), flux_By(nx, ny)
        call chimera_check(c_compute_fluxes(rho_L, rho_R, vx_L, vx_R, vy_L, vy_R, P_L, P_R, &
                                            Bx_L, Bx_R, By_L, By_R, gamma, nx, ny, flux_Mass, &
                                            flux_Momx, flux_Momy, flux_Energy, flux_By, &
                                            riemann_solver), 'compute_fluxes')
    end subroutine compute_fluxes

end module mhd_flux
