! Drop-in for the reference module mhd_derivatives (src/mhd_derivatives.f90): same
! public interface, each routine on the GPU.
module mhd_derivatives
    use iso_c_binding
    use mhd_config
    use chimera_c
    implicit none
    private
    public compute_gradients, apply_slope_limiter, reconstruction, thermal_pressure_check

contains

    ! src/mhd_derivatives.f90:15-47
    subroutine compute_gradients(f_pad, dx, nx, ny, f_dx, f_dy)
        integer, intent(in)  :: nx, ny
        real(8), intent(in)  :: f_pad(nx+2, ny+2)
        real(8), intent(in)  :: dx
        real(8), intent(out) :: f_dx(nx, ny), f_dy(nx, ny)
        call chimera_check(c_compute_gradients(f_pad, dx, nx, ny, f_dx, f_dy), 'compute_gradients')
    end subroutine compute_gradients

    ! src/mhd_derivatives.f90:50-127 (limiter from mhd_config)
    subroutine apply_slope_limiter(f_pad, dx, nx, ny, f_dx, f_dy)
        integer, intent(in)  :: nx, ny
        real(8), intent(in)  :: dx
        real(8), intent(in)  :: f_pad(nx+2, ny+2)
        real(8), intent(out) :: f_dx(nx, ny), f_dy(nx, ny)
        call chimera_check(c_apply_slope_limiter(f_pad, dx, nx, ny, f_dx, f_dy, slope_limiter), &
                           'apply_slope_limiter')
    end subroutine apply_slope_limiter

    ! src/mhd_derivatives.f90:130-229 (BC flags and MOOD switch from mhd_config)
    subroutine reconstruction(f, f_dx, f_dy, dx, nx, ny, f_XL, f_XR, f_YL, f_YR, rho_or_p)
        integer, intent(in)  :: nx, ny
        real(8), intent(in)  :: dx
        real(8), intent(in)  :: f(nx, ny), f_dx(nx, ny), f_dy(nx, ny)
        real(8), intent(out) :: f_XL(nx, ny), f_XR(nx, ny)
        real(8), intent(out) :: f_YL(nx, ny), f_YR(nx, ny)
        logical, intent(in)  :: rho_or_p
        call chimera_check(c_reconstruction(f, f_dx, f_dy, dx, nx, ny, f_XL, f_XR, f_YL, f_YR, &
                                            merge(1, 0, rho_or_p), current_bc(), &
                                            merge(1, 0, fallback_2_MOOD)), 'reconstruction')
    end subroutine reconstruction

    ! src/mhd_derivatives.f90:232-293
    subroutine thermal_pressure_check(nx, ny,                        &
                                       P_XL,  P_XR,  P_YL,  P_YR,   &
                                       Bx_XL, Bx_XR, Bx_YL, Bx_YR, &
                                       By_XL, By_XR, By_YL, By_YR, &
                                       rho_XL, rho_XR, rho_YL, rho_YR, &
                                       vx_XL,  vx_XR,  vx_YL,  vx_YR,  &
                                       vy_XL,  vy_XR,  vy_YL,  vy_YR,  &
                                       rho_0, vx_0, vy_0, P_0, Bx_0, By_0)
        integer, intent(in) :: nx, ny
        real(8), intent(inout) :: P_XL(nx,ny),   P_XR(nx,ny),   P_YL(nx,ny),   P_YR(nx,ny)
        real(8), intent(inout) :: Bx_XL(nx,ny),  Bx_XR(nx,ny),  Bx_YL(nx,ny),  Bx_YR(nx,ny)
        real(8), intent(inout) :: By_XL(nx,ny),  By_XR(nx,ny),  By_YL(nx,ny),  By_YR(nx,ny)
        real(8), intent(inout) :: rho_XL(nx,ny), rho_XR(nx,ny), rho_YL(nx,ny), rho_YR(nx,ny)
        real(8), intent(inout) :: vx_XL(nx,ny),  vx_XR(nx,ny),  vx_YL(nx,ny),  vx_YR(nx,ny)
        real(8), intent(inout) :: vy_XL(nx,ny),  vy_XR(nx,ny),  vy_YL(nx,ny),  vy_YR(nx,ny)
        real(8), intent(in)    :: rho_0(nx,ny), vx_0(nx,ny), vy_0(nx,ny)
        real(8), intent(in)    :: P_0(nx,ny),   Bx_0(nx,ny), By_0(nx,ny)
        call chimera_check(c_thermal_pressure_check(nx, ny, P_XL, P_XR, P_YL, P_YR, &
                Bx_XL, Bx_XR, Bx_YL, Bx_YR, By_XL, By_XR, By_YL, By_YR, &
                rho_XL, rho_XR, rho_YL, rho_YR, vx_XL, vx_XR, vx_YL, vx_YR, &
                vy_XL, vy_XR, vy_YL, vy_YR, rho_0, vx_0, vy_0, P_0, Bx_0, By_0), &
                'thermal_pressure_check')
    end subroutine thermal_pressure_check

end module mhd_derivatives
