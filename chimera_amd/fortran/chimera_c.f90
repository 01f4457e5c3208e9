! ISO_C_BINDING interface to libchimera_hip.so (include/chimera_hip.h) for Fortran
! callers.  The drop-in modules mhd_bc, mhd_derivatives, mhd_flux, mhd_field_ops and
! mhd_change_states in this directory keep the reference's public interfaces and
! forward every call through these bindings; the boundary flags and scheme choices
! the reference reads from module mhd_config are passed explicitly.
module chimera_c
    use iso_c_binding
    use mhd_config
    implicit none
    private
    public :: chimera_bc_t, chimera_scheme_t, current_bc, current_scheme, chimera_check

    type, bind(C) :: chimera_bc_t
        integer(c_int) :: xlo, xhi, ylo, yhi
    end type chimera_bc_t

    type, bind(C) :: chimera_scheme_t
        integer(c_int) :: riemann_solver, slope_limiter, use_slope_limiting, fallback_2_mood
    end type chimera_scheme_t

    public :: c_fill_ghost_cells, c_compute_gradients, c_apply_slope_limiter, &
              c_reconstruction, c_thermal_pressure_check, c_compute_fluxes, &
              c_update_conserved, c_constrained_transport, c_compute_curl_2d, &
              c_compute_divB, c_average_face_to_cell_B, c_get_conserved, &
              c_get_primitive, c_apply_conserved_floors

    interface
        function c_last_error() bind(C, name="chimera_last_error") result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function c_last_error

        integer(c_int) function c_fill_ghost_cells(f, f_pad, f_ambient, nx, ny, bc) &
                bind(C, name="chimera_fill_ghost_cells")
            import :: c_int, c_double, c_ptr, chimera_bc_t
            real(c_double), intent(in) :: f(*)
            real(c_double), intent(inout) :: f_pad(*)
            type(c_ptr), value :: f_ambient
            integer(c_int), value :: nx, ny
            type(chimera_bc_t), intent(in) :: bc
        end function c_fill_ghost_cells

        integer(c_int) function c_compute_gradients(f_pad, dx, nx, ny, f_dx, f_dy) &
                bind(C, name="chimera_compute_gradients")
            import :: c_int, c_double
            real(c_double), intent(in) :: f_pad(*)
            real(c_double), value :: dx
            integer(c_int), value :: nx, ny
            real(c_double), intent(inout) :: f_dx(*), f_dy(*)
        end function c_compute_gradients

        integer(c_int) function c_apply_slope_limiter(f_pad, dx, nx, ny, f_dx, f_dy, lim) &
                bind(C, name="chimera_apply_slope_limiter")
            import :: c_int, c_double
            real(c_double), intent(in) :: f_pad(*)
            real(c_double), value :: dx
            integer(c_int), value :: nx, ny, lim
            real(c_double), intent(inout) :: f_dx(*), f_dy(*)
        end function c_apply_slope_limiter

        integer(c_int) function c_reconstruction(f, f_dx, f_dy, dx, nx, ny, XL, XR, YL, YR, &
                rho_or_p, bc, mood) bind(C, name="chimera_reconstruction")
            import :: c_int, c_double, chimera_bc_t
            real(c_double), intent(in) :: f(*), f_dx(*), f_dy(*)
            real(c_double), value :: dx
            integer(c_int), value :: nx, ny, rho_or_p, mood
            real(c_double), intent(inout) :: XL(*), XR(*), YL(*), YR(*)
            type(chimera_bc_t), intent(in) :: bc
        end function c_reconstruction

        integer(c_int) function c_thermal_pressure_check(nx, ny, &
                P_XL, P_XR, P_YL, P_YR, Bx_XL, Bx_XR, Bx_YL, Bx_YR, By_XL, By_XR, By_YL, By_YR, &
                rho_XL, rho_XR, rho_YL, rho_YR, vx_XL, vx_XR, vx_YL, vx_YR, &
                vy_XL, vy_XR, vy_YL, vy_YR, rho_0, vx_0, vy_0, P_0, Bx_0, By_0) &
                bind(C, name="chimera_thermal_pressure_check")
            import :: c_int, c_double
            integer(c_int), value :: nx, ny
            real(c_double), intent(inout) :: P_XL(*), P_XR(*), P_YL(*), P_YR(*)
            real(c_double), intent(inout) :: Bx_XL(*), Bx_XR(*), Bx_YL(*), Bx_YR(*)
            real(c_double), intent(inout) :: By_XL(*), By_XR(*), By_YL(*), By_YR(*)
            real(c_double), intent(inout) :: rho_XL(*), rho_XR(*), rho_YL(*), rho_YR(*)
            real(c_double), intent(inout) :: vx_XL(*), vx_XR(*), vx_YL(*), vx_YR(*)
            real(c_double), intent(inout) :: vy_XL(*), vy_XR(*), vy_YL(*), vy_YR(*)
            real(c_double), intent(in) :: rho_0(*), vx_0(*), vy_0(*), P_0(*), Bx_0(*), By_0(*)
        end function c_thermal_pressure_check

        integer(c_int) function c_compute_fluxes(rho_L, rho_R, vx_L, vx_R, vy_L, vy_R, P_L, P_R, &
                Bx_L, Bx_R, By_L, By_R, gamma, nx, ny, f_mass, f_momx, f_momy, f_energy, f_by, &
                solver) bind(C, name="chimera_compute_fluxes")
            import :: c_int, c_double
            real(c_double), intent(in) :: rho_L(*), rho_R(*), vx_L(*), vx_R(*), vy_L(*), vy_R(*)
            real(c_double), intent(in) :: P_L(*), P_R(*), Bx_L(*), Bx_R(*), By_L(*), By_R(*)
            real(c_double), value :: gamma
            integer(c_int), value :: nx, ny, solver
            real(c_double), intent(inout) :: f_mass(*), f_momx(*), f_momy(*), f_energy(*), f_by(*)
        end function c_compute_fluxes

        integer(c_int) function c_update_conserved(F, fX, fY, dx, dt, nx, ny, bc) &
                bind(C, name="chimera_update_conserved")
            import :: c_int, c_double, chimera_bc_t
            real(c_double), intent(inout) :: F(*)
            real(c_double), intent(in) :: fX(*), fY(*)
            real(c_double), value :: dx, dt
            integer(c_int), value :: nx, ny
            type(chimera_bc_t), intent(in) :: bc
        end function c_update_conserved

        integer(c_int) function c_constrained_transport(bx, by, fByX, fBxY, dx, dt, nx, ny, bc) &
                bind(C, name="chimera_constrained_transport")
            import :: c_int, c_double, chimera_bc_t
            real(c_double), intent(inout) :: bx(*), by(*)
            real(c_double), intent(in) :: fByX(*), fBxY(*)
            real(c_double), value :: dx, dt
            integer(c_int), value :: nx, ny
            type(chimera_bc_t), intent(in) :: bc
        end function c_constrained_transport

        integer(c_int) function c_compute_curl_2d(Az, dx, nx, ny, bx, by, bc) &
                bind(C, name="chimera_compute_curl_2d")
            import :: c_int, c_double, chimera_bc_t
            real(c_double), intent(in) :: Az(*)
            real(c_double), value :: dx
            integer(c_int), value :: nx, ny
            real(c_double), intent(inout) :: bx(*), by(*)
            type(chimera_bc_t), intent(in) :: bc
        end function c_compute_curl_2d

        integer(c_int) function c_compute_divB(bx, by, dx, nx, ny, divB, bc) &
                bind(C, name="chimera_compute_divB")
            import :: c_int, c_double, chimera_bc_t
            real(c_double), intent(in) :: bx(*), by(*)
            real(c_double), value :: dx
            integer(c_int), value :: nx, ny
            real(c_double), intent(inout) :: divB(*)
            type(chimera_bc_t), intent(in) :: bc
        end function c_compute_divB

        integer(c_int) function c_average_face_to_cell_B(b_x, b_y, nx, ny, Bx, By, bc) &
                bind(C, name="chimera_average_face_to_cell_B")
            import :: c_int, c_double, chimera_bc_t
            real(c_double), intent(in) :: b_x(*), b_y(*)
            integer(c_int), value :: nx, ny
            real(c_double), intent(inout) :: Bx(*), By(*)
            type(chimera_bc_t), intent(in) :: bc
        end function c_average_face_to_cell_B

        integer(c_int) function c_get_conserved(rho, vx, vy, P, Bx, By, gamma, vol, nx, ny, &
                Mass, Momx, Momy, Energy) bind(C, name="chimera_get_conserved")
            import :: c_int, c_double
            real(c_double), intent(in) :: rho(*), vx(*), vy(*), P(*), Bx(*), By(*)
            real(c_double), value :: gamma, vol
            integer(c_int), value :: nx, ny
            real(c_double), intent(inout) :: Mass(*), Momx(*), Momy(*), Energy(*)
        end function c_get_conserved

        integer(c_int) function c_get_primitive(Mass, Momx, Momy, Energy, Bx, By, gamma, vol, &
                nx, ny, rho, vx, vy, P) bind(C, name="chimera_get_primitive")
            import :: c_int, c_double
            real(c_double), intent(in) :: Mass(*), Momx(*), Momy(*), Energy(*), Bx(*), By(*)
            real(c_double), value :: gamma, vol
            integer(c_int), value :: nx, ny
            real(c_double), intent(inout) :: rho(*), vx(*), vy(*), P(*)
        end function c_get_primitive

        integer(c_int) function c_apply_conserved_floors(Mass, Momx, Momy, Energy, Bx, By, vol, &
                gamma, nx, ny) bind(C, name="chimera_apply_conserved_floors")
            import :: c_int, c_double
            real(c_double), intent(inout) :: Mass(*), Momx(*), Momy(*), Energy(*)
            real(c_double), intent(in) :: Bx(*), By(*)
            real(c_double), value :: vol, gamma
            integer(c_int), value :: nx, ny
        end function c_apply_conserved_floors
    end interface

contains

    ! The per-side flags the reference keeps in mhd_config (mhd_config.f90:48-49).
    function current_bc() result(bc)
        type(chimera_bc_t) :: bc
        bc = chimera_bc_t(BC_xlo, BC_xhi, BC_ylo, BC_yhi)
    end function current_bc

    ! The compile-time scheme parameters of mhd_config (mhd_config.f90:27-37).
    function current_scheme() result(s)
        type(chimera_scheme_t) :: s
        s = chimera_scheme_t(riemann_solver, slope_limiter, merge(1, 0, useSlopeLimiting), &
                             merge(1, 0, fallback_2_MOOD))
    end function current_scheme

    ! The reference has no status codes; a failed GPU call ends the program with the
    ! library's message (error stop), as the reference does for fatal input.
    subroutine chimera_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer :: k
        if (rc == 0) return
        call c_f_pointer(c_last_error(), msg, [512])
        text = ''
        do k = 1, 512
            if (msg(k) == c_null_char) exit
            text(k:k) = msg(k)
        end do
        write(*, '(A, A, A, I0, A, A)') 'chimera: ', what, ' failed (status ', rc, '): ', trim(text)
        error stop 1
    end subroutine chimera_check

end module chimera_c
