! Harness over the reference CHIMERA hot-path modules -- TEST INFRASTRUCTURE ONLY.
!
! oracle/build_ref.sh compiles this file together with the UNMODIFIED reference
! modules mhd_field_ops, mhd_change_states, mhd_derivatives, mhd_bc and mhd_flux
! (from /root/reference/src) and a scheme variant of mhd_config (the compile-time
! riemann_solver / slope_limiter parameters substituted) into
! oracle/_ref/libchimera_ref_s<solver>_l<limiter>.so.
!
! Each ref_* entry point is a bind(C) wrapper around one reference routine, so
! the C oracle (oracle/chimera_oracle.c) can be compared with the reference
! routine by routine.  ref_run restates the time loop of src/main.f90:65-303 and
! compute_stage_fluxes (main.f90:346-403) on top of the reference routines; the
! reference main program itself cannot be linked here (its HDF5 writer needs a
! Fortran hdf5.mod the image does not have).
!
! Arrays are square N x N (the reference's fill_ghost_cells is sized by the
! module N); ref_set_grid sets N and the per-side BC flags.
module ref_harness
    use iso_c_binding
    use mhd_config
    use mhd_bc
    use mhd_derivatives
    use mhd_flux
    use mhd_field_ops
    use mhd_change_states
    implicit none

contains

    subroutine ref_set_grid(n_sq, bc) bind(C, name="ref_set_grid")
        integer(c_int), value      :: n_sq
        integer(c_int), intent(in) :: bc(4)
        N = n_sq
        BC_xlo = bc(1); BC_xhi = bc(2); BC_ylo = bc(3); BC_yhi = bc(4)
    end subroutine ref_set_grid

    ! Scheme the variant was built with: {riemann_solver, slope_limiter,
    ! useSlopeLimiting, fallback_2_MOOD}.
    subroutine ref_scheme(out) bind(C, name="ref_scheme")
        integer(c_int), intent(out) :: out(4)
        out(1) = riemann_solver
        out(2) = slope_limiter
        out(3) = merge(1, 0, useSlopeLimiting)
        out(4) = merge(1, 0, fallback_2_MOOD)
    end subroutine ref_scheme

    subroutine ref_fill_ghost_cells(f, f_pad) bind(C, name="ref_fill_ghost_cells")
        real(c_double), intent(in)  :: f(N, N)
        real(c_double), intent(out) :: f_pad(N+2, N+2)
        call fill_ghost_cells(f, f_pad)
    end subroutine ref_fill_ghost_cells

    subroutine ref_fill_ghost_cells_amb(f, f_pad, f_amb) bind(C, name="ref_fill_ghost_cells_amb")
        real(c_double), intent(in)  :: f(N, N), f_amb(N, N)
        real(c_double), intent(out) :: f_pad(N+2, N+2)
        call fill_ghost_cells(f, f_pad, f_amb)
    end subroutine ref_fill_ghost_cells_amb

    subroutine ref_compute_gradients(f_pad, dx, f_dx, f_dy) bind(C, name="ref_compute_gradients")
        real(c_double), intent(in)  :: f_pad(N+2, N+2)
        real(c_double), value       :: dx
        real(c_double), intent(out) :: f_dx(N, N), f_dy(N, N)
        call compute_gradients(f_pad, dx, N, N, f_dx, f_dy)
    end subroutine ref_compute_gradients

    subroutine ref_apply_slope_limiter(f_pad, dx, f_dx, f_dy) bind(C, name="ref_apply_slope_limiter")
        real(c_double), intent(in)  :: f_pad(N+2, N+2)
        real(c_double), value       :: dx
        real(c_double), intent(out) :: f_dx(N, N), f_dy(N, N)
        call apply_slope_limiter(f_pad, dx, N, N, f_dx, f_dy)
    end subroutine ref_apply_slope_limiter

    subroutine ref_reconstruction(f, f_dx, f_dy, dx, XL, XR, YL, YR, rho_or_p) &
            bind(C, name="ref_reconstruction")
        real(c_double), intent(in)  :: f(N, N), f_dx(N, N), f_dy(N, N)
        real(c_double), value       :: dx
        real(c_double), intent(out) :: XL(N, N), XR(N, N), YL(N, N), YR(N, N)
        integer(c_int), value       :: rho_or_p
        call reconstruction(f, f_dx, f_dy, dx, N, N, XL, XR, YL, YR, rho_or_p /= 0)
    end subroutine ref_reconstruction

    subroutine ref_thermal_pressure_check(P4, Bx4, By4, rho4, vx4, vy4, cells) &
            bind(C, name="ref_thermal_pressure_check")
        real(c_double), intent(inout) :: P4(N, N, 4), Bx4(N, N, 4), By4(N, N, 4)
        real(c_double), intent(inout) :: rho4(N, N, 4), vx4(N, N, 4), vy4(N, N, 4)
        real(c_double), intent(in)    :: cells(N, N, 6)   ! rho, vx, vy, P, Bx, By
        call thermal_pressure_check(N, N, &
            P4(:,:,1),   P4(:,:,2),   P4(:,:,3),   P4(:,:,4),   &
            Bx4(:,:,1),  Bx4(:,:,2),  Bx4(:,:,3),  Bx4(:,:,4),  &
            By4(:,:,1),  By4(:,:,2),  By4(:,:,3),  By4(:,:,4),  &
            rho4(:,:,1), rho4(:,:,2), rho4(:,:,3), rho4(:,:,4), &
            vx4(:,:,1),  vx4(:,:,2),  vx4(:,:,3),  vx4(:,:,4),  &
            vy4(:,:,1),  vy4(:,:,2),  vy4(:,:,3),  vy4(:,:,4),  &
            cells(:,:,1), cells(:,:,2), cells(:,:,3), cells(:,:,4), cells(:,:,5), cells(:,:,6))
    end subroutine ref_thermal_pressure_check

    ! in(:,:,1:12) = rho_L,rho_R,vx_L,vx_R,vy_L,vy_R,P_L,P_R,Bx_L,Bx_R,By_L,By_R
    ! out(:,:,1:5) = Mass, Momx, Momy, Energy, By
    subroutine ref_compute_fluxes(nx, ny, fin, gamma, fout) bind(C, name="ref_compute_fluxes")
        integer(c_int), value       :: nx, ny
        real(c_double), intent(in)  :: fin(nx, ny, 12)
        real(c_double), value       :: gamma
        real(c_double), intent(out) :: fout(nx, ny, 5)
        real(c_double) :: rL(nx, ny), rR(nx, ny), pL(nx, ny), pR(nx, ny)
        rL = fin(:,:,1); rR = fin(:,:,2); pL = fin(:,:,7); pR = fin(:,:,8)
        call compute_fluxes(rL, rR, fin(:,:,3), fin(:,:,4), fin(:,:,5), fin(:,:,6), pL, pR, &
                            fin(:,:,9), fin(:,:,10), fin(:,:,11), fin(:,:,12), gamma, nx, ny, &
                            fout(:,:,1), fout(:,:,2), fout(:,:,3), fout(:,:,4), fout(:,:,5))
    end subroutine ref_compute_fluxes

    subroutine ref_update_conserved(F, fX, fY, dx, dt) bind(C, name="ref_update_conserved")
        real(c_double), intent(inout) :: F(N, N)
        real(c_double), intent(in)    :: fX(N, N), fY(N, N)
        real(c_double), value         :: dx, dt
        call update_conserved(F, fX, fY, dx, dt, N, N)
    end subroutine ref_update_conserved

    subroutine ref_constrained_transport(bx, by, fByX, fBxY, dx, dt) &
            bind(C, name="ref_constrained_transport")
        real(c_double), intent(inout) :: bx(N, N), by(N, N)
        real(c_double), intent(in)    :: fByX(N, N), fBxY(N, N)
        real(c_double), value         :: dx, dt
        call constrained_transport(bx, by, fByX, fBxY, dx, dt, N, N)
    end subroutine ref_constrained_transport

    subroutine ref_compute_curl_2d(Az, dx, bx, by) bind(C, name="ref_compute_curl_2d")
        real(c_double), intent(in)  :: Az(N, N)
        real(c_double), value       :: dx
        real(c_double), intent(out) :: bx(N, N), by(N, N)
        call compute_curl_2d(Az, dx, N, N, bx, by)
    end subroutine ref_compute_curl_2d

    subroutine ref_compute_divB(bx, by, dx, divB) bind(C, name="ref_compute_divB")
        real(c_double), intent(in)  :: bx(N, N), by(N, N)
        real(c_double), value       :: dx
        real(c_double), intent(out) :: divB(N, N)
        call compute_divB(bx, by, dx, N, N, divB)
    end subroutine ref_compute_divB

    subroutine ref_average_face_to_cell_B(b_x, b_y, Bx, By) bind(C, name="ref_average_face_to_cell_B")
        real(c_double), intent(in)  :: b_x(N, N), b_y(N, N)
        real(c_double), intent(out) :: Bx(N, N), By(N, N)
        call average_face_to_cell_B(b_x, b_y, N, N, Bx, By)
    end subroutine ref_average_face_to_cell_B

    subroutine ref_get_conserved(prim, gamma, vol, cons) bind(C, name="ref_get_conserved")
        real(c_double), intent(in)  :: prim(N, N, 6)   ! rho, vx, vy, P, Bx, By
        real(c_double), value       :: gamma, vol
        real(c_double), intent(out) :: cons(N, N, 4)   ! Mass, Momx, Momy, Energy
        call get_conserved(prim(:,:,1), prim(:,:,2), prim(:,:,3), prim(:,:,4), prim(:,:,5), &
                           prim(:,:,6), gamma, vol, N, N, cons(:,:,1), cons(:,:,2), cons(:,:,3), &
                           cons(:,:,4))
    end subroutine ref_get_conserved

    subroutine ref_get_primitive(cons, Bx, By, gamma, vol, prim) bind(C, name="ref_get_primitive")
        real(c_double), intent(in)  :: cons(N, N, 4), Bx(N, N), By(N, N)
        real(c_double), value       :: gamma, vol
        real(c_double), intent(out) :: prim(N, N, 4)   ! rho, vx, vy, P
        call get_primitive(cons(:,:,1), cons(:,:,2), cons(:,:,3), cons(:,:,4), Bx, By, gamma, vol, &
                           N, N, prim(:,:,1), prim(:,:,2), prim(:,:,3), prim(:,:,4))
    end subroutine ref_get_primitive

    subroutine ref_apply_conserved_floors(cons, Bx, By, vol, gamma) &
            bind(C, name="ref_apply_conserved_floors")
        real(c_double), intent(inout) :: cons(N, N, 4)
        real(c_double), intent(in)    :: Bx(N, N), By(N, N)
        real(c_double), value         :: vol, gamma
        call apply_conserved_floors(cons(:,:,1), cons(:,:,2), cons(:,:,3), cons(:,:,4), Bx, By, &
                                    vol, gamma, N, N)
    end subroutine ref_apply_conserved_floors

    ! L(u) from the stage primitives prim(:,:,1:6) = rho, vx, vy, P, Bx, By.
    ! flux(:,:,1:10) = Mass_X, Mass_Y, Momx_X, Momx_Y, Momy_X, Momy_Y,
    !                  Energy_X, Energy_Y, By_X, Bx_Y.
    ! faces(:,:,4*(f-1)+s) returns the face states after reconstruction and the
    ! thermal check (f = rho..By, s = XL, XR, YL, YR).
    subroutine ref_stage_fluxes(prim, dx, gamma, flux, faces) bind(C, name="ref_stage_fluxes")
        real(c_double), intent(in)  :: prim(N, N, 6)
        real(c_double), value       :: dx, gamma
        real(c_double), intent(out) :: flux(N, N, 10)
        real(c_double), intent(out) :: faces(N, N, 24)
        real(c_double), allocatable :: pad(:,:), sx(:,:), sy(:,:)
        integer :: f
        allocate(pad(N+2, N+2), sx(N, N), sy(N, N))
        do f = 1, 6
            call fill_ghost_cells(prim(:,:,f), pad)
            call compute_gradients(pad, dx, N, N, sx, sy)
            if (useSlopeLimiting) call apply_slope_limiter(pad, dx, N, N, sx, sy)
            call reconstruction(prim(:,:,f), sx, sy, dx, N, N, faces(:,:,4*f-3), faces(:,:,4*f-2), &
                                faces(:,:,4*f-1), faces(:,:,4*f), (f == 1 .or. f == 4))
        end do
        ! thermal_pressure_check argument groups: P, Bx, By, rho, vx, vy
        call thermal_pressure_check(N, N, &
            faces(:,:,13), faces(:,:,14), faces(:,:,15), faces(:,:,16), &
            faces(:,:,17), faces(:,:,18), faces(:,:,19), faces(:,:,20), &
            faces(:,:,21), faces(:,:,22), faces(:,:,23), faces(:,:,24), &
            faces(:,:,1),  faces(:,:,2),  faces(:,:,3),  faces(:,:,4),  &
            faces(:,:,5),  faces(:,:,6),  faces(:,:,7),  faces(:,:,8),  &
            faces(:,:,9),  faces(:,:,10), faces(:,:,11), faces(:,:,12), &
            prim(:,:,1), prim(:,:,2), prim(:,:,3), prim(:,:,4), prim(:,:,5), prim(:,:,6))
        ! x faces: normal (vx, Bx)
        call compute_fluxes(faces(:,:,1), faces(:,:,2), faces(:,:,5), faces(:,:,6), &
                            faces(:,:,9), faces(:,:,10), faces(:,:,13), faces(:,:,14), &
                            faces(:,:,17), faces(:,:,18), faces(:,:,21), faces(:,:,22), gamma, N, N, &
                            flux(:,:,1), flux(:,:,3), flux(:,:,5), flux(:,:,7), flux(:,:,9))
        ! y faces, rotated: normal (vy, By), tangential (vx, Bx)
        call compute_fluxes(faces(:,:,3), faces(:,:,4), faces(:,:,11), faces(:,:,12), &
                            faces(:,:,7), faces(:,:,8), faces(:,:,15), faces(:,:,16), &
                            faces(:,:,23), faces(:,:,24), faces(:,:,19), faces(:,:,20), gamma, N, N, &
                            flux(:,:,2), flux(:,:,6), flux(:,:,4), flux(:,:,8), flux(:,:,10))
        deallocate(pad, sx, sy)
    end subroutine ref_stage_fluxes

    ! Time loop of src/main.f90:65-303 on state(:,:,1:6) = Mass, Momx, Momy,
    ! Energy, b_x, b_y.  dt_mode 0: CFL with the output-time clamp and the
    ! 1e-12 floor; 1: CFL and floor only; 2: dt_in(s).  Returns steps taken.
    integer(c_int) function ref_run(state, dx, gamma, nsteps, dt_mode, dt_in, dt_out, t_io, &
                                    out_count_io, t_end, divb_out) bind(C, name="ref_run")
        real(c_double), intent(inout) :: state(N, N, 6)
        real(c_double), value         :: dx, gamma, t_end
        integer(c_int), value         :: nsteps, dt_mode
        real(c_double), intent(in)    :: dt_in(*)
        real(c_double), intent(out)   :: dt_out(*)
        real(c_double), intent(inout) :: t_io
        integer(c_int), intent(inout) :: out_count_io
        real(c_double), intent(out)   :: divb_out(2, *)

        real(c_double), allocatable :: prim(:,:,:), u0(:,:,:), flux(:,:,:), faces(:,:,:)
        real(c_double), allocatable :: inv_rho(:,:), c0_sq(:,:), ca_sq(:,:), cf(:,:), divB(:,:)
        real(c_double) :: vol, dt, t, c_old, c_new
        integer :: s, k, q, outputCount

        allocate(prim(N,N,6), u0(N,N,6), flux(N,N,10), faces(N,N,24))
        allocate(inv_rho(N,N), c0_sq(N,N), ca_sq(N,N), cf(N,N), divB(N,N))
        vol = dx*dx
        t = t_io
        outputCount = out_count_io
        ref_run = 0

        do s = 1, nsteps
            if (t_end > 0.d0 .and. .not. (t < t_end)) exit

            call average_face_to_cell_B(state(:,:,5), state(:,:,6), N, N, prim(:,:,5), prim(:,:,6))
            call get_primitive(state(:,:,1), state(:,:,2), state(:,:,3), state(:,:,4), &
                               prim(:,:,5), prim(:,:,6), gamma, vol, N, N, &
                               prim(:,:,1), prim(:,:,2), prim(:,:,3), prim(:,:,4))
            if (dt_mode == 2) then
                dt = dt_in(s)
            else
                inv_rho = 1.0d0 / prim(:,:,1)
                c0_sq = gamma * (prim(:,:,4) - 0.5d0 * (prim(:,:,5)*prim(:,:,5) + prim(:,:,6)*prim(:,:,6))) &
                        * inv_rho
                ca_sq = (prim(:,:,5)*prim(:,:,5) + prim(:,:,6)*prim(:,:,6)) * inv_rho
                cf = sqrt(c0_sq + ca_sq)
                where (prim(:,:,1) < 1.0d-3) cf = min(cf, 50.0d0)
                dt = courant_fac * minval(dx / (cf + sqrt(prim(:,:,2)*prim(:,:,2) + prim(:,:,3)*prim(:,:,3))))
                if (dt_mode == 0) then
                    if (t + dt > outputCount * tOut) dt = outputCount * tOut - t
                end if
                dt = max(dt, 1e-12)
            end if
            dt_out(s) = dt
            u0 = state

            do k = 1, 3
                if (k > 1) then
                    call average_face_to_cell_B(state(:,:,5), state(:,:,6), N, N, prim(:,:,5), prim(:,:,6))
                    call get_primitive(state(:,:,1), state(:,:,2), state(:,:,3), state(:,:,4), &
                                       prim(:,:,5), prim(:,:,6), gamma, vol, N, N, &
                                       prim(:,:,1), prim(:,:,2), prim(:,:,3), prim(:,:,4))
                end if
                call ref_stage_fluxes(prim, dx, gamma, flux, faces)
                call update_conserved(state(:,:,1), flux(:,:,1), flux(:,:,2), dx, dt, N, N)
                call update_conserved(state(:,:,2), flux(:,:,3), flux(:,:,4), dx, dt, N, N)
                call update_conserved(state(:,:,3), flux(:,:,5), flux(:,:,6), dx, dt, N, N)
                call update_conserved(state(:,:,4), flux(:,:,7), flux(:,:,8), dx, dt, N, N)
                call apply_conserved_floors(state(:,:,1), state(:,:,2), state(:,:,3), state(:,:,4), &
                                            prim(:,:,5), prim(:,:,6), vol, gamma, N, N)
                call constrained_transport(state(:,:,5), state(:,:,6), flux(:,:,9), flux(:,:,10), &
                                           dx, dt, N, N)
                call outflow_fix(state, prim, gamma, dt, dx)
                if (k > 1) then
                    if (k == 2) then
                        c_old = 0.75d0;         c_new = 0.25d0
                    else
                        c_old = 1.d0/3.d0;      c_new = 2.d0/3.d0
                    end if
                    do q = 1, 6
                        state(:,:,q) = c_old*u0(:,:,q) + c_new*state(:,:,q)
                    end do
                    call apply_conserved_floors(state(:,:,1), state(:,:,2), state(:,:,3), state(:,:,4), &
                                                prim(:,:,5), prim(:,:,6), vol, gamma, N, N)
                end if
            end do

            t = t + dt
            call compute_divB(state(:,:,5), state(:,:,6), dx, N, N, divB)
            divb_out(1, s) = maxval(abs(divB))
            divb_out(2, s) = sum(abs(divB))
            if (t >= outputCount * tOut) outputCount = outputCount + 1
            ref_run = s
        end do

        t_io = t
        out_count_io = outputCount
        deallocate(prim, u0, flux, faces, inv_rho, c0_sq, ca_sq, cf, divB)
    end function ref_run

    ! Outflow correction on the first row/column (src/main.f90:116-148), with
    ! the stage primitives.
    subroutine outflow_fix(state, prim, gamma, dt, dx)
        real(c_double), intent(inout) :: state(N, N, 6)
        real(c_double), intent(in)    :: prim(N, N, 6)
        real(c_double), intent(in)    :: gamma, dt, dx
        real(c_double) :: hb(N), en(N)
        if (bc_xlo /= BC_PERIODIC) then
            hb = 0.5d0*(prim(1,:,5)**2 + prim(1,:,6)**2)
            en = (prim(1,:,4) - hb)/(gamma - 1.d0) &
               + 0.5d0*prim(1,:,1)*(prim(1,:,2)**2 + prim(1,:,3)**2) + hb
            state(1,:,1) = state(1,:,1) + dt*dx * prim(1,:,1)*prim(1,:,2)
            state(1,:,2) = state(1,:,2) + dt*dx * (prim(1,:,1)*prim(1,:,2)**2 &
                           + prim(1,:,4) - prim(1,:,5)**2)
            state(1,:,3) = state(1,:,3) + dt*dx * (prim(1,:,1)*prim(1,:,2)*prim(1,:,3) &
                           - prim(1,:,5)*prim(1,:,6))
            state(1,:,4) = state(1,:,4) + dt*dx * ((en + prim(1,:,4))*prim(1,:,2) &
                           - prim(1,:,5)*(prim(1,:,5)*prim(1,:,2) + prim(1,:,6)*prim(1,:,3)))
        end if
        if (bc_ylo /= BC_PERIODIC) then
            hb = 0.5d0*(prim(:,1,5)**2 + prim(:,1,6)**2)
            en = (prim(:,1,4) - hb)/(gamma - 1.d0) &
               + 0.5d0*prim(:,1,1)*(prim(:,1,2)**2 + prim(:,1,3)**2) + hb
            state(:,1,1) = state(:,1,1) + dt*dx * prim(:,1,1)*prim(:,1,3)
            state(:,1,3) = state(:,1,3) + dt*dx * (prim(:,1,1)*prim(:,1,3)**2 &
                           + prim(:,1,4) - prim(:,1,6)**2)
            state(:,1,2) = state(:,1,2) + dt*dx * (prim(:,1,1)*prim(:,1,3)*prim(:,1,2) &
                           - prim(:,1,6)*prim(:,1,5))
            state(:,1,4) = state(:,1,4) + dt*dx * ((en + prim(:,1,4))*prim(:,1,3) &
                           - prim(:,1,6)*(prim(:,1,6)*prim(:,1,3) + prim(:,1,5)*prim(:,1,2)))
        end if
    end subroutine outflow_fix

end module ref_harness
