! harness.f90 -- test driver around the reference's own Fortran hot path.
!
! TEST INFRASTRUCTURE ONLY. build_ref.sh compiles this program together with the
! reference's comum.f90, boundary.f90, equations.f90 and nonsymetric_mesh.f90 (taken from
! OpenMp_Multicore/ at build time, never stored in this repository) and writes the
! executable to oracle/_ref/. It is run only in the development container, by
! tests/golden/make_golden.py, to produce the golden vectors under tests/golden/.
!
! Input  (stream, native endian): harness_in.bin
!   int32  mode            1 = operators, 2 = pseudo-steps, 3 = physical steps, 4 = mesh,
!                          5 = mixture properties (properties_CH4.f90) at nsteps temperatures
!   int32  nsteps, nphys, itc_max, steady
!   real64 dtau, dt, beta, b_art, tinf, eps_conv
!   geometry (modes 1-3): x(imax), y(jmax), xm(imax+1), ym(jmax+1), areau_n(2:imax),
!          areau_e(jmax), areav_n(imax), areav_e(2:jmax), epsilon1(imax,jmax), liga_poros(imax,jmax)
!   state  (modes 1-3): um(imax+1,jmax), vm(imax,jmax+1), p(imax,jmax), T(imax,jmax), C(imax,jmax)
! Output (stream): harness_out.bin, layout per mode (see make_golden.py).
!
! The driver loops restate OpenMp_Multicore/main.f90:116-157 with the unset levels pinned:
! pn = p, T_tau = T, C_tau = C before the first pseudo-step.
program ref_harness
    use comum
    implicit none
    integer(4) :: mode, nsteps, nphys, itc_max4, steady
    real(8) :: eps_conv
    real(8), dimension(1:imax+1,1:jmax) :: um, um_n, um_tau, um_n_tau, ru_a, ub
    real(8), dimension(1:imax,1:jmax+1) :: vm, vm_n, vm_tau, vm_n_tau, rv_a, vb
    real(8), dimension(1:imax,1:jmax) :: p, pn, T, T_tau, T_n_tau, C, C_tau, C_n_tau, rz_a, rc_a, pb
    real(8) :: r_u, r_v, r_p, error
    real(8), allocatable :: hist(:,:)
    integer(4), allocatable :: taken(:)
    integer :: k, j, i, itc, tr

    open(unit=21, file='harness_in.bin', access='stream', form='unformatted', status='old')
    read(21) mode, nsteps, nphys, itc_max4, steady
    read(21) dtau, dt, beta, b_art, tinf, eps_conv

    open(unit=22, file='harness_out.bin', access='stream', form='unformatted', status='replace')

    if (mode == 4) then
        call mesh
        write(22) x, y, xm, ym, areau_n, areau_e, areav_n, areav_e, epsilon1, liga_poros
        close(22)
        stop
    end if

    if (mode == 5) then
        ! Too(k), k = 1..nsteps, follow the header; per temperature the reference's mixture
        ! properties: k_tot, rho_tot, cp_tot, nu_tot, alpha_tot, Pr
        YF_b = 1.d0
        do k = 1, nsteps
            read(21) Too
            call properties
            write(22) k_tot, rho_tot, cp_tot, nu_tot, alpha_tot, Pr
        end do
        close(21)
        close(22)
        stop
    end if

    read(21) x, y, xm, ym, areau_n, areau_e, areav_n, areav_e, epsilon1, liga_poros
    areau_s = areau_n
    areau_w = areau_e
    areav_s = areav_n
    areav_w = areav_e
    read(21) um, vm, p, T, C
    close(21)

    if (mode == 1) then
        ru_a = 0.d0; rv_a = 0.d0; rz_a = 0.d0; rc_a = 0.d0
        call RESU(um, vm, p, ru_a)
        call RESV(um, vm, p, T, rv_a)
        call RESZ(um, vm, T, rz_a)
        call RESC(um, vm, C, rc_a)
        write(22) ru_a, rv_a, rz_a, rc_a
        ! first-order upwind rows on every row (the global-upwind operator)
        ru_a = 0.d0; rv_a = 0.d0
        do j = 2, jmax-1
            call upwind_Ui(um, vm, p, ru_a, j)
        end do
        do j = 2, jmax
            call upwind_Vi(um, vm, p, rv_a, T, j)
        end do
        write(22) ru_a, rv_a
        ub = um; vb = vm
        call bcUV(ub, vb)
        write(22) ub, vb
        pb = p
        call bcP(pb)
        write(22) pb
        pb = T
        call bcZ(pb)
        write(22) pb
        pb = C
        call bcC(pb)
        write(22) pb
        ! one solve of each kind, in the order of the driver
        um_tau = um; vm_tau = vm; pn = p; T_tau = T; C_tau = C
        um_n_tau = 0.d0; vm_n_tau = 0.d0
        call solve_U(um, vm, um_n, um_tau, vm_tau, um_n_tau, pn, r_u)
        write(22) um_n_tau, vm_tau, r_u
        call solve_V(um, vm, vm_n, um_tau, vm_tau, vm_n_tau, pn, T, r_v)
        write(22) vm_n_tau, um_tau, r_v
        call solve_P(p, um_n_tau, vm_n_tau, pn, r_p)
        write(22) pn, r_p
        call solve_Z(um_n_tau, vm_n_tau, T, T_n_tau, T_tau)
        write(22) T_n_tau
        call solve_C(um_n_tau, vm_n_tau, C, C_n_tau, C_tau)
        write(22) C_n_tau
        close(22)
        stop
    end if

    um_tau = um; vm_tau = vm; pn = p; T_tau = T; C_tau = C
    um_n_tau = 0.d0; vm_n_tau = 0.d0; T_n_tau = 0.d0; C_n_tau = 0.d0

    if (mode == 2) then
        allocate(hist(3, nsteps))
        hist = 0.d0
        do k = 1, nsteps
            if (steady /= 0) then
                um = um_tau; vm = vm_tau; T = T_tau; C = C_tau
            end if
            call solve_U(um, vm, um_n, um_tau, vm_tau, um_n_tau, pn, r_u)
            call solve_V(um, vm, vm_n, um_tau, vm_tau, vm_n_tau, pn, T, r_v)
            call solve_P(p, um_n_tau, vm_n_tau, pn, r_p)
            call solve_Z(um_n_tau, vm_n_tau, T, T_n_tau, T_tau)
            call solve_C(um_n_tau, vm_n_tau, C, C_n_tau, C_tau)
            hist(1, k) = r_u; hist(2, k) = r_v; hist(3, k) = r_p
            um_tau = um_n_tau
            vm_tau = vm_n_tau
            p = pn
            T_tau = T_n_tau
            C_tau = C_n_tau
            ! with eps_conv > 0, stop after the first step past the first whose residual
            ! max(r_u, r_v, r_p) is below it (main.f90:139-150's exit test); the state
            ! written is the one after that step, the later history rows stay zero
            if (eps_conv > 0.d0 .and. k > 1 .and. max(r_u, r_v, r_p) < eps_conv) exit
        end do
        write(22) hist
        write(22) um_tau, vm_tau, p, T_tau, C_tau
        close(22)
        stop
    end if

    if (mode == 3) then
        allocate(taken(nphys))
        itc = 1
        do tr = 1, nphys
            taken(tr) = 0
            do while (itc .lt. itc_max4)
                call solve_U(um, vm, um_n, um_tau, vm_tau, um_n_tau, pn, r_u)
                call solve_V(um, vm, vm_n, um_tau, vm_tau, vm_n_tau, pn, T, r_v)
                call solve_P(p, um_n_tau, vm_n_tau, pn, r_p)
                call solve_Z(um_n_tau, vm_n_tau, T, T_n_tau, T_tau)
                call solve_C(um_n_tau, vm_n_tau, C, C_n_tau, C_tau)
                taken(tr) = taken(tr) + 1
                error = max(r_u, r_v, r_p)
                if (itc .ne. 1 .and. error .lt. eps_conv) exit
                um_tau = um_n_tau
                vm_tau = vm_n_tau
                p = pn
                T_tau = T_n_tau
                C_tau = C_n_tau
                itc = itc + 1
            end do
            um = um_n_tau
            vm = vm_n_tau
            T = T_n_tau
            C = C_n_tau
            itc = 0
        end do
        write(22) taken
        write(22) um, vm, um_tau, vm_tau, p, pn, T, T_tau, C, C_tau
        write(22) r_u, r_v, r_p
        close(22)
        stop
    end if
    close(22)
end program ref_harness
