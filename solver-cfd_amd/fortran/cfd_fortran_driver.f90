! cfd_fortran_driver.f90 -- a Fortran main program that runs the physical-time loop of
! OpenMp_Multicore/main.f90:116-157 through libcfd_amd.so's Fortran entry points
! (include/cfd_fortran.h). The CALLs below are the reference's, unchanged: solve_U,
! solve_V, solve_P, solve_Z and solve_C are external procedures with implicit interfaces,
! resolved at link time to the library's solve_u_, ..., solve_c_. The module state that
! comum.f90 holds in the reference is handed over once with cfd_fortran_setup.
!
! Usage: cfd_fortran_driver <input> <output>
! Input (stream, native endian):
!   cfd_params                        (the bind(C) type of module cfd_amd, component order)
!   int32  nphys, itc_max             physical steps; pseudo-step cap per physical step
!   real64 eps_conv                   exit tolerance of main.f90:139-142
!   real64 x(imax), y(jmax), epsilon1(imax,jmax), liga_poros(imax,jmax)
!   real64 um(imax+1,jmax), vm(imax,jmax+1), p(imax,jmax), T(imax,jmax), C(imax,jmax)
! Output (stream): int32 taken(nphys) (pseudo-steps run per physical step), then
!   um, vm, um_tau, vm_tau, p, pn, T, T_tau, C, C_tau, and residual_u, residual_v, residual_p.
! The levels main.f90 reads before writing on the first pseudo-step start from the physical
! level: um_tau = um, vm_tau = vm, pn = p, T_tau = T, C_tau = C (the n_tau outputs from 0).
program cfd_fortran_driver
    use, intrinsic :: iso_c_binding
    use, intrinsic :: ieee_arithmetic, only: ieee_is_nan
    use cfd_amd
    implicit none
    type(cfd_params) :: prm
    integer(c_int32_t) :: nphys, itc_max
    real(c_double) :: eps_conv, residual_u, residual_v, residual_p, error
    real(c_double), allocatable :: x(:), y(:), epsilon1(:,:), liga_poros(:,:)
    real(c_double), allocatable, dimension(:,:) :: um, um_n, um_tau, um_n_tau
    real(c_double), allocatable, dimension(:,:) :: vm, vm_n, vm_tau, vm_n_tau
    real(c_double), allocatable, dimension(:,:) :: p, pn, T, T_tau, T_n_tau, C, C_tau, C_n_tau
    integer(c_int32_t), allocatable :: taken(:)
    integer :: imax, jmax, itc, tr, rc
    character(len=4096) :: fin, fout
    external :: solve_U, solve_V, solve_P, solve_Z, solve_C

    if (command_argument_count() /= 2) then
        write(*, '(a)') 'usage: cfd_fortran_driver <input> <output>'
        stop 2
    end if
    call get_command_argument(1, fin)
    call get_command_argument(2, fout)

    open(unit=21, file=trim(fin), access='stream', form='unformatted', status='old')
    read(21) prm
    read(21) nphys, itc_max
    read(21) eps_conv
    imax = prm%imax
    jmax = prm%jmax
    allocate(x(imax), y(jmax), epsilon1(imax, jmax), liga_poros(imax, jmax))
    allocate(um(imax+1, jmax), um_n(imax+1, jmax), um_tau(imax+1, jmax), um_n_tau(imax+1, jmax))
    allocate(vm(imax, jmax+1), vm_n(imax, jmax+1), vm_tau(imax, jmax+1), vm_n_tau(imax, jmax+1))
    allocate(p(imax, jmax), pn(imax, jmax), T(imax, jmax), T_tau(imax, jmax), T_n_tau(imax, jmax))
    allocate(C(imax, jmax), C_tau(imax, jmax), C_n_tau(imax, jmax), taken(nphys))
    read(21) x, y, epsilon1, liga_poros
    read(21) um, vm, p, T, C
    close(21)

    rc = cfd_fortran_setup(prm, x, y, epsilon1, liga_poros)
    if (rc /= 0) then
        write(*, '(a,i0)') 'cfd_fortran_setup failed: ', rc
        stop 1
    end if

    um_tau = um; vm_tau = vm; pn = p; T_tau = T; C_tau = C
    um_n = 0.d0; vm_n = 0.d0
    um_n_tau = 0.d0; vm_n_tau = 0.d0; T_n_tau = 0.d0; C_n_tau = 0.d0
    residual_u = 0.d0; residual_v = 0.d0; residual_p = 0.d0

    itc = 1
    do tr = 1, nphys
        taken(tr) = 0
        do while (itc .lt. itc_max)
            call solve_U(um, vm, um_n, um_tau, vm_tau, um_n_tau, pn, residual_u)
            call solve_V(um, vm, vm_n, um_tau, vm_tau, vm_n_tau, pn, T, residual_v)
            call solve_P(p, um_n_tau, vm_n_tau, pn, residual_p)
            call solve_Z(um_n_tau, vm_n_tau, T, T_n_tau, T_tau)
            call solve_C(um_n_tau, vm_n_tau, C, C_n_tau, C_tau)
            taken(tr) = taken(tr) + 1
            error = max(residual_u, residual_v, residual_p)
            if (ieee_is_nan(error)) then
                write(*, '(a,i0,a,i0)') 'non-finite residual in physical step ', tr, ', pseudo-step ', taken(tr)
                stop 1
            end if
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

    open(unit=22, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(22) taken
    write(22) um, vm, um_tau, vm_tau, p, pn, T, T_tau, C, C_tau
    write(22) residual_u, residual_v, residual_p
    close(22)
    call cfd_fortran_finalize()
    write(*, '(a,i0,a)') 'cfd_fortran_driver: ', nphys, ' physical steps done'
end program cfd_fortran_driver
