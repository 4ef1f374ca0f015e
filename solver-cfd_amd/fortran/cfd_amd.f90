! cfd_amd.f90 -- iso_c_binding interface to libcfd_amd.so for a Fortran driver.
!
! A Fortran main program like OpenMp_Multicore/main.f90 replaces its calls to
! solve_U/solve_V/solve_P/solve_Z/solve_C (main.f90:122-132) by one context and
! cfd_pseudo_steps / cfd_physical_step, passing its own column-major arrays with
! CFD_LAYOUT_FORTRAN: um(1:imax+1,1:jmax), vm(1:imax,1:jmax+1), p/T/C(1:imax,1:jmax).
! The derived type mirrors struct cfd_params of include/cfd.h field by field.
module cfd_amd
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int32_t), parameter :: CFD_SCHEME_QUICK = 0, CFD_SCHEME_UPWIND = 1
    integer(c_int32_t), parameter :: CFD_BC_REFERENCE = 0, CFD_BC_CHANNEL = 1
    integer(c_int32_t), parameter :: CFD_BUOY_REFERENCE = 0, CFD_BUOY_BOUSSINESQ = 1
    integer(c_int32_t), parameter :: CFD_COMPAT_FORTRAN = 0, CFD_COMPAT_CUDA = 1
    integer(c_int32_t), parameter :: CFD_PREC_FP64 = 0, CFD_PREC_FP32 = 1
    integer(c_int32_t), parameter :: CFD_MU_CONST = 0, CFD_MU_POLY4 = 1, CFD_K_CONST = 0, CFD_K_POLY4 = 1
    integer(c_int32_t), parameter :: CFD_LAYOUT_C = 0, CFD_LAYOUT_FORTRAN = 1
    integer(c_int32_t), parameter :: CFD_F_UM = 0, CFD_F_VM = 1, CFD_F_UM_TAU = 2, CFD_F_VM_TAU = 3, &
                                     CFD_F_P = 4, CFD_F_T = 5, CFD_F_T_TAU = 6, CFD_F_C = 7, &
                                     CFD_F_C_TAU = 8, CFD_F_EPS = 9, CFD_F_LP = 10, CFD_F_PN = 11

    type, bind(C) :: cfd_params
        integer(c_int32_t) :: imax, jmax, scheme, bc, buoyancy, compat, steady, mu_model
        real(c_double) :: re, pe, sc, darcy, cf, g, invfr2, ri, t_ref, v_i, tinf
        real(c_double) :: dtau, dt, beta, b_art, u_in, t_bottom, t_top
        real(c_double) :: slot_lo, slot_hi, slot_value
        real(c_double) :: mu_coef(5)
        real(c_double) :: k_coef(5)
        integer(c_int32_t) :: k_model, reserved
    end type cfd_params

    interface
        subroutine cfd_params_default(p, imax, jmax) bind(C, name="cfd_params_default")
            import :: cfd_params, c_int32_t
            type(cfd_params), intent(out) :: p
            integer(c_int32_t), value :: imax, jmax
        end subroutine

        integer(c_int) function cfd_ctx_create(p, precision, ctx) bind(C, name="cfd_ctx_create")
            import :: cfd_params, c_int, c_int32_t, c_ptr
            type(cfd_params), intent(in) :: p
            integer(c_int32_t), value :: precision
            type(c_ptr), intent(out) :: ctx
        end function

        integer(c_int) function cfd_ctx_destroy(ctx) bind(C, name="cfd_ctx_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        ! x(0:imax), y(0:jmax): element 0 unused, as the C layout
        integer(c_int) function cfd_ctx_set_mesh(ctx, x, y) bind(C, name="cfd_ctx_set_mesh")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: x(*), y(*)
        end function

        integer(c_int) function cfd_ctx_upload(ctx, field, host, layout) bind(C, name="cfd_ctx_upload")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: field, layout
            real(c_double), intent(in) :: host(*)
        end function

        integer(c_int) function cfd_ctx_download(ctx, field, host, layout) bind(C, name="cfd_ctx_download")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: field, layout
            real(c_double), intent(out) :: host(*)
        end function

        integer(c_int) function cfd_ctx_sync_levels(ctx) bind(C, name="cfd_ctx_sync_levels")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function cfd_pseudo_steps(ctx, nsteps, res, history) bind(C, name="cfd_pseudo_steps")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nsteps
            real(c_double), intent(out) :: res(3)
            type(c_ptr), value :: history
        end function

        integer(c_int) function cfd_run_to_tol(ctx, tol, itc_max, check_every, diverge_factor, iters, res) &
                bind(C, name="cfd_run_to_tol")
            import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: tol, diverge_factor
            integer(c_int64_t), value :: itc_max
            integer(c_int32_t), value :: check_every
            integer(c_int64_t), intent(out) :: iters
            real(c_double), intent(out) :: res(3)
        end function

        integer(c_int) function cfd_physical_step(ctx, eps, itc_max, first, iters, res) &
                bind(C, name="cfd_physical_step")
            import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: eps
            integer(c_int64_t), value :: itc_max
            integer(c_int32_t), value :: first
            integer(c_int64_t), intent(out) :: iters
            real(c_double), intent(out) :: res(3)
        end function

        type(c_ptr) function cfd_last_error() bind(C, name="cfd_last_error")
            import :: c_ptr
        end function

        ! include/cfd_fortran.h: hands the module state of comum.f90 (parameters, nodes
        ! x(1:imax), y(1:jmax), epsilon1 and liga_poros) to the library once; afterwards a
        ! driver CALLs solve_U, solve_V, solve_P, solve_Z, solve_C, RESU, ..., bcUV, ... as
        ! the reference's main.f90 does (external procedures, implicit interface, host arrays).
        integer(c_int) function cfd_fortran_setup(p, x, y, epsilon1, liga_poros) bind(C, name="cfd_fortran_setup")
            import :: cfd_params, c_int, c_double
            type(cfd_params), intent(in) :: p
            real(c_double), intent(in) :: x(*), y(*), epsilon1(*), liga_poros(*)
        end function

        subroutine cfd_fortran_finalize() bind(C, name="cfd_fortran_finalize")
        end subroutine
    end interface
end module cfd_amd
