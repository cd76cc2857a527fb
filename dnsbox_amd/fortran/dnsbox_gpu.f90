!-----------------------------------------------------------------------------
! dnsbox_gpu: iso_c_binding interface to libdnsbox_hip.so (include/dnsbox.h).
!
! Replaces, for a Fortran host, the module procedures of the reference's hot
! path: rhs_int_fact (rhs.f90:8-16), rhs_nonlinear (rhs.f90:18-135), the loop
! body of dnsboxheun.f90:81-157, state_project / state_dealias
! (state.f90:120-209), io_save_stats (io.f90:628-661) and rhsAdj
! (rhsAdj.f90:356-381).  Spectral arrays keep the reference layout:
! complex(8) uhat(Nz, Ny, Nx/2+1).
!-----------------------------------------------------------------------------
module dnsbox_gpu
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: DNSBOX_NDIAG = 11
    integer(c_int), parameter :: D_EKIN = 1, D_DISP = 2, D_PROD = 3, D_COURANT = 4, &
                                 D_DIVMAX = 5, D_EZERO = 6, D_EONE = 7, D_ESQRT2 = 8, &
                                 D_EBAND = 9, D_TIME = 10, D_DT = 11

    ! dnsbox_params (runtime replacement of parameters.f90)
    type, bind(C) :: dnsbox_params
        real(c_double) :: nu, Q, Pin, kcut, kolmo_amp, deltak
        real(c_double) :: cfl_min, cfl_max, dt_max, dt
        integer(c_int) :: n, precision, kolmo_k, bandlim, spherical, parity_mode
        integer(c_int) :: adaptive_dt, state_precision
        real(c_double) :: alpha_x, alpha_y, alpha_z
        integer(c_int) :: nx, ny, nz
    end type dnsbox_params

    interface
        integer(c_int) function dnsbox_params_default(p) bind(C, name="dnsbox_params_default")
            import :: c_int, dnsbox_params
            type(dnsbox_params), intent(out) :: p
        end function
        integer(c_int) function dnsbox_create(p, ctx) bind(C, name="dnsbox_create")
            import :: c_int, c_ptr, dnsbox_params
            type(dnsbox_params), intent(in) :: p
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function dnsbox_destroy(ctx) bind(C, name="dnsbox_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function dnsbox_nspec(ctx) bind(C, name="dnsbox_nspec")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function dnsbox_set_state(ctx, u, v, w) bind(C, name="dnsbox_set_state")
            import :: c_int, c_ptr, c_double_complex
            type(c_ptr), value :: ctx
            complex(c_double_complex), intent(in) :: u(*), v(*), w(*)
        end function
        integer(c_int) function dnsbox_get_state(ctx, u, v, w) bind(C, name="dnsbox_get_state")
            import :: c_int, c_ptr, c_double_complex
            type(c_ptr), value :: ctx
            complex(c_double_complex), intent(out) :: u(*), v(*), w(*)
        end function
        integer(c_int) function dnsbox_set_state_physical(ctx, u, v, w) &
                bind(C, name="dnsbox_set_state_physical")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: u(*), v(*), w(*)
        end function
        integer(c_int) function dnsbox_get_state_physical(ctx, u, v, w) &
                bind(C, name="dnsbox_get_state_physical")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: u(*), v(*), w(*)
        end function
        integer(c_int) function dnsbox_init_random(ctx, seed, box, k0, energy) &
                bind(C, name="dnsbox_init_random")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: seed
            integer(c_int), value :: box
            real(c_double), value :: k0, energy
        end function
        integer(c_int) function dnsbox_set_pruning(ctx, enable) bind(C, name="dnsbox_set_pruning")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: enable
        end function
        integer(c_int) function dnsbox_get_pruning(ctx, active) bind(C, name="dnsbox_get_pruning")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(out) :: active
        end function
        integer(c_int) function dnsbox_project(ctx, dealias, project) bind(C, name="dnsbox_project")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: dealias, project
        end function
        integer(c_int) function dnsbox_int_fact(ctx, dt) bind(C, name="dnsbox_int_fact")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        integer(c_int) function dnsbox_rhs_nonlinear(ctx) bind(C, name="dnsbox_rhs_nonlinear")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function dnsbox_get_nonlin(ctx, u, v, w) bind(C, name="dnsbox_get_nonlin")
            import :: c_int, c_ptr, c_double_complex
            type(c_ptr), value :: ctx
            complex(c_double_complex), intent(out) :: u(*), v(*), w(*)
        end function
        integer(c_int) function dnsbox_step(ctx, nsteps, stepper) bind(C, name="dnsbox_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps, stepper
        end function
        ! dnsboxAdj.f90:76-126: nsteps Heun steps in fictitious time dtau on rhsAdj, then state_project
        integer(c_int) function dnsbox_adj_descent_step(ctx, dtau, nsteps, lin_sign) &
                bind(C, name="dnsbox_adj_descent_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dtau
            integer(c_int), value :: nsteps, lin_sign
        end function
        integer(c_int) function dnsbox_diagnostics(ctx, diag, espec) bind(C, name="dnsbox_diagnostics")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: diag(*), espec(*)
        end function
        integer(c_int) function dnsbox_get_time(ctx, t, dt) bind(C, name="dnsbox_get_time")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: t, dt
        end function
        type(c_ptr) function dnsbox_last_error_c() bind(C, name="dnsbox_last_error")
            import :: c_ptr
        end function
        ! groups: many boxes advanced by one launch per pass (no reference counterpart)
        integer(c_int) function dnsbox_group_create(p, nbox, grp) bind(C, name="dnsbox_group_create")
            import :: c_int, c_ptr, dnsbox_params
            type(dnsbox_params), intent(in) :: p(*)
            integer(c_int), value :: nbox
            type(c_ptr), intent(out) :: grp
        end function
        integer(c_int) function dnsbox_group_destroy(grp) bind(C, name="dnsbox_group_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: grp
        end function
        integer(c_int) function dnsbox_group_box(grp, i, ctx) bind(C, name="dnsbox_group_box")
            import :: c_int, c_ptr
            type(c_ptr), value :: grp
            integer(c_int), value :: i
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function dnsbox_group_step(grp, nsteps) bind(C, name="dnsbox_group_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: grp
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function dnsbox_group_sync(grp) bind(C, name="dnsbox_group_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: grp
        end function
    end interface

contains

    ! message of the last failed call, as a Fortran string
    function dnsbox_last_error() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: s(:)
        integer :: n
        interface
            integer(c_size_t) function c_strlen(p) bind(C, name="strlen")
                import :: c_size_t, c_ptr
                type(c_ptr), value :: p
            end function
        end interface
        type(c_ptr) :: p
        p = dnsbox_last_error_c()
        n = int(c_strlen(p))
        call c_f_pointer(p, s, [n])
        allocate(character(len=n) :: msg)
        msg = transfer(s(1:n), msg)
    end function

    subroutine dnsbox_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc /= 0) then
            print *, 'dnsbox error in ', what, ': ', dnsbox_last_error()
            error stop 1
        end if
    end subroutine

end module dnsbox_gpu
