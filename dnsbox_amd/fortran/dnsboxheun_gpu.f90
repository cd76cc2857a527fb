!-----------------------------------------------------------------------------
! The time loop of dnsboxheun.f90:81-191 with the hot path on the GPU: IF-Heun
! steps, CFL control, statistics every iSaveRate2 steps (io_save_stats,
! io.f90:628-661, written like Ekin.dat), spectrum at the end (io_save_spectrum).
!
! usage: dnsboxheun_gpu [n] [nsteps] [precision] [nu]
!-----------------------------------------------------------------------------
program dnsboxheun_gpu
    use iso_c_binding
    use dnsbox_gpu
    implicit none
    type(dnsbox_params) :: p
    type(c_ptr) :: ctx
    integer :: n, nsteps, prec, it, nspec, iSaveRate2, k
    real(c_double) :: nu, diag(DNSBOX_NDIAG)
    real(c_double), allocatable :: espec(:)
    character(len=32) :: arg

    n = 32; nsteps = 100; prec = 64; nu = 0.05d0; iSaveRate2 = 10
    if (command_argument_count() >= 1) then; call get_command_argument(1, arg); read(arg, *) n; end if
    if (command_argument_count() >= 2) then; call get_command_argument(2, arg); read(arg, *) nsteps; end if
    if (command_argument_count() >= 3) then; call get_command_argument(3, arg); read(arg, *) prec; end if
    if (command_argument_count() >= 4) then; call get_command_argument(4, arg); read(arg, *) nu; end if

    call dnsbox_check(dnsbox_params_default(p), 'params_default')
    p%n = n; p%precision = prec; p%nu = nu
    p%kolmo_amp = 1.0d0; p%kolmo_k = 4; p%bandlim = 0
    p%adaptive_dt = 0; p%dt = 0.005d0
    call dnsbox_check(dnsbox_create(p, ctx), 'create')
    nspec = dnsbox_nspec(ctx)
    allocate(espec(nspec))

    ! initial field (state_init_rand replaced by the seeded generator), dealias, project
    call dnsbox_check(dnsbox_init_random(ctx, 0_c_int64_t, 0_c_int, 4.0d0, 0.5d0), 'init_random')
    call dnsbox_check(dnsbox_project(ctx, 1_c_int, 1_c_int), 'project')
    call dnsbox_check(dnsbox_int_fact(ctx, p%dt), 'int_fact')

    do it = 1, nsteps / iSaveRate2
        call dnsbox_check(dnsbox_step(ctx, iSaveRate2, 1_c_int), 'step')
        call dnsbox_check(dnsbox_diagnostics(ctx, diag, espec), 'diagnostics')
        ! columns of Ekin.dat (io.f90:651-658)
        write(*, '(8e20.12)') diag(D_TIME), diag(D_EKIN), diag(D_DISP), diag(D_PROD), &
                              diag(D_COURANT), diag(D_DIVMAX), diag(D_EZERO), diag(D_EBAND)
    end do
    do k = 1, nspec
        write(*, '(a, 2e20.12)') 'spec', real(k, c_double), espec(k)
    end do
    call dnsbox_check(dnsbox_destroy(ctx), 'destroy')
end program dnsboxheun_gpu
