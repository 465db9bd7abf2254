!------------------------------------------------------------------------------
! isph_driver -- Fortran driver of the gfx950 ISPH step.
!
! Keeps the reference's time loop (main.f90:41-66) and its subroutine
! order; each call goes through the iso_c_binding interface of libisph.so.
!
!   isph_driver <case.bin> <result.bin> <nsteps>
!
! case.bin (stream, little-endian): isph_params, int32 n, int32 ns, then the
! n-particle arrays x, y, vx, vy, mass, density, oden, pressure (float64) and
! tid, flags, pid (int32), then the ns soil arrays x, y, mass, density.
! result.bin: int32 n, float64 t, then x, y, vx, vy, pressure (float64) and
! pid (int32) in the context's final order.
!------------------------------------------------------------------------------
program isph_driver
    use, intrinsic :: iso_c_binding
    use isph_capi
    implicit none

    type(isph_params) :: prm
    type(c_ptr) :: ctx
    integer(c_int32_t) :: n, ns, nsteps, iters, outer, step, nout
    real(c_double), allocatable :: x(:), y(:), vx(:), vy(:), mass(:), density(:), oden(:), p(:)
    real(c_double), allocatable :: sx(:), sy(:), sm(:), sd(:)
    integer(c_int32_t), allocatable :: tid(:), flags(:), pid(:)
    real(c_double) :: umax, dt, t, relres
    integer(c_int) :: ierr
    character(len=512) :: fin, fout, arg
    integer :: u

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    call get_command_argument(3, arg)
    read(arg, *) nsteps

    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) prm
    read(u) n, ns
    allocate(x(n), y(n), vx(n), vy(n), mass(n), density(n), oden(n), p(n), tid(n), flags(n), &
             pid(n), sx(max(ns, 1)), sy(max(ns, 1)), sm(max(ns, 1)), sd(max(ns, 1)))
    read(u) x, y, vx, vy, mass, density, oden, p, tid, flags, pid
    if (ns > 0) read(u) sx(1:ns), sy(1:ns), sm(1:ns), sd(1:ns)
    close(u)

    ! init, gen_boundary, gen_fluid, memalloc (main.f90:23-33)
    ierr = isph_create(prm, n, max(ns, 1), ctx)
    if (ierr /= ISPH_OK) then
        write(*, *) 'isph_create failed', ierr
        stop 1
    end if
    call isph_check(ctx, isph_set_particles(ctx, n, x, y, vx, vy, mass, density, oden, p, tid, &
                                            flags, pid), 'set_particles')
    call isph_check(ctx, isph_set_soil(ctx, ns, sx, sy, sm, sd), 'set_soil')
    ! effpor before the loop (main.f90:35-37); umax from the initial field
    call isph_check(ctx, isph_effpor(ctx), 'effpor')
    call isph_check(ctx, isph_timestep(ctx), 'timestep')

    do step = 1, nsteps
        ! told = t; t = t + dt (main.f90:42-43)
        call isph_check(ctx, isph_get_scalars(ctx, umax, dt, t, iters, outer, relres), 'scalars')
        call isph_check(ctx, isph_set_scalars(ctx, umax, dt, t + dt), 'scalars')
        call isph_check(ctx, isph_projection(ctx), 'projection')
        call isph_check(ctx, isph_cellshift(ctx), 'cellshift')
        call isph_check(ctx, isph_resetid(ctx), 'resetid')
        call isph_check(ctx, isph_neighbour(ctx), 'neighbour')
        call isph_check(ctx, isph_effpor(ctx), 'effpor')
        call isph_check(ctx, isph_eddyvis(ctx), 'eddyvis')
        call isph_check(ctx, isph_int_vel(ctx), 'int_vel')
        call isph_check(ctx, isph_ppesolve(ctx, iters, relres), 'ppesolve')
        call isph_check(ctx, isph_comp_vel(ctx), 'comp_vel')
        call isph_check(ctx, isph_comp_pos(ctx), 'comp_pos')
        call isph_check(ctx, isph_cellshiftalt(ctx), 'cellshiftalt')
        call isph_check(ctx, isph_neighbour(ctx), 'neighbour')
        call isph_check(ctx, isph_freesurf(ctx), 'freesurf')
        call isph_check(ctx, isph_effpor(ctx), 'effpor')
        call isph_check(ctx, isph_opt2_shift(ctx), 'opt2_shift')
        call isph_check(ctx, isph_timestep(ctx), 'timestep')
        write(*, '(a,i6,a,i6,a,es12.4)') 'step ', step, '  PPE iterations ', iters, &
            '  relres ', relres
    end do

    call isph_check(ctx, isph_get_particles(ctx, nout, x, y, vx, vy, mass, density, oden, p, tid, &
                                            flags, pid), 'get_particles')
    call isph_check(ctx, isph_get_scalars(ctx, umax, dt, t, iters, outer, relres), 'scalars')
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) nout, t
    write(u) x(1:nout), y(1:nout), vx(1:nout), vy(1:nout), p(1:nout), pid(1:nout)
    close(u)
    ierr = isph_destroy(ctx)
end program isph_driver
