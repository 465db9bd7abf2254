!------------------------------------------------------------------------------
! isph_capi -- iso_c_binding interface to libisph.so (include/isph_capi.h).
!
! The reference driver (main.f90:41-66) calls zero-argument module
! subroutines that act on module globals.  With this module the same driver
! keeps its call order and replaces each call by the entry point of the same
! name, e.g. `call neighbour` -> `ierr = isph_neighbour(ctx)`; the particle
! state lives on the GPU inside `ctx`.
!------------------------------------------------------------------------------
module isph_capi
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int), parameter :: ISPH_OK = 0
    integer(c_int), parameter :: ISPH_KERNEL_HYBRID = 0, ISPH_KERNEL_WENDLAND_C2 = 1
    integer(c_int), parameter :: ISPH_DRAG_REFERENCE = 0, ISPH_DRAG_ERGUN = 1
    integer(c_int), parameter :: ISPH_PREC_FP64 = 0, ISPH_PREC_MIXED = 1, ISPH_PREC_MIXED_V64 = 2
    integer(c_int), parameter :: ISPH_PREC_MIXED_MG = 3, ISPH_PREC_FGMRES = 4, ISPH_PREC_IR = 5
    integer(c_int), parameter :: ISPH_PREC_FP64_MG = 6

    ! same member order as the C struct: doubles first, then int32s
    type, bind(C) :: isph_params
        real(c_double) :: h, cell_size, xmin, ymin, xmax, ymax, dl, lam, dt, rho0, rel_den, g, nu
        real(c_double) :: csh, maxshift, lamfs, maxdivr, por, alpha, beta, d50, tol, inner_tol
        real(c_double) :: xlcutoff, ytcutoff, sig1, mu, rhomin, tschmidt, salt_xbuf
        real(c_double) :: xrelease_lo, xrelease_hi
        integer(c_int32_t) :: kernel, drag_law, precision, maxit, inner_maxit, max_outer, max_nbr
        integer(c_int32_t) :: shifting, adaptive_dt, sink, visc_law, salinity, cell_order, reserved
    end type isph_params

    ! a buffer zone (include/isph_capi.h, isph_buffer_zone)
    type, bind(C) :: isph_buffer_zone
        integer(c_int32_t) :: ntemplate, ncells
        type(c_ptr) :: x, y, vx, vy, pressure   ! real(c_double) arrays of ntemplate
        type(c_ptr) :: flags                    ! integer(c_int32_t) array of ntemplate
        type(c_ptr) :: cells                    ! integer(c_int32_t) (2, ncells)
        real(c_double) :: mass, density, ovol, con, period, xcut
        integer(c_int32_t) :: remove_right, reserved
    end type isph_buffer_zone

    interface
        integer(c_int) function isph_create(prm, cap, cap_soil, ctx) bind(C, name="isph_create")
            import :: c_int, c_int32_t, c_ptr, isph_params
            type(isph_params), intent(in) :: prm
            integer(c_int32_t), value :: cap, cap_soil
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function isph_destroy(ctx) bind(C, name="isph_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function isph_last_error(ctx) bind(C, name="isph_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_set_particles(ctx, n, x, y, vx, vy, mass, density, oden, &
                                                   pressure, tid, flags, pid) &
                                                   bind(C, name="isph_set_particles")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: x(*), y(*), vx(*), vy(*), mass(*), density(*), oden(*)
            real(c_double), intent(in) :: pressure(*)
            integer(c_int32_t), intent(in) :: tid(*), flags(*), pid(*)
        end function
        integer(c_int) function isph_get_particles(ctx, n, x, y, vx, vy, mass, density, oden, &
                                                   pressure, tid, flags, pid) &
                                                   bind(C, name="isph_get_particles")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: n
            real(c_double), intent(out) :: x(*), y(*), vx(*), vy(*), mass(*), density(*), oden(*)
            real(c_double), intent(out) :: pressure(*)
            integer(c_int32_t), intent(out) :: tid(*), flags(*), pid(*)
        end function
        integer(c_int) function isph_set_soil(ctx, ns, x, y, mass, density) &
                                              bind(C, name="isph_set_soil")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: ns
            real(c_double), intent(in) :: x(*), y(*), mass(*), density(*)
        end function
        integer(c_int) function isph_get_scalars(ctx, umax, dt, t, iters, outer, relres) &
                                                 bind(C, name="isph_get_scalars")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: umax, dt, t, relres
            integer(c_int32_t), intent(out) :: iters, outer
        end function
        integer(c_int) function isph_set_scalars(ctx, umax, dt, t) bind(C, name="isph_set_scalars")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: umax, dt, t
        end function
        ! one entry per reference subroutine (main.f90:41-66)
        integer(c_int) function isph_projection(ctx) bind(C, name="isph_projection")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_cellshift(ctx) bind(C, name="isph_cellshift")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_resetid(ctx) bind(C, name="isph_resetid")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_neighbour(ctx) bind(C, name="isph_neighbour")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_effpor(ctx) bind(C, name="isph_effpor")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_eddyvis(ctx) bind(C, name="isph_eddyvis")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_int_vel(ctx) bind(C, name="isph_int_vel")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_ppesolve(ctx, iters, relres) bind(C, name="isph_ppesolve")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: iters
            real(c_double), intent(out) :: relres
        end function
        integer(c_int) function isph_comp_vel(ctx) bind(C, name="isph_comp_vel")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_comp_pos(ctx) bind(C, name="isph_comp_pos")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_cellshiftalt(ctx) bind(C, name="isph_cellshiftalt")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_freesurf(ctx) bind(C, name="isph_freesurf")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_opt2_shift(ctx) bind(C, name="isph_opt2_shift")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_timestep(ctx) bind(C, name="isph_timestep")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_step(ctx, nsteps) bind(C, name="isph_step")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nsteps
        end function
        ! salinity transport and variable density (scalar.f90, main.f90:160-263)
        integer(c_int) function isph_scalart(ctx) bind(C, name="isph_scalart")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_scalarupdate(ctx, stime) bind(C, name="isph_scalarupdate")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: stime
        end function
        integer(c_int) function isph_densityupdate(ctx) bind(C, name="isph_densityupdate")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_eddyvis_lagged(ctx) bind(C, name="isph_eddyvis_lagged")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function isph_salt_init(ctx, xl, yl, line_grad) &
                                               bind(C, name="isph_salt_init")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: xl, yl, line_grad
        end function
        ! buffer zones (internal.f90:475-614, main.f90:103-121,218-236)
        integer(c_int) function isph_set_buffer_zones(ctx, nzones, zones, time_shift) &
                                                      bind(C, name="isph_set_buffer_zones")
            import :: c_int, c_int32_t, c_ptr, c_double, isph_buffer_zone
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nzones
            type(isph_buffer_zone), intent(in) :: zones(*)
            real(c_double), value :: time_shift
        end function
        integer(c_int) function isph_remove_buffer(ctx, zone) bind(C, name="isph_remove_buffer")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: zone
        end function
        integer(c_int) function isph_insert_buffer(ctx, zone) bind(C, name="isph_insert_buffer")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: zone
        end function
        integer(c_int) function isph_get_substeps(ctx, solsteps, dtsol) &
                                                  bind(C, name="isph_get_substeps")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: solsteps
            real(c_double), intent(out) :: dtsol
        end function
    end interface

contains

    ! message of the last failed call as a Fortran string
    function isph_message(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=256) :: msg
        character(kind=c_char), pointer :: s(:)
        integer :: k
        msg = ' '
        call c_f_pointer(isph_last_error(ctx), s, [256])
        do k = 1, 256
            if (s(k) == c_null_char) exit
            msg(k:k) = s(k)
        end do
    end function

    subroutine isph_check(ctx, ierr, where)
        type(c_ptr), intent(in) :: ctx
        integer(c_int), intent(in) :: ierr
        character(len=*), intent(in) :: where
        if (ierr /= ISPH_OK) then
            write(*, '(a,a,a,i0,a,a)') 'isph: ', where, ' failed with ', ierr, ': ', &
                trim(isph_message(ctx))
            stop 1
        end if
    end subroutine

end module isph_capi
