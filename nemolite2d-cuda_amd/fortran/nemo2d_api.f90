!> ISO_C_BINDING interfaces to the nemo2d C-ABI (include/nemo2d.h).
!!
!! The reference binds its CUDA port the same way for the parameter set-up
!! and boundary call (cuda/nemolite2d.f90:62-127) and through implicit
!! interfaces for the rest; here every entry point has an explicit bind(C)
!! interface, an opaque handle and an integer status result.
module nemo2d_api
    use, intrinsic :: iso_c_binding, only: c_int, c_double, c_ptr, c_char
    implicit none
    private
    public :: nemo2d_create, nemo2d_set_grid, nemo2d_set_state, &
              nemo2d_continuity, nemo2d_momentum_u, nemo2d_momentum_v, &
              nemo2d_bc_ssh, nemo2d_bc_solid_u, nemo2d_bc_solid_v, &
              nemo2d_bc_flather_u, nemo2d_bc_flather_v, nemo2d_next, &
              nemo2d_step, nemo2d_run, nemo2d_record_checksums, nemo2d_checksum_log, &
              nemo2d_checksums, nemo2d_retrieve_results, nemo2d_retrieve_aux, &
              nemo2d_timer_start, nemo2d_timer_stop, nemo2d_destroy, &
              nemo2d_set_output_grid, nemo2d_snapshot_begin, nemo2d_snapshot_wait, &
              nemo2d_stage_timers, nemo2d_stage_times, nemo2d_set_stage_fusion, nemo2d_check

    interface
        integer(c_int) function nemo2d_create(m, jpi, jpj, dx, dy, dep_const, nit000, nitend, &
                                              irecord, rdt, cbfr, visc) bind(C, name="nemo2d_create")
            import :: c_int, c_double, c_ptr
            type(c_ptr) :: m
            integer(c_int), value :: jpi, jpj, nit000, nitend, irecord
            real(c_double), value :: dx, dy, dep_const, rdt, cbfr, visc
        end function

        integer(c_int) function nemo2d_set_grid(m, pt, e1t, e2t, e1u, e2u, e1v, e2v, e12t, e12u, &
                                                e12v, gphiu, gphiv, ht, hu, hv) &
                                                bind(C, name="nemo2d_set_grid")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: m
            integer(c_int) :: pt(*)
            real(c_double) :: e1t(*), e2t(*), e1u(*), e2u(*), e1v(*), e2v(*)
            real(c_double) :: e12t(*), e12u(*), e12v(*), gphiu(*), gphiv(*), ht(*), hu(*), hv(*)
        end function

        integer(c_int) function nemo2d_set_state(m, sshn, sshn_u, sshn_v, un, vn) &
                                                 bind(C, name="nemo2d_set_state")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: m
            real(c_double) :: sshn(*), sshn_u(*), sshn_v(*), un(*), vn(*)
        end function

        integer(c_int) function nemo2d_continuity(m) bind(C, name="nemo2d_continuity")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function
        integer(c_int) function nemo2d_momentum_u(m) bind(C, name="nemo2d_momentum_u")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function
        integer(c_int) function nemo2d_momentum_v(m) bind(C, name="nemo2d_momentum_v")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function
        integer(c_int) function nemo2d_bc_ssh(m, rtime) bind(C, name="nemo2d_bc_ssh")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double), value :: rtime
        end function
        integer(c_int) function nemo2d_bc_solid_u(m) bind(C, name="nemo2d_bc_solid_u")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function
        integer(c_int) function nemo2d_bc_solid_v(m) bind(C, name="nemo2d_bc_solid_v")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function
        integer(c_int) function nemo2d_bc_flather_u(m) bind(C, name="nemo2d_bc_flather_u")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function
        integer(c_int) function nemo2d_bc_flather_v(m) bind(C, name="nemo2d_bc_flather_v")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function
        integer(c_int) function nemo2d_next(m) bind(C, name="nemo2d_next")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function

        integer(c_int) function nemo2d_step(m, istp) bind(C, name="nemo2d_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
            integer(c_int), value :: istp
        end function
        integer(c_int) function nemo2d_run(m, nit000, nitend) bind(C, name="nemo2d_run")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
            integer(c_int), value :: nit000, nitend
        end function

        integer(c_int) function nemo2d_record_checksums(m, capacity) &
                                                        bind(C, name="nemo2d_record_checksums")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
            integer(c_int), value :: capacity
        end function
        integer(c_int) function nemo2d_checksum_log(m, out, max_steps, nsteps) &
                                                    bind(C, name="nemo2d_checksum_log")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double) :: out(*)
            integer(c_int), value :: max_steps
            integer(c_int) :: nsteps
        end function
        integer(c_int) function nemo2d_checksums(m, out) bind(C, name="nemo2d_checksums")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double) :: out(3)
        end function

        integer(c_int) function nemo2d_retrieve_results(m, sshn, un, vn, ua, va) &
                                                        bind(C, name="nemo2d_retrieve_results")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double) :: sshn(*), un(*), vn(*), ua(*), va(*)
        end function
        integer(c_int) function nemo2d_retrieve_aux(m, sshn_u, sshn_v, ssha) &
                                                    bind(C, name="nemo2d_retrieve_aux")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double) :: sshn_u(*), sshn_v(*), ssha(*)
        end function

        integer(c_int) function nemo2d_timer_start(m) bind(C, name="nemo2d_timer_start")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function
        integer(c_int) function nemo2d_timer_stop(m, ms) bind(C, name="nemo2d_timer_stop")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double) :: ms
        end function
        integer(c_int) function nemo2d_destroy(m) bind(C, name="nemo2d_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function

        ! record the nine stage calls and run them with next as one fused
        ! pass (on /= 0, the default) or run each call at once
        integer(c_int) function nemo2d_set_stage_fusion(m, on) bind(C, name="nemo2d_set_stage_fusion")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
            integer(c_int), value :: on
        end function
        ! device timers of the Continuity/Momentum/BCs/Next regions and the
        ! fused step (regions 1..5 of the arrays)
        integer(c_int) function nemo2d_stage_timers(m, on) bind(C, name="nemo2d_stage_timers")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
            integer(c_int), value :: on
        end function
        integer(c_int) function nemo2d_stage_times(m, counts, total_s, stderr_s) &
            bind(C, name="nemo2d_stage_times")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            integer(c_int) :: counts(*)
            real(c_double) :: total_s(*), stderr_s(*)
        end function

        ! go2d snapshots written behind the steps (xt, yt, ht: T arrays)
        integer(c_int) function nemo2d_set_output_grid(m, xt, yt, ht) &
            bind(C, name="nemo2d_set_output_grid")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double) :: xt(*), yt(*), ht(*)
        end function
        integer(c_int) function nemo2d_snapshot_begin(m, istp, dir) bind(C, name="nemo2d_snapshot_begin")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: m
            integer(c_int), value :: istp
            character(kind=c_char) :: dir(*)
        end function
        integer(c_int) function nemo2d_snapshot_wait(m) bind(C, name="nemo2d_snapshot_wait")
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function

        type(c_ptr) function nemo2d_last_error_c() bind(C, name="nemo2d_last_error")
            import :: c_ptr
        end function
    end interface

contains

    !> Stop with the library's message when a call failed.
    subroutine nemo2d_check(rc, what)
        use, intrinsic :: iso_c_binding, only: c_f_pointer
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer :: k
        if (rc == 0) return
        call c_f_pointer(nemo2d_last_error_c(), msg, [512])
        text = ' '
        do k = 1, 512
            if (msg(k) == achar(0)) exit
            text(k:k) = msg(k)
        end do
        write (*, '(A,A,A,I0,A,A)') 'nemo2d: ', what, ' failed (status ', rc, '): ', trim(text)
        stop 1
    end subroutine nemo2d_check

end module nemo2d_api
