!> NEMOLite2D driver for the nemo2d gfx950 library.
!!
!! Keeps the serial model's program structure (original/nemolite2d.f90:1-90):
!! the &namctl namelist read from ./namelist, the grid and initial state
!! built on the host, the go2d_NNNNN.dat snapshots at istp=0 and every
!! irecord steps, and the final 'ua checksum' / 'va checksum' lines.  The step
!! calls the library's nine stage entry points in the serial model's order
!! (continuity, momentum_u, momentum_v, bc_ssh, bc_solid_u, bc_solid_v,
!! bc_flather_u, bc_flather_v, next) or, with fused=.true. in the optional
!! &namhip group, the single fused step.  The library records the nine calls
!! and runs them with next as one fused pass (stage_fusion=.false. runs a
!! kernel per call instead; the results are the same).  Per-step checksums sum|ssh|, sum|u|,
!! sum|v| are written to the file named by csfile.  With async_output=.true.
!! the snapshots are copied and written by the library behind the following
!! steps (same files) instead of retrieve-then-write.
program nemolite2d_hip
    use, intrinsic :: iso_c_binding, only: c_int, c_double, c_ptr, c_null_ptr
    use nemo2d_api
    implicit none

    integer, parameter :: wp = c_double

    integer(c_int), allocatable :: pt(:, :)
    real(wp), allocatable :: e1t(:, :), e2t(:, :), e1u(:, :), e2u(:, :), e1v(:, :), e2v(:, :)
    real(wp), allocatable :: e12t(:, :), e12u(:, :), e12v(:, :), gphiu(:, :), gphiv(:, :)
    real(wp), allocatable :: xt(:, :), yt(:, :), ht(:, :), hu(:, :), hv(:, :)
    real(wp), allocatable :: sshn(:, :), sshn_u(:, :), sshn_v(:, :)
    real(wp), allocatable :: un(:, :), vn(:, :), ua(:, :), va(:, :)
    real(wp), allocatable :: cslog(:)

    integer :: jpiglo, jpjglo, jpi, jpj, jphgr_msh, nit000, nitend, irecord
    real(wp) :: dx, dy, dep_const, rdt, cbfr, visc
    logical :: fused, write_output, async_output, stage_fusion
    character(len=256) :: csfile, grid_file

    type(c_ptr) :: model = c_null_ptr
    integer :: istp, nsteps, k
    integer(c_int) :: nlogged, tcount(5)
    integer(8) :: c0, c1, crate
    real(wp) :: t_loop, t_dev_ms, cells, ttotal(5), terr(5)
    character(len=32), parameter :: region(5) = [character(len=32) :: 'Continuity', 'Momentum', &
                                                 'BCs', 'Next', 'Fused step']

    call setup
    call grid
    call initialisation

    call nemo2d_check(nemo2d_create(model, jpi, jpj, dx, dy, dep_const, nit000, nitend, irecord, &
                                    rdt, cbfr, visc), 'nemo2d_create')
    call nemo2d_check(nemo2d_set_grid(model, pt, e1t, e2t, e1u, e2u, e1v, e2v, e12t, e12u, e12v, &
                                      gphiu, gphiv, ht, hu, hv), 'nemo2d_set_grid')
    call nemo2d_check(nemo2d_set_state(model, sshn, sshn_u, sshn_v, un, vn), 'nemo2d_set_state')
    nsteps = max(nitend - nit000 + 1, 0)
    call nemo2d_check(nemo2d_record_checksums(model, int(nsteps, c_int)), 'nemo2d_record_checksums')

    if (write_output .and. async_output) &
        call nemo2d_check(nemo2d_set_output_grid(model, xt, yt, ht), 'nemo2d_set_output_grid')
    istp = 0
    if (write_output) call snapshot

    call system_clock(c0, crate)
    call nemo2d_check(nemo2d_set_stage_fusion(model, merge(1_c_int, 0_c_int, stage_fusion)), &
                      'nemo2d_set_stage_fusion')
    call nemo2d_check(nemo2d_stage_timers(model, 1_c_int), 'nemo2d_stage_timers')
    call nemo2d_check(nemo2d_timer_start(model), 'nemo2d_timer_start')
    do istp = nit000, nitend
        call step
    end do
    call nemo2d_check(nemo2d_timer_stop(model, t_dev_ms), 'nemo2d_timer_stop')
    call system_clock(c1)
    t_loop = real(c1 - c0, wp)/real(crate, wp)

    call nemo2d_check(nemo2d_retrieve_results(model, sshn, un, vn, ua, va), 'nemo2d_retrieve_results')
    if (write_output .and. async_output) &
        call nemo2d_check(nemo2d_snapshot_wait(model), 'nemo2d_snapshot_wait')

    ! per-step checksums, then the serial model's two final lines
    allocate (cslog(3*max(nsteps, 1)))
    call nemo2d_check(nemo2d_checksum_log(model, cslog, int(nsteps, c_int), nlogged), &
                      'nemo2d_checksum_log')
    open (2, file=trim(csfile), status='replace')
    write (2, '(A)') '#  istp        sum|ssh|                 sum|u|                   sum|v|'
    do k = 1, min(int(nlogged), nsteps)
        write (2, '(I8,3ES25.16E3)') nit000 + k - 1, cslog(3*k - 2), cslog(3*k - 1), cslog(3*k)
    end do
    close (2)
    if (nsteps > 0) then
        write (*, "('ua checksum = ',E16.8)") cslog(3*nsteps - 1)
        write (*, "('va checksum = ',E16.8)") cslog(3*nsteps)
    end if

    ! timing report in dl_timer's layout (shared/dl_timer/src/dl_timer.f90:713-773)
    call nemo2d_check(nemo2d_stage_times(model, tcount, ttotal, terr), 'nemo2d_stage_times')
    cells = real(jpi, wp)*real(jpj, wp)*real(nsteps, wp)
    write (*, '(A)') '==================================== Timing report ===================================='
    write (*, '(A)') 'Timed using HIP events on the device. Units are seconds.'
    write (*, "(88('-'))")
    write (*, "('Region',26x,'Counts',6x,'Total',7x,'Average*   Avg/repeat*  Std Err')")
    write (*, "(88('-'))")
    ! labels left-justified in 32 columns, as dl_timer writes label(1:32) with (A)
    write (*, '(A,1x,I6,1x,E12.5,1x,E12.5,1x,E12.5,1x,E9.2)') label32('Time-stepping'), 1, &
        t_dev_ms*1.0e-3_wp, t_dev_ms*1.0e-3_wp, t_dev_ms*1.0e-3_wp/real(max(nsteps, 1), wp), 0.0_wp
    do k = 1, 5
        if (tcount(k) > 0) write (*, '(A,1x,I6,1x,E12.5,1x,E12.5,1x,E12.5,1x,E9.2)') region(k), &
            tcount(k), ttotal(k), ttotal(k)/real(tcount(k), wp), ttotal(k)/real(tcount(k), wp), terr(k)
    end do
    write (*, "(88('-'))")
    write (*, '(A)') '* per call, device event times (no systematic-error correction)'
    write (*, '(A,F12.3,A,ES12.4,A)') 'host wall time of the loop: ', t_loop, ' s;  device rate ', &
        cells/(t_dev_ms*1.0e-3_wp), ' cell-updates/s'
    write (*, '(A)') '========================================================================================'

    call nemo2d_check(nemo2d_destroy(model), 'nemo2d_destroy')
    write (*, *) 'Simulation finished!!'

contains

    !> A label in 32 columns, blank-padded on the right.
    pure function label32(s) result(l)
        character(len=*), intent(in) :: s
        character(len=32) :: l
        l = s
    end function label32

    !> &namctl as original/nemolite2d.f90:96-141 (same defaults, same -2
    !! correction), plus the optional &namhip group of this driver.
    subroutine setup
        integer :: ios
        namelist /namctl/ jpiglo, jpjglo, jphgr_msh, dx, dy, dep_const, &
            nit000, nitend, irecord, rdt, cbfr, visc
        namelist /namhip/ fused, write_output, async_output, stage_fusion, csfile, grid_file

        jpiglo = 258; jpjglo = 258; jphgr_msh = 1
        dx = 1000._wp; dy = 1000._wp; dep_const = 100._wp
        nit000 = 1; nitend = 1000; irecord = 1
        rdt = 10._wp; cbfr = 0.001_wp; visc = 100._wp
        fused = .false.; write_output = .true.; async_output = .false.; stage_fusion = .true.
        csfile = 'nemo2d_checksums.txt'
        grid_file = 'grid.txt'

        open (1, file='namelist', status='OLD')
        rewind (1)
        read (1, nml=namctl, iostat=ios)
        if (ios /= 0) stop "err found in reading namelist file"
        rewind (1)
        read (1, nml=namhip, iostat=ios)   ! optional group
        close (1)
        write (*, nml=namctl)
        ! jphgr_msh = 0 reads the grid from grid_file (the serial model stops
        ! there, original/nemolite2d.f90:155-158); 1 builds the uniform basin
        if (jphgr_msh /= 0 .and. jphgr_msh /= 1) stop "Wrong grid defination type, check your setup !!!!"
        jpiglo = jpiglo - 2
        jpjglo = jpjglo - 2
        jpi = jpiglo
        jpj = jpjglo
        if (jpi < 1 .or. jpj < 1) stop "grid too small: jpiglo and jpjglo must be at least 3"
    end subroutine setup

    !> Mask, metrics, latitudes and depths of the jphgr_msh=1 grid
    !! (original/nemolite2d.f90:180-279): a land ring with an open south edge.
    subroutine grid
        integer :: ji, jj
        allocate (pt(0:jpi + 1, 0:jpj + 1))
        allocate (e1t(jpi, jpj), e2t(jpi, jpj), e12t(jpi, jpj), xt(jpi, jpj), yt(jpi, jpj), ht(jpi, jpj))
        allocate (e1u(0:jpi, jpj), e2u(0:jpi, jpj), e12u(0:jpi, jpj), gphiu(0:jpi, jpj), hu(0:jpi, jpj))
        allocate (e1v(jpi, 0:jpj), e2v(jpi, 0:jpj), e12v(jpi, 0:jpj), gphiv(jpi, 0:jpj), hv(jpi, 0:jpj))
        allocate (sshn(jpi, jpj), sshn_u(0:jpi, jpj), sshn_v(jpi, 0:jpj))
        allocate (un(0:jpi, jpj), vn(jpi, 0:jpj), ua(0:jpi, jpj), va(jpi, 0:jpj))

        if (jphgr_msh == 0) then
            call read_grid_file
        else
            pt = 1
            pt(0, :) = 0
            pt(jpi + 1, :) = 0
            pt(:, jpj + 1) = 0
            pt(1:jpi, 0) = -1
            pt(0, 0) = 0
            pt(jpi + 1, 0) = 0

            e1t = dx; e2t = dy; e1u = dx; e2u = dy; e1v = dx; e2v = dy
            gphiu = 50._wp; gphiv = 50._wp
            ht = dep_const; hu = dep_const; hv = dep_const
        end if
        ! T-point coordinates: cell centres, each half a cell width on from
        ! the last (for uniform widths, the serial model's xt(ji-1) + dx)
        do jj = 1, jpj
            xt(1, jj) = 0.0_wp + 0.5_wp*e1t(1, jj)
            do ji = 2, jpi
                xt(ji, jj) = xt(ji - 1, jj) + 0.5_wp*(e1t(ji - 1, jj) + e1t(ji, jj))
            end do
        end do
        yt(:, 1) = 0.0_wp + 0.5_wp*e2t(:, 1)
        do jj = 2, jpj
            yt(:, jj) = yt(:, jj - 1) + 0.5_wp*(e2t(:, jj - 1) + e2t(:, jj))
        end do
        e12t = e1t*e2t
        e12u = e1u*e2u
        e12v = e1v*e2v
    end subroutine grid

    !> Grid from a text file (jphgr_msh = 0): list-directed, the interior
    !! sizes jpi jpj, then pt(0:jpi+1,0:jpj+1), e1t, e2t (1:jpi,1:jpj), e1u,
    !! e2u (0:jpi,1:jpj), e1v, e2v (1:jpi,0:jpj), gphiu (U), gphiv (V), ht (T),
    !! hu (U), hv (V), each in storage order (first index fastest).  The areas
    !! e12* are formed as the products, as for the built-in grid.
    subroutine read_grid_file
        integer :: ios, fi, fj
        open (3, file=trim(grid_file), status='OLD', action='READ', iostat=ios)
        if (ios /= 0) stop "grid file (grid_file in &namhip) not found"
        read (3, *, iostat=ios) fi, fj
        if (ios /= 0) stop "grid file: cannot read its sizes"
        if (fi /= jpi .or. fj /= jpj) stop "grid file: sizes differ from jpiglo-2, jpjglo-2"
        read (3, *, iostat=ios) pt, e1t, e2t, e1u, e2u, e1v, e2v, gphiu, gphiv, ht, hu, hv
        if (ios /= 0) stop "grid file: too few or unreadable values"
        close (3)
    end subroutine read_grid_file

    !> Rest state (original/nemolite2d.f90:324-371).
    subroutine initialisation
        integer :: ji, jj, i1, i2
        sshn = 0.0_wp
        do ji = 0, jpi
            do jj = 1, jpj
                i1 = min(ji + 1, jpi)
                i2 = max(ji, 1)
                sshn_u(ji, jj) = 0.5_wp*(e12t(i1, jj)*sshn(i1, jj) + e12t(i2, jj)*sshn(i2, jj))/e12u(ji, jj)
            end do
        end do
        do ji = 1, jpi
            do jj = 0, jpj
                i1 = min(jj + 1, jpj)
                i2 = max(jj, 1)
                sshn_v(ji, jj) = 0.5_wp*(e12t(ji, i1)*sshn(ji, i1) + e12t(ji, i2)*sshn(ji, i2))/e12v(ji, jj)
            end do
        end do
        un = 0.0_wp
        vn = 0.0_wp
        ua = 0.0_wp
        va = 0.0_wp
    end subroutine initialisation

    !> One step (original/nemolite2d.f90:375-386) through the library.
    subroutine step
        real(wp) :: rtime
        rtime = real(istp, wp)*rdt
        if (fused) then
            call nemo2d_check(nemo2d_step(model, istp), 'nemo2d_step')
        else
            call nemo2d_check(nemo2d_continuity(model), 'nemo2d_continuity')
            call nemo2d_check(nemo2d_momentum_u(model), 'nemo2d_momentum_u')
            call nemo2d_check(nemo2d_momentum_v(model), 'nemo2d_momentum_v')
            call nemo2d_check(nemo2d_bc_ssh(model, rtime), 'nemo2d_bc_ssh')
            call nemo2d_check(nemo2d_bc_solid_u(model), 'nemo2d_bc_solid_u')
            call nemo2d_check(nemo2d_bc_solid_v(model), 'nemo2d_bc_solid_v')
            call nemo2d_check(nemo2d_bc_flather_u(model), 'nemo2d_bc_flather_u')
            call nemo2d_check(nemo2d_bc_flather_v(model), 'nemo2d_bc_flather_v')
            call nemo2d_check(nemo2d_next(model), 'nemo2d_next')
        end if
        if (write_output .and. mod(istp, irecord) == 0) call snapshot
    end subroutine step

    !> Snapshot of the current state: retrieved and written here, or (async)
    !! queued in the library and written behind the following steps.
    subroutine snapshot
        if (async_output) then
            call nemo2d_check(nemo2d_snapshot_begin(model, int(istp, c_int), '.'//achar(0)), &
                              'nemo2d_snapshot_begin')
        else
            call nemo2d_check(nemo2d_retrieve_results(model, sshn, un, vn, ua, va), &
                              'nemo2d_retrieve_results')
            call output
        end if
    end subroutine snapshot

    !> go2d_NNNNN.dat snapshot in the serial model's format
    !! (original/nemolite2d.f90:773-799).
    subroutine output
        character(len=5) :: fname
        integer :: ji, jj
        write (fname, '(I5.5)') istp
        open (1, file='go2d_'//fname//'.dat', status='UNKNOWN')
        rewind (1)
        do jj = 1, jpj
            do ji = 1, jpi
                write (1, '(6e20.7e4)') xt(ji, jj), yt(ji, jj), ht(ji, jj), sshn(ji, jj), &
                    0.5_wp*(un(ji - 1, jj) + un(ji, jj)), 0.5_wp*(vn(ji, jj - 1) + vn(ji, jj))
            end do
            write (1, *)
        end do
        close (1)
    end subroutine output

end program nemolite2d_hip
