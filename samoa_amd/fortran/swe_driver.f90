! swe_driver.f90 -- the Fortran host: scenario set-up and the SWE time loop
! over the HIP C ABI.
!
! `swe_time_loop` is the main loop of swe_run (SWE.f90:380-429) for a frozen
! mesh: it applies pre_traversal_grid_op's clamps (SWE_euler_timestep.f90:93-99),
! calls `swe_step` where the reference calls swe%euler%traverse(grid), then
! post_traversal_grid_op's bookkeeping (:115-128: t += dt, dt = courant *
! dt_new) and flips the traversal direction (SFC_generic_traversal_ringbuffer
! .f90:420).  Without per-step host work the loop is handed to `swe_run`,
! which evaluates the same dt chain on the device.
!
! The bind(c) wrappers export the host routines to the Python harness.
module swe_driver
    use iso_c_binding
    use swe_mi355x
    use swe_vtu
    use swe_gauges
    use swe_sfc_mesh
    use swe_scenarios
    implicit none
    private
    public :: swe_time_loop, initial_dt, leg_size, adaptive_loop, tsunami_adaptive_setup

    real(c_double), parameter :: HUGE_SR = huge(1.0_c_double)

contains

    !> sqrt(0.5)**k as the reference's parameter tables hold it
    !> (SFC_data_types.f90:557).
    pure real(c_double) function leg_size(k)
        integer, intent(in) :: k
        integer :: e
        real(c_double) :: r, base
        e = abs(k)
        r = 1.0_c_double
        base = sqrt(0.5_c_double)
        do while (e > 0)
            if (iand(e, 1) == 1) r = r * base
            base = base * base
            e = ishft(e, -1)
        end do
        if (k < 0) r = 1.0_c_double / r
        leg_size = r
    end function

    !> Initial time step of the init traversal (SWE_initialize.f90:327-335,
    !> :240): courant * min over cells of vol / (sum(len) * max wave speed),
    !> in unit-square units (the reference does not apply cfg%scaling here).
    real(c_double) function initial_dt(n, depth, eta, hu, hv, b, g, courant)
        integer(c_int64_t), intent(in) :: n
        integer(c_int8_t), intent(in)  :: depth(n)
        real(c_double), intent(in)     :: eta(n), hu(n), hv(n), b(n), g, courant
        integer(c_int64_t) :: i
        real(c_double) :: h, s, dmin_, vol, sumlen
        integer :: d
        dmin_ = HUGE_SR
        do i = 1, n
            h = eta(i) - b(i)
            if (h > 0.0_c_double) then
                s = sqrt(g * h) + sqrt((hu(i) * hu(i) + hv(i) * hv(i)) / (h * h))
            else
                s = 0.0_c_double
            end if
            d = depth(i)
            vol = 0.5_c_double ** (d + 1)
            sumlen = (leg_size(d) + leg_size(d - 1)) + leg_size(d)
            dmin_ = min(dmin_, vol / (sumlen * s))
        end do
        initial_dt = courant * dmin_
    end function

    !> SWE.f90:380-429 on a frozen mesh.  Runs until time >= tmax (tmax >= 0)
    !> or nmax steps (nmax >= 0); dt, time and forward are updated in place.
    !> If host_steps is false and no time limit is set, the steps run on the
    !> device through swe_run.  Returns a SWE_* status.
    integer(c_int) function swe_time_loop(handle, courant, tmax, tout, nmax, host_steps, &
                                          dt, time, forward, steps, refine) result(rc)
        type(c_ptr), intent(in)            :: handle
        real(c_double), intent(in)         :: courant, tmax, tout
        integer(c_int64_t), intent(in)     :: nmax
        logical, intent(in)                :: host_steps
        real(c_double), intent(inout), target :: dt, time
        logical, intent(inout)             :: forward
        integer(c_int64_t), intent(out)    :: steps
        !> int8 [n] or c_null_ptr: refinement flags of the last step (host steps only)
        type(c_ptr), intent(in), optional  :: refine
        type(c_ptr) :: rf
        real(c_double), target :: dt_cfl
        real(c_double) :: clamp
        integer(c_int64_t) :: nref
        integer(c_int) :: fw

        steps = 0
        rc = SWE_OK
        rf = c_null_ptr
        if (present(refine)) rf = refine
        if (.not. host_steps .and. tmax < 0.0_c_double .and. nmax >= 0 .and. .not. c_associated(rf)) then
            clamp = -1.0_c_double
            if (tout > 0.0_c_double) clamp = tout
            if (tout > 0.0_c_double) dt = min(tout, dt)
            fw = merge(1_c_int, 0_c_int, forward)
            rc = swe_run(handle, nmax, dt, fw, clamp, c_loc(time), c_loc(dt_cfl))
            if (rc /= SWE_OK) return
            dt = dt_cfl
            steps = nmax
            if (mod(nmax, 2_c_int64_t) == 1) forward = .not. forward
            return
        end if
        do
            if ((tmax >= 0.0_c_double .and. time >= tmax) .or. (nmax >= 0 .and. steps >= nmax)) exit
            steps = steps + 1
            if (tmax > 0.0_c_double) dt = min(tmax, dt)
            if (tout > 0.0_c_double) dt = min(tout, dt)
            fw = merge(1_c_int, 0_c_int, forward)
            rc = swe_step(handle, dt, fw, dt_cfl, nref, rf)
            if (rc /= SWE_OK) return
            time = time + dt
            dt = courant * dt_cfl
            forward = .not. forward
        end do
    end function

    !> SWE.f90:380-429 with remeshing before every nadapt-th step
    !> (:385-390), the flags taken from the surface-elevation jumps between
    !> wet neighbours (refine above refine_tol, coarsen below coarsen_tol) of
    !> the state each step reads, which the step leaves on the device as the
    !> reference's Euler traversal leaves its flags (swe_keep_eta_jump_flags).
    !> k0 numbers the first step from 0, so that calls continue one another:
    !> step k remeshes when mod(k + 1, nadapt) = 0, as time step k + 1 of the
    !> reference does; before step 0 the flags are the initial phase's, all 0,
    !> so that remeshing only flips the direction.  dt > 0 starts the device dt chain
    !> at (dt, time); dt <= 0 continues it.  Nothing waits for the steps
    !> until the end, where time and dt (the next step's) are read back.
    !> cells: the cells stepped, summed over the steps.  Returns a SWE_* status.
    integer(c_int) function adaptive_loop(handle, nsteps, nadapt, refine_tol, coarsen_tol, k0, &
                                          dt, time, forward, cells) result(rc)
        type(c_ptr), intent(in)         :: handle
        integer(c_int64_t), intent(in)  :: nsteps, nadapt, k0
        real(c_double), intent(in)      :: refine_tol, coarsen_tol
        real(c_double), intent(inout), target :: dt, time
        logical, intent(inout)          :: forward
        integer(c_int64_t), intent(out) :: cells
        integer(c_int64_t) :: k, n_new
        integer(c_int) :: sweeps
        real(c_double) :: dt_first
        dt_first = dt
        cells = 0
        rc = SWE_OK
        if (nadapt > 0) then
            rc = swe_keep_eta_jump_flags(handle, refine_tol, coarsen_tol)
            if (rc /= SWE_OK) return
        end if
        do k = k0, k0 + nsteps - 1
            if (nadapt > 0) then
                if (mod(k + 1, nadapt) == 0) then
                    if (k > 0) then
                        rc = swe_adapt(handle, c_null_ptr, n_new, sweeps)
                        if (rc /= SWE_OK) return
                    end if
                    forward = .not. forward  ! the adaption traversal flips it too
                end if
            end if
            cells = cells + swe_cells(handle)
            if (k == k0 .and. dt_first > 0.0_c_double) then
                rc = swe_run(handle, 1_c_int64_t, dt_first, merge(1_c_int, 0_c_int, forward), &
                             -1.0_c_double, c_loc(time), c_null_ptr)
            else
                rc = swe_run(handle, 1_c_int64_t, -1.0_c_double, merge(1_c_int, 0_c_int, forward), &
                             -1.0_c_double, c_null_ptr, c_null_ptr)
            end if
            if (rc /= SWE_OK) return
            forward = .not. forward
        end do
        rc = swe_run(handle, 0_c_int64_t, -1.0_c_double, 0_c_int, -1.0_c_double, c_loc(time), c_loc(dt))
    end function

    !> The adaptive tsunami's initial mesh and state on `handle`, which holds
    !> a uniform mesh and any state: the initial phase of SWE.f90:230-262
    !> with the tsunami's criterion, refining every cell shallower than dmax
    !> whose corners see the source's surface elevation differ by more than
    !> tol, until none does; then the scenario state at the centroids, the
    !> bathymetry raster for the cells later remeshings split, and dt0
    !> (initial_dt).  n_out: the cells.  Returns a SWE_* status.
    integer(c_int) function tsunami_adaptive_setup(handle, seed, tol, dmax, drytol, courant, &
                                                   n_out, dt0) result(rc)
        type(c_ptr), intent(in)         :: handle
        integer(c_int64_t), intent(in)  :: seed
        real(c_double), intent(in)      :: tol, drytol, courant
        integer, intent(in)             :: dmax
        integer(c_int64_t), intent(out) :: n_out
        real(c_double), intent(out)     :: dt0
        integer(c_int8_t), allocatable, target :: depth(:), plotter(:), refine(:)
        real(c_double), allocatable, target :: coords(:, :, :), eta(:), hu(:), hv(:), b(:)
        real(c_double), allocatable, target :: grid(:, :)
        real(c_double) :: amp, ratio, x0, y0, a, c, e(3), p(2)
        integer(c_int64_t) :: n, i, n_new
        integer :: depth_i, j, ierr
        integer(c_int) :: sweeps
        call tsunami_draw(seed, amp, ratio, x0, y0, depth_i, dmax, dmax)
        a = 1.0e5_c_double
        c = a / ratio
        do
            n = swe_cells(handle)
            if (allocated(depth)) deallocate(depth, plotter, refine, coords)
            allocate(depth(n), plotter(n), refine(n), coords(2, 3, n))
            rc = swe_download_mesh(handle, c_loc(depth), c_loc(plotter), c_null_ptr)
            if (rc /= SWE_OK) return
            call sfc_mesh_from_depths(n, depth, plotter, coords, ierr)
            if (ierr /= 0) then
                rc = SWE_ERR_STATE
                return
            end if
            do i = 1, n
                refine(i) = 0
                if (depth(i) >= dmax) cycle
                do j = 1, 3
                    p = TSU_SCALING * coords(:, j, i)
                    e(j) = amp * exp(-(((p(1) - x0) / a) ** 2 + ((p(2) - y0) / c) ** 2))
                end do
                if (maxval(e) - minval(e) > tol) refine(i) = 1
            end do
            if (count(refine == 1) == 0) exit
            rc = swe_adapt(handle, c_loc(refine), n_new, sweeps)
            if (rc /= SWE_OK) return
        end do
        allocate(eta(n), hu(n), hv(n), b(n), grid(TSU_G, TSU_G))
        call tsunami_state(n, coords, depth, int(dmax), seed, amp, ratio, x0, y0, drytol, eta, hu, hv, b)
        rc = swe_upload(handle, eta, hu, hv, b)
        if (rc /= SWE_OK) return
        call tsunami_raster(seed, grid)
        rc = swe_set_bathymetry_raster(handle, int(TSU_G, c_int64_t), int(TSU_G, c_int64_t), &
                                       0.0_c_double, 0.0_c_double, TSU_DX, TSU_DX, 0.0_c_double, &
                                       0.0_c_double, c_loc(grid))
        if (rc /= SWE_OK) return
        dt0 = initial_dt(n, depth, eta, hu, hv, b, 9.80665_c_double, courant)
        n_out = n
    end function

    ! ------------------------------------------------------------------
    ! C entry points for the Python harness

    integer(c_int) function swe_host_adaptive_loop(handle, nsteps, nadapt, refine_tol, coarsen_tol, k0, &
                                                   dt, time, forward, cells) bind(c, name='swe_host_adaptive_loop')
        type(c_ptr), value               :: handle
        integer(c_int64_t), value        :: nsteps, nadapt, k0
        real(c_double), value            :: refine_tol, coarsen_tol
        real(c_double), intent(inout)    :: dt, time
        integer(c_int), intent(inout)    :: forward
        integer(c_int64_t), intent(out)  :: cells
        logical :: fw
        fw = forward /= 0
        swe_host_adaptive_loop = adaptive_loop(handle, nsteps, nadapt, refine_tol, coarsen_tol, k0, &
                                               dt, time, fw, cells)
        forward = merge(1_c_int, 0_c_int, fw)
    end function

    integer(c_int) function swe_host_tsunami_adaptive_setup(handle, seed, tol, dmax, drytol, courant, &
                                                            n_out, dt0) bind(c, name='swe_host_tsunami_adaptive_setup')
        type(c_ptr), value               :: handle
        integer(c_int64_t), value        :: seed
        real(c_double), value            :: tol, drytol, courant
        integer(c_int), value            :: dmax
        integer(c_int64_t), intent(out)  :: n_out
        real(c_double), intent(out)      :: dt0
        swe_host_tsunami_adaptive_setup = tsunami_adaptive_setup(handle, seed, tol, int(dmax), drytol, &
                                                                 courant, n_out, dt0)
    end function

    real(c_double) function swe_host_initial_dt(n, depth, eta, hu, hv, b, g, courant) &
            bind(c, name='swe_host_initial_dt')
        integer(c_int64_t), value     :: n
        integer(c_int8_t), intent(in) :: depth(n)
        real(c_double), intent(in)    :: eta(n), hu(n), hv(n), b(n)
        real(c_double), value         :: g, courant
        swe_host_initial_dt = initial_dt(n, depth, eta, hu, hv, b, g, courant)
    end function

    subroutine swe_host_reference_dam_break(n, coords, eta, hu, hv, b) &
            bind(c, name='swe_host_reference_dam_break')
        integer(c_int64_t), value   :: n
        real(c_double), intent(in)  :: coords(2, 3, n)
        real(c_double), intent(out) :: eta(n), hu(n), hv(n), b(n)
        integer(c_int64_t) :: i
        real(c_double) :: x(2)
        do i = 1, n
            x = 10.0_c_double * centroid(coords(:, :, i)) - 5.0_c_double
            eta(i) = merge(1.0_c_double, 0.0_c_double, x(1) < 0.0_c_double)
        end do
        hu = 0.0_c_double
        hv = 0.0_c_double
        b = 0.0_c_double
    end subroutine

    subroutine swe_host_radial_dam_break(n, coords, eta, hu, hv, b) &
            bind(c, name='swe_host_radial_dam_break')
        integer(c_int64_t), value   :: n
        real(c_double), intent(in)  :: coords(2, 3, n)
        real(c_double), intent(out) :: eta(n), hu(n), hv(n), b(n)
        integer(c_int64_t) :: i
        real(c_double) :: x(2)
        do i = 1, n
            x = centroid(coords(:, :, i))
            eta(i) = merge(10.0_c_double, 5.0_c_double, norm2(x - 0.5_c_double) < 0.1_c_double)
        end do
        hu = 0.0_c_double
        hv = 0.0_c_double
        b = 0.0_c_double
    end subroutine

    subroutine swe_host_tsunami_draw(seed, dlo, dhi, amp, ratio, x0, y0, depth) &
            bind(c, name='swe_host_tsunami_draw')
        integer(c_int64_t), value   :: seed
        integer(c_int), value       :: dlo, dhi
        real(c_double), intent(out) :: amp, ratio, x0, y0
        integer(c_int), intent(out) :: depth
        integer :: d
        call tsunami_draw(seed, amp, ratio, x0, y0, d, int(dlo), int(dhi))
        depth = int(d, c_int)
    end subroutine

    !> The tsunami scenario's bathymetry raster (tsunami_raster): ng x ng
    !> values, x fastest, node spacing dx from the origin.  Returns ng.
    integer(c_int) function swe_host_tsunami_raster(seed, values, dx) &
            bind(c, name='swe_host_tsunami_raster')
        integer(c_int64_t), value   :: seed
        real(c_double), intent(out) :: values(TSU_G, TSU_G)
        real(c_double), intent(out) :: dx
        call tsunami_raster(seed, values)
        dx = TSU_DX
        swe_host_tsunami_raster = TSU_G
    end function

    subroutine swe_host_tsunami(n, coords, depth, dmax, seed, amp, ratio, x0, y0, drytol, eta, hu, &
                                hv, b) bind(c, name='swe_host_tsunami')
        integer(c_int64_t), value   :: n, seed
        real(c_double), intent(in)  :: coords(2, 3, n)
        integer(c_int8_t), intent(in) :: depth(n)
        integer(c_int), value       :: dmax
        real(c_double), value       :: amp, ratio, x0, y0, drytol
        real(c_double), intent(out) :: eta(n), hu(n), hv(n), b(n)
        call tsunami_state(n, coords, depth, int(dmax), seed, amp, ratio, x0, y0, drytol, eta, hu, hv, b)
    end subroutine

    !> The time loop for a handle created by the caller.  forward is in/out
    !> (0/1).  Returns a SWE_* status.
    integer(c_int) function swe_host_time_loop(handle, courant, tmax, tout, nmax, host_steps, &
                                               dt, time, forward, steps) bind(c, name='swe_host_time_loop')
        type(c_ptr), value                :: handle
        real(c_double), value             :: courant, tmax, tout
        integer(c_int64_t), value         :: nmax
        integer(c_int), value             :: host_steps
        real(c_double), intent(inout)     :: dt, time
        integer(c_int), intent(inout)     :: forward
        integer(c_int64_t), intent(out)   :: steps
        logical :: fw
        fw = forward /= 0
        swe_host_time_loop = swe_time_loop(handle, courant, tmax, tout, nmax, host_steps /= 0, &
                                           dt, time, fw, steps)
        forward = merge(1_c_int, 0_c_int, fw)
    end function
    !> VTU file of a cell state (swe_vtu); path is a NUL-terminated string.
    !> forward: 1 if the output traversal runs forward.  Returns 0 or an I/O
    !> status.
    integer(c_int) function swe_host_write_vtu(path, n, coords, plotter, depth, eta, b, hu, hv, &
                                               refine, time, scaling, offset, forward) &
            bind(c, name='swe_host_write_vtu')
        character(kind=c_char), intent(in) :: path(*)
        integer(c_int64_t), value          :: n
        real(c_double), intent(in)         :: coords(2, 3, n), eta(n), b(n), hu(n), hv(n)
        integer(c_int8_t), intent(in)      :: plotter(n), depth(n), refine(n)
        real(c_double), value              :: time, scaling, offset
        integer(c_int), value              :: forward
        integer :: ierr
        call vtu_write(c_string(path), n, coords, plotter, depth, eta, b, hu, hv, refine, time, &
                       scaling, offset, forward /= 0, ierr)
        swe_host_write_vtu = ierr
    end function

    !> The single-piece .pvtu master file naming `piece`.
    integer(c_int) function swe_host_write_pvtu(path, piece) bind(c, name='swe_host_write_pvtu')
        character(kind=c_char), intent(in) :: path(*), piece(*)
        integer :: ierr
        call pvtu_write(c_string(path), c_string(piece), ierr)
        swe_host_write_pvtu = ierr
    end function

    !> Probe values of a cell state (swe_gauges): out(8, npt) =
    !> x_unit, y_unit, p(1), p(2), h, b, centroid distance, time in days.
    subroutine swe_host_gauges_sample(n, coords, eta, hu, hv, b, npt, px, py, scaling, offset, &
                                      time, out) bind(c, name='swe_host_gauges_sample')
        integer(c_int64_t), value      :: n
        real(c_double), intent(in)     :: coords(2, 3, n), eta(n), hu(n), hv(n), b(n)
        integer(c_int), value          :: npt
        real(c_double), intent(in)     :: px(npt), py(npt)
        real(c_double), value          :: scaling, offset, time
        real(c_double), intent(out)    :: out(8, npt)
        call gauges_sample(n, coords, eta, hu, hv, b, int(npt), px, py, scaling, offset, time, out)
    end subroutine

    !> The probe file; full /= 0 writes 17 significant digits instead of F0.15.
    integer(c_int) function swe_host_gauges_write(path, npt, px, py, out, full) &
            bind(c, name='swe_host_gauges_write')
        character(kind=c_char), intent(in) :: path(*)
        integer(c_int), value          :: npt, full
        real(c_double), intent(in)     :: px(npt), py(npt), out(8, npt)
        integer :: ierr
        call gauges_write(c_string(path), int(npt), px, py, out, full /= 0, ierr)
        swe_host_gauges_write = ierr
    end function

    function c_string(cs) result(s)
        character(kind=c_char), intent(in) :: cs(*)
        character(len=:), allocatable :: s
        integer :: k
        k = 0
        do while (cs(k + 1) /= c_null_char)
            k = k + 1
        end do
        allocate(character(len=k) :: s)
        s = transfer(cs(1:k), s)
    end function
end module swe_driver
