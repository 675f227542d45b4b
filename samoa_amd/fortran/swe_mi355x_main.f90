! swe_mi355x_main.f90 -- command-line SWE run on the MI355X, the counterpart
! of the reference's `samoa` binary built with scenario=swe (SFC_main.f90,
! SWE.f90:205-444).
!
!   swe_mi355x -dmax 14 -nmax 100 [-dmin d] [-dstart d] [-nadapt k]
!              [-tmax t] [-tout t] [-nout k]
!              [-courant c] [-drytolerance d] [-scenario dambreak|radial|tsunami]
!              [-seed s] [-boundary outflow|wall]
!              [-solver aug_riemann|fwave|lf|llf|lfbath|llfbath]
!              [-dump prefix] [-hoststeps] [-xmloutput .true.] [-output_dir dir]
!              [-stestpoints "x y,x y,..."] [-pointoutput_full .true.]
!              [-dam_left hL] [-dam_right hR] [-refine_tol m]
!
! The mesh starts uniform at depth -dstart (default: -dmax).  Below -dmax the
! initial phase of SWE.f90:230-262 refines it: a cell shallower than -dmin,
! or one whose corners see different initial water heights
! (SWE_initialize.f90:292-320), is refined, closed by the conformity rules,
! until no cell asks (dambreak and radial scenarios); the tsunami scenario
! refines where its source varies by more than -refine_tol metres across a
! cell instead.  With -nadapt k the mesh is remeshed on the device before
! every k-th step from the refinement flags of the step before
! (SWE.f90:387-390), or for the tsunami from the surface-elevation jumps
! between wet neighbours (refine above -refine_tol / 2, coarsen below
! -refine_tol / 16); each remeshing flips the traversal direction, as the
! reference's adaption traversal does.  The dam
! break's heights default to the reference's 1 m | dry (SWE_initialize.f90:354).
!
! Like the reference CLI (Config.f90:213-296, M_kracken rget), -courant,
! -drytolerance, -tmax and -tout are read as default REAL before they are
! stored in double precision.  With -nout k, every k-th step is followed by an
! output traversal: the direction flips once more and, with -dump, the state
! is written to <prefix>_<i>.bin (int64 n, float64 time, eta, hu, hv, b);
! with -xmloutput .true., <output_dir>/swe_<i>_r0_s1.vtu and swe_<i>.pvtu are
! written in the reference's VTK layout (SWE_xml_output.f90), cells in the
! order of the output traversal; with -stestpoints, the probe file
! <output_dir>/swe_d<dmax>_cou<c>_dry<d>_pointoutput_<i>.txt
! (SWE_point_output.f90; 17 significant digits with -pointoutput_full).  As
! in SWE.f90:405-418, every output traversal flips the direction: the VTK (or
! -dump) output once, the probe output once more.
program swe_mi355x_main
    use iso_c_binding
    use swe_mi355x
    use swe_sfc_mesh
    use swe_scenarios
    use swe_driver
    use swe_vtu
    use swe_gauges
    implicit none

    integer :: dmax, dmin, dstart, nadapt, nout, i, narg, ierr, depth_i
    integer(c_int64_t) :: nmax, n, seed, steps, k, nsteps_done, chunk
    real :: r_courant, r_drytol, r_tmax, r_tout
    real(c_double) :: courant, drytol, tmax, tout, scaling, offset, dt, time, amp, ratio, x0, y0
    real(c_double) :: dam_left, dam_right, refine_tol
    integer(c_int64_t) :: n_new
    integer(c_int) :: sweeps
    character(len=256) :: arg, val, scenario, dump, bcname, solvername, outdir
    character(len=4096) :: probes
    logical :: host_steps, forward, xml, probe_full
    real(c_double), allocatable :: px(:), py(:), pvals(:, :)
    integer :: npt
    integer(c_int8_t), allocatable, target :: plotter(:), depth(:), refine(:)
    integer(c_int32_t), allocatable, target :: nbr(:, :)
    real(c_double), allocatable, target :: coords(:, :, :), eta(:), hu(:), hv(:), b(:)
    type(swe_mesh_desc) :: desc
    type(c_ptr) :: h
    integer(c_int) :: rc
    integer(c_int64_t) :: t0, t1, rate
    integer :: iout

    dmax = 14; dmin = -1; dstart = -1; nadapt = -1; nmax = -1; nout = -1
    dam_left = 1.0_c_double; dam_right = 0.0_c_double; refine_tol = 1.5e-3_c_double
    r_courant = 0.45; r_drytol = 0.01; r_tmax = 3600.0; r_tout = -1.0
    scenario = 'dambreak'; dump = ''; bcname = 'outflow'; solvername = 'aug_riemann'
    seed = 0; host_steps = .false.; xml = .false.; outdir = 'output'; probes = ''
    probe_full = .false.
    narg = command_argument_count()
    i = 1
    do while (i <= narg)
        call get_command_argument(i, arg)
        val = ''
        if (i < narg) call get_command_argument(i + 1, val)
        select case (trim(arg))
        case ('-dmax'); read(val, *) dmax; i = i + 1
        case ('-dstart'); read(val, *) dstart; i = i + 1
        case ('-dmin'); read(val, *) dmin; i = i + 1
        case ('-nadapt'); read(val, *) nadapt; i = i + 1
        case ('-dam_left'); read(val, *) dam_left; i = i + 1
        case ('-dam_right'); read(val, *) dam_right; i = i + 1
        case ('-refine_tol'); read(val, *) refine_tol; i = i + 1
        case ('-nmax'); read(val, *) nmax; i = i + 1
        case ('-nout'); read(val, *) nout; i = i + 1
        case ('-tmax'); read(val, *) r_tmax; i = i + 1
        case ('-tout'); read(val, *) r_tout; i = i + 1
        case ('-courant'); read(val, *) r_courant; i = i + 1
        case ('-drytolerance'); read(val, *) r_drytol; i = i + 1
        case ('-scenario'); scenario = val; i = i + 1
        case ('-seed'); read(val, *) seed; i = i + 1
        case ('-boundary'); bcname = val; i = i + 1
        case ('-solver'); solvername = val; i = i + 1
        case ('-dump'); dump = val; i = i + 1
        case ('-hoststeps'); host_steps = .true.
        case ('-xmloutput'); read(val, *) xml; i = i + 1
        case ('-output_dir'); outdir = val; i = i + 1
        case ('-stestpoints'); probes = val; i = i + 1
        case ('-pointoutput_full'); read(val, *) probe_full; i = i + 1
        case default
            print '(A, A)', 'swe_mi355x: unknown argument ', trim(arg)
            stop 2
        end select
        i = i + 1
    end do
    if (dmin < 0) dmin = dmax
    if (dstart < 0) dstart = dmax
    courant = r_courant; drytol = r_drytol; tmax = r_tmax; tout = r_tout
    call parse_probes(probes)
    select case (trim(scenario))
    case ('dambreak')
        scaling = 10.0_c_double; offset = -5.0_c_double
    case ('radial')
        scaling = 1.0_c_double; offset = 0.0_c_double
    case ('tsunami')
        scaling = 1.0e6_c_double; offset = 0.0_c_double
    case default
        print '(A, A)', 'swe_mi355x: unknown scenario ', trim(scenario)
        stop 2
    end select

    n = ishft(1_c_int64_t, dstart + 1)
    allocate(plotter(n), depth(n), nbr(3, n), coords(2, 3, n), eta(n), hu(n), hv(n), b(n), refine(n))
    refine = 0
    call sfc_uniform_mesh(dstart, plotter, nbr, coords, ierr)
    if (ierr /= 0) then
        print '(A, I0)', 'swe_mi355x: mesh generation failed: ', ierr
        stop 1
    end if
    depth = int(dstart, c_int8_t)

    desc%n_cells = n
    desc%depth = c_loc(depth)
    desc%plotter = c_loc(plotter)
    desc%nbr = c_loc(nbr)
    desc%scaling = scaling
    desc%g = 9.80665_c_double
    desc%dry_tolerance = drytol
    desc%courant = courant
    desc%min_depth = dmin
    desc%max_depth = dmax
    ! the reference's flux_solver names (SConstruct:45-46, :224-237)
    select case (trim(solvername))
    case ('aug_riemann'); desc%solver = SWE_SOLVER_AUG_RIEMANN
    case ('fwave'); desc%solver = SWE_SOLVER_FWAVE
    case ('lf'); desc%solver = SWE_SOLVER_LF
    case ('llf'); desc%solver = SWE_SOLVER_LLF
    case ('lfbath'); desc%solver = SWE_SOLVER_LF_BATH
    case ('llfbath'); desc%solver = SWE_SOLVER_LLF_BATH
    case default
        write(*, '(A, A)') 'unknown -solver ', trim(solvername)
        stop 2
    end select
    desc%boundary = merge(SWE_BC_WALL, SWE_BC_OUTFLOW, trim(bcname) == 'wall')
    rc = swe_create(desc, h)
    call check(rc, 'swe_create')

    ! initial phase (SWE.f90:230-262): refine until no cell asks
    if (trim(scenario) == 'tsunami' .and. dstart < dmax) then
        ! the tsunami refines where the source varies by more than -refine_tol
        ! across a cell, and samples b from its raster for later splits
        eta = 0; hu = 0; hv = 0; b = 0
        rc = swe_upload(h, eta, hu, hv, b)
        call check(rc, 'swe_upload')
        rc = tsunami_adaptive_setup(h, seed, refine_tol, dmax, drytol, courant, n_new, dt)
        call check(rc, 'tsunami_adaptive_setup')
        call sync_mesh()
        rc = swe_download(h, c_loc(eta), c_loc(hu), c_loc(hv), c_loc(b))
        call check(rc, 'swe_download')
    else if (dstart < dmax .or. dstart < dmin) then
        eta = 0; hu = 0; hv = 0; b = 0
        rc = swe_upload(h, eta, hu, hv, b)
        call check(rc, 'swe_upload')
        do
            do k = 1, n
                refine(k) = 0
                if (depth(k) < dmin) then
                    refine(k) = 1
                else if (depth(k) < dmax) then
                    if (corners_differ(coords(:, :, k))) refine(k) = 1
                end if
            end do
            if (count(refine(1:n) == 1) == 0) exit
            rc = swe_adapt(h, c_loc(refine), n_new, sweeps)
            call check(rc, 'swe_adapt')
            call sync_mesh()
        end do
        refine = 0
    end if
    select case (trim(scenario))
    case ('dambreak')
        call swe_host_reference_dam_break_f(n, coords, eta, hu, hv, b)
    case ('radial')
        do k = 1, n
            eta(k) = radial_eta(centroid(coords(:, :, k)))
        end do
        hu = 0; hv = 0; b = 0
    case ('tsunami')
        if (dstart == dmax) then
            call tsunami_draw(seed, amp, ratio, x0, y0, depth_i, dmax, dmax)
            call tsunami_state(n, coords, depth, dmax, seed, amp, ratio, x0, y0, drytol, eta, hu, hv, b)
        end if
    end select
    if (.not. (trim(scenario) == 'tsunami' .and. dstart < dmax)) then
        rc = swe_upload(h, eta, hu, hv, b)
        call check(rc, 'swe_upload')
    end if
    if (nadapt > 0) then
        if (trim(scenario) == 'tsunami') then
            ! |grad eta| of the state each step reads: a jump to a neighbour
            ! spans about half a corner range
            rc = swe_keep_eta_jump_flags(h, refine_tol / 2.0_c_double, refine_tol / 16.0_c_double)
            call check(rc, 'swe_keep_eta_jump_flags')
        else
            rc = swe_keep_refinement_flags(h, 1_c_int)
            call check(rc, 'swe_keep_refinement_flags')
        end if
    end if

    ! init_grid leaves the grid reversed; init_b runs backward, init_dofs
    ! forward (SWE.f90:230-234); the initial output traversal flips once more.
    forward = .false.
    time = 0.0_c_double
    dt = initial_dt(n, depth, eta, hu, hv, b, 9.80665_c_double, courant)
    iout = 0
    if (nout > 0) then
        call output(iout)
        iout = iout + 1
    end if

    call system_clock(t0, rate)
    nsteps_done = 0
    do
        if ((tmax >= 0.0_c_double .and. time >= tmax) .or. (nmax >= 0 .and. nsteps_done >= nmax)) exit
        chunk = -1
        if (nout > 0) chunk = int(nout, c_int64_t)
        if (nmax >= 0) then
            if (chunk < 0) then
                chunk = nmax - nsteps_done
            else
                chunk = min(chunk, nmax - nsteps_done)
            end if
        end if
        if (nadapt > 0) then
            ! one step at a time, remeshing before every nadapt-th (SWE.f90:385-390);
            ! before the first step the flags are the initial phase's, all 0
            if (mod(nsteps_done + 1, int(nadapt, c_int64_t)) == 0) then
                if (nsteps_done == 0) then
                    refine = 0
                    rc = swe_adapt(h, c_loc(refine), n_new, sweeps)
                else
                    rc = swe_adapt(h, c_null_ptr, n_new, sweeps)
                end if
                call check(rc, 'swe_adapt')
                forward = .not. forward
                if (n_new /= size(refine, kind=c_int64_t)) then
                    deallocate(refine)
                    allocate(refine(n_new))
                end if
            end if
            rc = swe_time_loop(h, courant, tmax, tout, 1_c_int64_t, .true., dt, time, forward, steps, &
                               c_loc(refine))
        else if (xml .and. dmin < dmax) then
            ! the output carries the refinement flags of the last step
            rc = swe_time_loop(h, courant, tmax, tout, chunk, .true., dt, time, forward, steps, &
                               c_loc(refine))
        else
            rc = swe_time_loop(h, courant, tmax, tout, chunk, host_steps .or. tmax >= 0.0_c_double, &
                               dt, time, forward, steps)
        end if
        call check(rc, 'swe_time_loop')
        nsteps_done = nsteps_done + steps
        if (nout > 0 .and. mod(nsteps_done, int(nout, c_int64_t)) == 0) then
            if (nadapt > 0 .and. (len_trim(dump) > 0 .or. xml .or. npt > 0)) call sync_mesh()
            if (len_trim(dump) > 0 .or. xml .or. npt > 0) then
                rc = swe_download(h, c_loc(eta), c_loc(hu), c_loc(hv), c_loc(b))
                call check(rc, 'swe_download')
            end if
            call output(iout)
            iout = iout + 1
        end if
        if (steps == 0) exit
    end do
    call system_clock(t1)
    if (nadapt > 0) call sync_mesh()
    rc = swe_download(h, c_loc(eta), c_loc(hu), c_loc(hv), c_loc(b))
    call check(rc, 'swe_download')
    if (len_trim(dump) > 0 .and. nout <= 0) call write_dump(0)
    print '(A, I0, A, I0, A, ES22.15)', ' SWE: done. cells: ', n, ' steps: ', nsteps_done, ' time: ', time
    if (t1 > t0) print '(A, F14.4, A)', ' Cell update throughput: ', &
        1.0d-6 * real(n * nsteps_done, c_double) / (real(t1 - t0, c_double) / real(rate, c_double)), ' M/s'
    rc = swe_destroy(h)

contains

    subroutine check(code, what)
        integer(c_int), intent(in) :: code
        character(*), intent(in)   :: what
        if (code /= SWE_OK) then
            print '(A, A, A, I0, A, A)', 'swe_mi355x: ', what, ' failed (', code, '): ', trim(swe_error_message())
            stop 1
        end if
    end subroutine

    ! The handle's current mesh on the host: depths, plotter types and
    ! neighbours from the device, corners from the depth sequence; the state
    ! arrays resized to match (their contents are downloaded when needed).
    ! The refinement flags of the last step are kept.
    subroutine sync_mesh()
        integer(c_int8_t), allocatable :: keep(:)
        n = swe_cells(h)
        if (size(depth, kind=c_int64_t) /= n) then
            allocate(keep(min(n, size(refine, kind=c_int64_t))))
            keep = refine(1:size(keep))
            deallocate(plotter, depth, nbr, coords, eta, hu, hv, b, refine)
            allocate(plotter(n), depth(n), nbr(3, n), coords(2, 3, n), eta(n), hu(n), hv(n), b(n), &
                     refine(n))
            refine = 0
            refine(1:size(keep)) = keep
        end if
        rc = swe_download_mesh(h, c_loc(depth), c_loc(plotter), c_loc(nbr))
        call check(rc, 'swe_download_mesh')
        call sfc_mesh_from_depths(n, depth, plotter, coords, ierr)
        if (ierr /= 0) then
            print '(A)', 'swe_mi355x: the device mesh is not a leaf sequence'
            stop 1
        end if
    end subroutine

    ! The initial water height of the scenario at a point of the unit square.
    real(c_double) function initial_h(x)
        real(c_double), intent(in) :: x(2)
        real(c_double) :: xs(2)
        if (trim(scenario) == 'radial') then
            initial_h = radial_eta(x)
        else
            xs = scaling * x + offset
            initial_h = merge(dam_left, dam_right, xs(1) < 0.0_c_double)
        end if
    end function

    ! SWE_initialize.f90:308-319: refine where the corners see different
    ! initial water heights.
    logical function corners_differ(x)
        real(c_double), intent(in) :: x(2, 3)
        real(c_double) :: h1, h2, h3
        h1 = initial_h(x(:, 1))
        h2 = initial_h(x(:, 2))
        h3 = initial_h(x(:, 3))
        corners_differ = max(h1, h2, h3) > min(h1, h2, h3)
    end function

    real(c_double) function radial_eta(x)
        real(c_double), intent(in) :: x(2)
        radial_eta = merge(10.0_c_double, 5.0_c_double, norm2(x - 0.5_c_double) < 0.1_c_double)
    end function

    subroutine write_dump(idx)
        integer, intent(in) :: idx
        character(len=512) :: fname
        integer :: u
        write(fname, '(A, A, I0, A)') trim(dump), '_', idx, '.bin'
        open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
        write(u) n, time, eta, hu, hv, b
        close(u)
    end subroutine

    ! One output phase (SWE.f90:405-418): the VTK output traversal (or the
    ! -dump file standing for it), then the probe traversal; each flips the
    ! traversal direction.
    subroutine output(idx)
        integer, intent(in) :: idx
        integer :: ierr3
        if (len_trim(dump) > 0) call write_dump(idx)
        if (xml) call write_xml(idx)
        if (len_trim(dump) > 0 .or. xml) forward = .not. forward
        if (npt > 0) then
            call gauges_sample(n, coords, eta, hu, hv, b, npt, px, py, scaling, offset, time, pvals)
            call gauges_write(gauges_file_name(trim(outdir) // '/swe', dmax, courant, drytol, idx), &
                              npt, px, py, pvals, probe_full, ierr3)
            if (ierr3 /= 0) then
                print '(A)', 'swe_mi355x: cannot write the probe file'
                stop 1
            end if
            forward = .not. forward
        end if
    end subroutine

    ! "x1 y1,x2 y2,..." (Config.f90:605-734)
    subroutine parse_probes(str)
        character(*), intent(in) :: str
        integer :: p0, p1, k
        npt = 0
        if (len_trim(str) == 0) then
            allocate(px(0), py(0), pvals(8, 0))
            return
        end if
        npt = 1
        do k = 1, len_trim(str)
            if (str(k:k) == ',') npt = npt + 1
        end do
        allocate(px(npt), py(npt), pvals(8, npt))
        p0 = 1
        do k = 1, npt
            p1 = index(str(p0:), ',')
            if (p1 == 0) then
                p1 = len_trim(str) + 1
            else
                p1 = p0 + p1 - 1
            end if
            read(str(p0:p1 - 1), *) px(k), py(k)
            p0 = p1 + 1
        end do
    end subroutine

    subroutine write_xml(idx)
        integer, intent(in) :: idx
        character(len=512) :: piece, fname
        integer :: ierr2
        write(piece, '(A, I0, A)') 'swe_', idx, '_r0_s1.vtu'
        call vtu_write(trim(outdir) // '/' // trim(piece), n, coords, plotter, depth, eta, b, hu, hv, &
                       refine, time, scaling, offset, forward, ierr2)
        if (ierr2 /= 0) then
            print '(A, A)', 'swe_mi355x: cannot write ', trim(outdir) // '/' // trim(piece)
            stop 1
        end if
        write(fname, '(A, A, I0, A)') trim(outdir), '/swe_', idx, '.pvtu'
        call pvtu_write(trim(fname), trim(piece), ierr2)
        if (ierr2 /= 0) stop 1
    end subroutine

    subroutine swe_host_reference_dam_break_f(nc, x, e, u, v, bb)
        integer(c_int64_t), intent(in) :: nc
        real(c_double), intent(in)     :: x(2, 3, nc)
        real(c_double), intent(out)    :: e(nc), u(nc), v(nc), bb(nc)
        integer(c_int64_t) :: j
        real(c_double) :: p(2)
        do j = 1, nc
            p = 10.0_c_double * centroid(x(:, :, j)) - 5.0_c_double
            e(j) = merge(dam_left, dam_right, p(1) < 0.0_c_double)
        end do
        u = 0; v = 0; bb = 0
    end subroutine
end program swe_mi355x_main
