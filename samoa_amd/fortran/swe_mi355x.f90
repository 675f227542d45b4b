! swe_mi355x.f90 -- ISO_C_BINDING interface to the HIP SWE library.
!
! The thin layer through which a Fortran host (the reference's own language)
! calls the MI355X Euler step.  Each interface mirrors one declaration of
! include/swe_mi355x.h; `swe_step` replaces `swe%euler%traverse(grid)`
! (SWE.f90:393).
module swe_mi355x
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: SWE_OK = 0
    integer(c_int), parameter :: SWE_ERR_ARG = -1, SWE_ERR_HIP = -2, SWE_ERR_NOMEM = -3, SWE_ERR_STATE = -4
    integer(c_int32_t), parameter :: SWE_SOLVER_AUG_RIEMANN = 0, SWE_SOLVER_FWAVE = 1
    integer(c_int32_t), parameter :: SWE_SOLVER_LF = 2, SWE_SOLVER_LLF = 3, SWE_SOLVER_LF_BATH = 4, &
                                     SWE_SOLVER_LLF_BATH = 5
    integer(c_int32_t), parameter :: SWE_BC_OUTFLOW = 0, SWE_BC_WALL = 1

    type, bind(c) :: swe_mesh_desc
        integer(c_int64_t) :: n_cells
        type(c_ptr)        :: depth
        type(c_ptr)        :: plotter
        type(c_ptr)        :: nbr
        real(c_double)     :: scaling
        real(c_double)     :: g
        real(c_double)     :: dry_tolerance
        real(c_double)     :: courant
        integer(c_int32_t) :: min_depth
        integer(c_int32_t) :: max_depth
        integer(c_int32_t) :: solver
        integer(c_int32_t) :: boundary
    end type

    interface
        integer(c_int) function swe_create(desc, handle) bind(c, name='swe_create')
            import :: c_int, c_ptr, swe_mesh_desc
            type(swe_mesh_desc), intent(in) :: desc
            type(c_ptr), intent(out)        :: handle
        end function

        integer(c_int) function swe_upload(handle, eta, hu, hv, b) bind(c, name='swe_upload')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value        :: handle
            real(c_double), intent(in) :: eta(*), hu(*), hv(*), b(*)
        end function

        integer(c_int) function swe_download(handle, eta, hu, hv, b) bind(c, name='swe_download')
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
            type(c_ptr), value :: eta, hu, hv, b
        end function

        !> swe%adaption%traverse (SWE.f90:387-390); refine = c_null_ptr takes
        !> the last step's flags left on the device (swe_keep_refinement_flags)
        integer(c_int) function swe_adapt(handle, refine, n_cells, sweeps) bind(c, name='swe_adapt')
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value        :: handle
            type(c_ptr), value        :: refine
            integer(c_int64_t)        :: n_cells
            integer(c_int)            :: sweeps
        end function

        integer(c_int) function swe_keep_refinement_flags(handle, on) bind(c, name='swe_keep_refinement_flags')
            import :: c_int, c_ptr
            type(c_ptr), value      :: handle
            integer(c_int), value   :: on
        end function

        !> every step leaves flags from the elevation jumps of the state it reads
        integer(c_int) function swe_keep_eta_jump_flags(handle, refine_tol, coarsen_tol) &
                bind(c, name='swe_keep_eta_jump_flags')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: handle
            real(c_double), value   :: refine_tol, coarsen_tol
        end function

        integer(c_int64_t) function swe_cells(handle) bind(c, name='swe_cells')
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: handle
        end function

        integer(c_int) function swe_download_mesh(handle, depth, plotter, nbr) bind(c, name='swe_download_mesh')
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
            type(c_ptr), value :: depth, plotter, nbr
        end function

        integer(c_int) function swe_step(handle, dt, forward, dt_cfl, n_refine, refine) bind(c, name='swe_step')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value          :: handle
            real(c_double), value       :: dt
            integer(c_int), value       :: forward
            real(c_double), intent(out) :: dt_cfl
            integer(c_int64_t), intent(out) :: n_refine
            type(c_ptr), value          :: refine
        end function

        !> time, dt_next: c_loc of the values, or c_null_ptr (both: do not wait);
        !> dt0 <= 0 continues the device dt chain
        integer(c_int) function swe_run(handle, n_steps, dt0, forward, dt_clamp, time, dt_next) bind(c, name='swe_run')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value          :: handle
            integer(c_int64_t), value   :: n_steps
            real(c_double), value       :: dt0
            integer(c_int), value       :: forward
            real(c_double), value       :: dt_clamp
            type(c_ptr), value          :: time, dt_next
        end function

        !> refinement flags from surface-elevation jumps, for swe_adapt(h, c_null_ptr, ...)
        integer(c_int) function swe_flag_eta_jumps(handle, refine_tol, coarsen_tol) bind(c, name='swe_flag_eta_jumps')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: handle
            real(c_double), value   :: refine_tol, coarsen_tol
        end function

        !> bathymetry of split cells: bilinear in values(nx, ny) at world (x0 + i dx, y0 + j dy)
        integer(c_int) function swe_set_bathymetry_raster(handle, nx, ny, x0, y0, dx, dy, offset_x, offset_y, &
                                                          values) bind(c, name='swe_set_bathymetry_raster')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value          :: handle
            integer(c_int64_t), value   :: nx, ny
            real(c_double), value       :: x0, y0, dx, dy, offset_x, offset_y
            type(c_ptr), value          :: values
        end function

        integer(c_int) function swe_destroy(handle) bind(c, name='swe_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function

        integer(c_size_t) function swe_last_error(buf, size) bind(c, name='swe_last_error')
            import :: c_size_t, c_char
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_size_t), value            :: size
        end function
    end interface

contains

    !> The library's last error message as a Fortran string.
    function swe_error_message() result(msg)
        character(len=512) :: msg
        character(kind=c_char) :: buf(512)
        integer(c_size_t) :: n
        integer :: i
        n = swe_last_error(buf, int(size(buf), c_size_t))
        msg = ' '
        do i = 1, min(int(n), 511)
            if (buf(i) == c_null_char) exit
            msg(i:i) = buf(i)
        end do
    end function
end module swe_mi355x
