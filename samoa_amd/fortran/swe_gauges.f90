! swe_gauges.f90 -- probe (gauge) output of the SWE state, as the reference's
! point-output traversal computes and writes it (SWE_point_output.f90:68-265).
!
! Each probe is given in world coordinates and mapped to the unit square,
! (x - offset) / scaling.  The cell that contains it, within a tolerance of
! 0.001 world-x units mapped into the cell's local frame, and whose centroid is
! nearest, supplies p(1), p(2), h (the surface elevation Q%h), b and the
! centroid distance.  Time is written in days since the start of the year of
! the reference's buoy comparison: t / 86400 + 70.2422.  A probe that no cell
! contains keeps distance huge(1.0) and zeros elsewhere, as in the reference.
! The reference prints 15 decimals (F0.15); `full` selects 17 significant
! digits instead, enough to read every double back exactly.
module swe_gauges
    use iso_c_binding
    implicit none
    private
    public :: gauges_sample, gauges_write, gauges_file_name

contains

    !> out(1:8, i) = x_unit, y_unit, p(1), p(2), h, b, dist, time (days) of probe i.
    subroutine gauges_sample(n, coords, eta, hu, hv, b, npt, px, py, scaling, offset, time, out)
        integer(c_int64_t), intent(in) :: n
        real(c_double), intent(in)     :: coords(2, 3, n), eta(n), hu(n), hv(n), b(n)
        integer, intent(in)            :: npt
        real(c_double), intent(in)     :: px(npt), py(npt), scaling, offset, time
        real(c_double), intent(out)    :: out(8, npt)
        integer(c_int64_t) :: c
        integer :: i
        real(c_double) :: a(2), bb(2), o(2), det, xi, et, eps, d(2), dist, rx, ry
        out = 0.0_c_double
        do i = 1, npt
            out(1, i) = (px(i) - offset) / scaling
            out(2, i) = (py(i) - offset) / scaling
            out(7, i) = huge(1.0)
        end do
        do c = 1, n
            ! local frame of the reference's element transform: corner 1 at
            ! (1,0), corner 2 at the origin, corner 3 at (0,1)
            o = coords(:, 2, c)
            a = coords(:, 1, c) - o
            bb = coords(:, 3, c) - o
            det = a(1) * bb(2) - a(2) * bb(1)
            ! |A^-1 (0.001, 0)|: the tolerance in local units
            eps = sqrt((bb(2) * 0.001_c_double / det) ** 2 + (a(2) * 0.001_c_double / det) ** 2)
            do i = 1, npt
                rx = out(1, i) - o(1)
                ry = out(2, i) - o(2)
                xi = (bb(2) * rx - bb(1) * ry) / det
                et = (a(1) * ry - a(2) * rx) / det
                if (xi + eps >= 0.0_c_double .and. et + eps >= 0.0_c_double .and. &
                    xi + et - eps <= 1.0_c_double) then
                    d = (o + a / 3.0_c_double + bb / 3.0_c_double) - [out(1, i), out(2, i)]
                    dist = sqrt(dot_product(d, d))
                    if (dist < out(7, i)) then
                        out(3, i) = hu(c)
                        out(4, i) = hv(c)
                        out(5, i) = eta(c)
                        out(6, i) = b(c)
                        out(7, i) = dist
                        out(8, i) = time / (24.0_c_double * 60.0_c_double * 60.0_c_double) + 70.2422_c_double
                    end if
                end if
            end do
        end do
    end subroutine

    !> The reference's file name: <stamp>_d<dmax>_cou<courant>_dry<drytol>_pointoutput_<i>.txt
    function gauges_file_name(stamp, dmax, courant, drytol, idx) result(name)
        character(*), intent(in) :: stamp
        integer, intent(in) :: dmax, idx
        real(c_double), intent(in) :: courant, drytol
        character(len=:), allocatable :: name
        character(len=1024) :: buf
        write(buf, "(A, A, I0, A, F10.3, A, F10.3, A, I0, A)") trim(stamp), "_d", dmax, "_cou", courant, &
            "_dry", drytol, "_pointoutput_", idx, ".txt"
        name = trim(buf)
    end function

    subroutine gauges_write(path, npt, px, py, out, full, ierr)
        character(*), intent(in) :: path
        integer, intent(in) :: npt
        real(c_double), intent(in) :: px(npt), py(npt), out(8, npt)
        logical, intent(in) :: full
        integer, intent(out) :: ierr
        integer :: u, i
        open(newunit=u, file=path, action="write", status="replace", iostat=ierr)
        if (ierr /= 0) return
        write(u, "(A)") "x, y, z, p(1), p(2), h, b, dist_to_cell_center, time"
        do i = 1, npt
            if (full) then
                write(u, "(8(ES24.16E3,A),ES24.16E3)") px(i), ",", py(i), ",", 0.0_c_double, ",", &
                    out(3, i), ",", out(4, i), ",", out(5, i), ",", out(6, i), ",", out(7, i), ",", out(8, i)
            else
                write(u, "(8(F0.15,A),F0.15)") px(i), ",", py(i), ",", 0.0_c_double, ",", &
                    out(3, i), ",", out(4, i), ",", out(5, i), ",", out(6, i), ",", out(7, i), ",", out(8, i)
            end if
        end do
        close(u)
    end subroutine
end module swe_gauges
