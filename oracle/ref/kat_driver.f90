! Known-answer driver for the reference GeoClaw solvers.
!
! TEST INFRASTRUCTURE ONLY.  Linked by oracle/Makefile against the reference's
! own src/geoclaw/riemannsolvers.f and src/geoclaw/c_bind_riemannsolvers.f90
! (compiled from where they lie in the reference tree).  It reads input
! records from stdin-named files and writes the reference's answers:
!
!   kat_driver geoclaw <solver> <in> <out>
!       in : n records of 10 float64 (hL hR huL huR hvL hvR bL bR drytol g)
!       out: n records of 7 float64 (netL(3) netR(3) maxWaveSpeed)
!       solver: 1 = GEOCLAW_FWAVE, 3 = GEOCLAW_AUG_RIEMANN
!       (c_bind_riemannsolvers.f90:16-176, called with maxiter=1, mwaves=3
!        as SWE_euler_timestep.f90:469-471 does)
!   kat_driver riemanntype <in> <out>
!       in : n records of 6 float64 (hL hR uL uR drytol g)
!       out: n records of 5 float64 (hm s1m s2m rare1 rare2)
!       (riemannsolvers.f:507-618 with maxiter=1)
program kat_driver
    use c_bind_riemannsolvers
    implicit none
    character(len=512) :: mode, fin, fout, ssolver
    integer :: solver, n, i, u_in, u_out, ios
    integer(kind=8) :: fsize
    double precision :: rec10(10), rec6(6), o7(7), o5(5)
    double precision :: hL, hR, huL, huR, hvL, hvR, bL, bR
    double precision :: netL(3), netR(3), smax, hm, s1m, s2m
    logical :: rare1, rare2

    call get_command_argument(1, mode)
    if (trim(mode) == 'geoclaw') then
        call get_command_argument(2, ssolver)
        call get_command_argument(3, fin)
        call get_command_argument(4, fout)
        read(ssolver, *) solver
        open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
        inquire(unit=u_in, size=fsize)
        n = int(fsize / 80)
        open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
        do i = 1, n
            read(u_in, iostat=ios) rec10
            if (ios /= 0) exit
            hL = rec10(1); hR = rec10(2); huL = rec10(3); huR = rec10(4)
            hvL = rec10(5); hvR = rec10(6); bL = rec10(7); bR = rec10(8)
            call c_bind_geoclaw_solver(solver, 1, 3, hL, hR, huL, huR, hvL, hvR, bL, bR, &
                                       rec10(9), rec10(10), netL, netR, smax)
            o7(1:3) = netL
            o7(4:6) = netR
            o7(7) = smax
            write(u_out) o7
        end do
        close(u_in)
        close(u_out)
    else if (trim(mode) == 'riemanntype') then
        call get_command_argument(2, fin)
        call get_command_argument(3, fout)
        open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
        inquire(unit=u_in, size=fsize)
        n = int(fsize / 48)
        open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
        do i = 1, n
            read(u_in, iostat=ios) rec6
            if (ios /= 0) exit
            call riemanntype(rec6(1), rec6(2), rec6(3), rec6(4), hm, s1m, s2m, rare1, rare2, 1, rec6(5), rec6(6))
            o5(1) = hm
            o5(2) = s1m
            o5(3) = s2m
            o5(4) = merge(1.0d0, 0.0d0, rare1)
            o5(5) = merge(1.0d0, 0.0d0, rare2)
            write(u_out) o5
        end do
        close(u_in)
        close(u_out)
    else
        print *, 'usage: kat_driver geoclaw <solver> <in> <out> | riemanntype <in> <out>'
        stop 1
    end if
end program
