! Dumps the geometry tables exactly as the reference declares them.
!
! TEST INFRASTRUCTURE ONLY.  The array constructors below are the reference's
! parameter expressions (SFC_data_types.f90:514,524,557), compiled by the same
! compiler, so the folded constants are the ones the reference uses.
! Output: one line per index with the IEEE-754 bit pattern in hex.
program tables
    implicit none
    integer, parameter :: GRID_SR = kind(1.0d0)
    integer, parameter :: GRID_DI = selected_int_kind(16)
    integer, parameter :: MAX_DEPTH = bit_size(1_GRID_DI) - 4
    integer(kind=1) :: i
    integer :: j
    real(kind=GRID_SR), parameter, dimension(-1:MAX_DEPTH) :: r_leg_sizes = [ (sqrt(0.5_GRID_SR) ** i, i = -1, MAX_DEPTH) ]
    real(kind=GRID_SR), parameter, dimension(0:MAX_DEPTH) :: r_volumes = [ (0.5_GRID_SR ** (i + 1), i = 0, MAX_DEPTH) ]
    real(kind=GRID_SR), parameter, dimension(-1:MAX_DEPTH) :: r_scalings = [ (0.5_GRID_SR ** i, 0.5_GRID_SR ** i, i = 0, MAX_DEPTH/2) ]
    do j = -1, MAX_DEPTH
        write(*, '(A, 1X, I0, 1X, Z16.16)') 'leg', j, r_leg_sizes(j)
    end do
    do j = 0, MAX_DEPTH
        write(*, '(A, 1X, I0, 1X, Z16.16)') 'vol', j, r_volumes(j)
    end do
    do j = -1, MAX_DEPTH
        write(*, '(A, 1X, I0, 1X, Z16.16)') 'scl', j, r_scalings(j)
    end do
end program
