! swe_sfc_mesh.f90 -- uniform Sierpinski meshes in the reference's SFC order.
!
! The unit square is two root triangles (SFC_node_traversal.f90:90-97),
! refined by newest-vertex bisection; the first child takes nodes
! (x1, m, x2), the second (x2, m, x3) with m the hypotenuse midpoint
! (:138-143), and the plotter types follow the child table of :257,270.
! Neighbours are matched through the edge midpoints, which lie exactly on a
! lattice of spacing 2^-(ceil(d/2)+1); a direct-address table over that
! lattice pairs the two half-edges of every interior edge in O(N).
module swe_sfc_mesh
    use iso_c_binding
    implicit none
    private
    public :: sfc_uniform_mesh, swe_host_mesh_uniform, sfc_mesh_from_depths, swe_host_mesh_from_depths

    integer(c_int8_t), parameter :: plotter_child(-8:8) = &
        int([3, 2, 1, 8, 7, 6, 5, 4, 0, -6, -7, -8, -1, -2, -3, -4, -5], c_int8_t)

contains

    recursive subroutine refine(p, x, levels, next, plotter, coords)
        integer, intent(in)                :: p, levels
        real(c_double), intent(in)         :: x(2, 3)
        integer(c_int64_t), intent(inout)  :: next
        integer(c_int8_t), intent(inout)   :: plotter(:)
        real(c_double), intent(inout)      :: coords(:, :, :)
        real(c_double) :: m(2), first(2, 3), second(2, 3)

        if (levels == 0) then
            next = next + 1
            plotter(next) = int(p, c_int8_t)
            coords(:, :, next) = x
            return
        end if
        m = 0.5_c_double * (x(:, 1) + x(:, 3))
        first(:, 1) = x(:, 1); first(:, 2) = m; first(:, 3) = x(:, 2)
        second(:, 1) = x(:, 2); second(:, 2) = m; second(:, 3) = x(:, 3)
        call refine(int(plotter_child(p)), first, levels - 1, next, plotter, coords)
        call refine(-int(plotter_child(-p)), second, levels - 1, next, plotter, coords)
    end subroutine

    !> Uniform mesh of depth d: n = 2**(d+1) cells in forward SFC order.
    !> plotter(n), nbr(3, n) (0-based neighbour index, -1 on the boundary),
    !> coords(2, 3, n) node positions in the unit square.  ierr /= 0 if the
    !> depth is out of range or an edge cannot be paired.
    subroutine sfc_uniform_mesh(depth, plotter, nbr, coords, ierr)
        integer, intent(in)             :: depth
        integer(c_int8_t), intent(out)  :: plotter(:)
        integer(c_int32_t), intent(out) :: nbr(:, :)
        real(c_double), intent(out)     :: coords(:, :, :)
        integer, intent(out)            :: ierr
        real(c_double), parameter :: root1(2, 3) = reshape([0d0, 0d0, 1d0, 0d0, 1d0, 1d0], [2, 3])
        real(c_double), parameter :: root2(2, 3) = reshape([1d0, 1d0, 0d0, 1d0, 0d0, 0d0], [2, 3])
        integer(c_int32_t), allocatable :: owner(:, :)
        integer(c_int64_t) :: n, next, c, slot, other
        integer :: k, a, b, s, ix, iy, ko
        real(c_double) :: mx, my, fs

        ierr = 0
        if (depth < 0 .or. depth > 26) then
            ierr = 1
            return
        end if
        n = ishft(1_c_int64_t, depth + 1)
        if (size(plotter, kind=c_int64_t) < n .or. size(nbr, 2, kind=c_int64_t) < n .or. &
            size(coords, 3, kind=c_int64_t) < n) then
            ierr = 2
            return
        end if
        next = 0
        call refine(-2, root1, depth, next, plotter, coords)
        call refine(-6, root2, depth, next, plotter, coords)

        s = 2 ** ((depth + 1) / 2 + 1)
        fs = real(s, c_double)
        allocate(owner(0:s, 0:s))
        owner = -1
        nbr(:, 1:n) = -1
        do c = 1, n
            do k = 1, 3
                ! edge k is opposite node k
                a = mod(k, 3) + 1
                b = mod(k + 1, 3) + 1
                mx = 0.5_c_double * (coords(1, a, c) + coords(1, b, c)) * fs
                my = 0.5_c_double * (coords(2, a, c) + coords(2, b, c)) * fs
                ix = nint(mx)
                iy = nint(my)
                if (real(ix, c_double) /= mx .or. real(iy, c_double) /= my) then
                    ierr = 3
                    return
                end if
                slot = 3 * (c - 1) + (k - 1)
                if (owner(ix, iy) < 0) then
                    owner(ix, iy) = int(slot, c_int32_t)
                else
                    other = owner(ix, iy)
                    ko = int(mod(other, 3_c_int64_t)) + 1
                    nbr(k, c) = int(other / 3, c_int32_t)
                    nbr(ko, other / 3 + 1) = int(c - 1, c_int32_t)
                    owner(ix, iy) = -2
                end if
            end do
        end do
        deallocate(owner)
    end subroutine

    !> C entry: arrays sized for 2**(depth+1) cells; nbr is [3n] row-major by
    !> cell (edge fastest), coords [6n].  Returns n, or a negative error.
    integer(c_int64_t) function swe_host_mesh_uniform(depth, plotter, nbr, coords) &
            bind(c, name='swe_host_mesh_uniform')
        integer(c_int), value          :: depth
        integer(c_int8_t), intent(out) :: plotter(*)
        integer(c_int32_t), intent(out) :: nbr(3, *)
        real(c_double), intent(out)    :: coords(2, 3, *)
        integer :: ierr
        integer(c_int64_t) :: n
        if (depth < 0 .or. depth > 26) then
            swe_host_mesh_uniform = -1
            return
        end if
        n = ishft(1_c_int64_t, depth + 1)
        call sfc_uniform_mesh(int(depth), plotter(1:n), nbr(:, 1:n), coords(:, :, 1:n), ierr)
        if (ierr /= 0) then
            swe_host_mesh_uniform = -ierr
        else
            swe_host_mesh_uniform = n
        end if
    end function

    !> Plotter types and corners of an adaptive mesh from its depth sequence
    !> alone.  The leaves of the two root trees, read in curve order, fix the
    !> tree: a node whose depth is below the next leaf's depth was split.
    !> ierr = 1 if the sequence is not the leaf sequence of the two roots
    !> (a leaf shallower than its place in the tree, too few or too many cells).
    subroutine sfc_mesh_from_depths(n, depth, plotter, coords, ierr)
        integer(c_int64_t), intent(in)  :: n
        integer(c_int8_t), intent(in)   :: depth(n)
        integer(c_int8_t), intent(out)  :: plotter(n)
        real(c_double), intent(out)     :: coords(2, 3, n)
        integer, intent(out)            :: ierr
        real(c_double), parameter :: root1(2, 3) = reshape([0d0, 0d0, 1d0, 0d0, 1d0, 1d0], [2, 3])
        real(c_double), parameter :: root2(2, 3) = reshape([1d0, 1d0, 0d0, 1d0, 0d0, 0d0], [2, 3])
        ! a depth-first stack holds at most one pending sibling per level
        integer, parameter :: MAXD = 62
        real(c_double) :: sx(2, 3, 2 * MAXD + 4), x(2, 3), m(2)
        integer :: sp(2 * MAXD + 4), sd(2 * MAXD + 4), top, p, d
        integer(c_int64_t) :: next

        ierr = 0
        top = 2
        sx(:, :, 1) = root2
        sp(1) = -6
        sd(1) = 0
        sx(:, :, 2) = root1
        sp(2) = -2
        sd(2) = 0
        next = 0
        do while (top > 0)
            x = sx(:, :, top)
            p = sp(top)
            d = sd(top)
            top = top - 1
            if (next >= n) then
                ierr = 1
                return
            end if
            if (depth(next + 1) == d) then
                next = next + 1
                plotter(next) = int(p, c_int8_t)
                coords(:, :, next) = x
            else if (depth(next + 1) > d .and. d < MAXD) then
                m = 0.5_c_double * (x(:, 1) + x(:, 3))
                top = top + 1
                sx(:, 1, top) = x(:, 2); sx(:, 2, top) = m; sx(:, 3, top) = x(:, 3)
                sp(top) = -int(plotter_child(-p))
                sd(top) = d + 1
                top = top + 1
                sx(:, 1, top) = x(:, 1); sx(:, 2, top) = m; sx(:, 3, top) = x(:, 2)
                sp(top) = int(plotter_child(p))
                sd(top) = d + 1
            else
                ierr = 1
                return
            end if
        end do
        if (next /= n) ierr = 1
    end subroutine

    !> C entry of sfc_mesh_from_depths: plotter [n], coords [6n].  Returns 0,
    !> or 1 if depth[] is not a leaf sequence of the two root triangles.
    integer(c_int) function swe_host_mesh_from_depths(n, depth, plotter, coords) &
            bind(c, name='swe_host_mesh_from_depths')
        integer(c_int64_t), value       :: n
        integer(c_int8_t), intent(in)   :: depth(*)
        integer(c_int8_t), intent(out)  :: plotter(*)
        real(c_double), intent(out)     :: coords(2, 3, *)
        integer :: ierr
        if (n <= 0) then
            swe_host_mesh_from_depths = 1
            return
        end if
        call sfc_mesh_from_depths(n, depth(1:n), plotter(1:n), coords(:, :, 1:n), ierr)
        swe_host_mesh_from_depths = ierr
    end function
end module swe_sfc_mesh
