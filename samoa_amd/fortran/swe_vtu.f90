! swe_vtu.f90 -- VTK XML output of the SWE cell state, in the layout the
! reference's output traversal writes (SWE_xml_output.f90:154-250 through
! LIB_VTK_IO's 'binary' mode: raw appended data, big-endian, a 4-byte length
! before every array), so that the files of this build and of the reference
! can be compared byte for byte and opened by the same ParaView pipelines.
!
! One piece per file (rank 0, section 1), triangles with three points each,
! cell data: water height (Q%h, the surface elevation), bathymetry, momentum
! (3 components, z = 0), rank, section index, depth, refinement flag; field
! data: time.  The cells are written in the order of the output traversal
! (forward or backward along the curve); each cell's corners are written in
! the order of the forward mesh for positive plotter types and reversed for
! negative ones, as the reference's element transform yields them.
module swe_vtu
    use iso_c_binding
    implicit none
    private
    public :: vtu_write, pvtu_write

contains

    !> Writes `path`.  coords(2,3,n): unit-square corners of each cell in
    !> forward SFC order; world = scaling * coords + offset.
    subroutine vtu_write(path, n, coords, plotter, depth, eta, b, hu, hv, refine, time, &
                         scaling, offset, forward, ierr)
        character(*), intent(in)        :: path
        integer(c_int64_t), intent(in)  :: n
        real(c_double), intent(in)      :: coords(2, 3, n), eta(n), b(n), hu(n), hv(n)
        integer(c_int8_t), intent(in)   :: plotter(n), depth(n), refine(n)
        real(c_double), intent(in)      :: time, scaling, offset
        logical, intent(in)             :: forward
        integer, intent(out)            :: ierr
        integer(c_int64_t) :: np, off(12), c, k, j, i
        integer(c_int8_t), allocatable :: buf(:)
        integer(c_int64_t) :: pos
        character(len=:), allocatable :: head
        character(len=64) :: s_np, s_nc
        character(len=32) :: so(12)
        integer :: u

        np = 3 * n
        ! byte offsets of the arrays in the appended block (each: 4-byte length + data)
        off(1) = 0                                  ! time
        off(2) = off(1) + 4 + 8                     ! Point
        off(3) = off(2) + 4 + 24 * np               ! connectivity
        off(4) = off(3) + 4 + 4 * np                ! offsets
        off(5) = off(4) + 4 + 4 * n                 ! types
        off(6) = off(5) + 4 + n                     ! water height
        off(7) = off(6) + 4 + 8 * n                 ! bathymetry
        off(8) = off(7) + 4 + 8 * n                 ! momentum
        off(9) = off(8) + 4 + 24 * n                ! rank
        off(10) = off(9) + 4 + 4 * n                ! section index
        off(11) = off(10) + 4 + 4 * n               ! depth
        off(12) = off(11) + 4 + n                   ! refinement flag
        do i = 1, 12
            write(so(i), '(I0)') off(i)
        end do
        write(s_np, '(I0)') np
        write(s_nc, '(I0)') n

        head = '<?xml version="1.0"?>' // achar(10) // &
            '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="BigEndian">' // achar(10) // &
            '  <UnstructuredGrid>' // achar(10) // &
            '    <FieldData>' // achar(10) // &
            '      <DataArray type="Float64" Name="time" NumberOfComponents="1" format="appended" ' // &
            'NumberOfTuples="1" offset="' // trim(so(1)) // '">' // achar(10) // &
            '      </DataArray>' // achar(10) // &
            '    </FieldData>' // achar(10) // &
            '    <Piece NumberOfPoints="' // trim(s_np) // '" NumberOfCells="' // trim(s_nc) // '">' // achar(10) // &
            '      <Points>' // achar(10) // &
            '        <DataArray type="Float64" NumberOfComponents="3" Name="Point" format="appended" offset="' // &
            trim(so(2)) // '">' // achar(10) // &
            '        </DataArray>' // achar(10) // &
            '      </Points>' // achar(10) // &
            '      <Cells>' // achar(10) // &
            '        <DataArray type="Int32" Name="connectivity" format="appended" offset="' // trim(so(3)) // &
            '">' // achar(10) // &
            '        </DataArray>' // achar(10) // &
            '        <DataArray type="Int32" Name="offsets" format="appended" offset="' // trim(so(4)) // '">' // &
            achar(10) // &
            '        </DataArray>' // achar(10) // &
            '        <DataArray type="Int8" Name="types" format="appended" offset="' // trim(so(5)) // '">' // &
            achar(10) // &
            '        </DataArray>' // achar(10) // &
            '      </Cells>' // achar(10) // &
            '      <PointData>' // achar(10) // &
            '      </PointData>' // achar(10) // &
            '      <CellData>' // achar(10) // &
            cell_array('Float64', 'water height', 1, s_nc, so(6)) // &
            cell_array('Float64', 'bathymetry', 1, s_nc, so(7)) // &
            cell_array('Float64', 'momentum', 3, s_nc, so(8)) // &
            cell_array('Int32', 'rank', 1, s_nc, so(9)) // &
            cell_array('Int32', 'section index', 1, s_nc, so(10)) // &
            cell_array('Int8', 'depth', 1, s_nc, so(11)) // &
            cell_array('Int8', 'refinement flag', 1, s_nc, so(12)) // &
            '      </CellData>' // achar(10) // &
            '    </Piece>' // achar(10) // &
            '  </UnstructuredGrid>' // achar(10) // &
            '  <AppendedData encoding="raw">' // achar(10) // '_'

        allocate(buf(off(12) + 4 + n))
        pos = 0
        call put_i32(buf, pos, 8_c_int64_t)
        call put_f64(buf, pos, time)
        call put_i32(buf, pos, 24 * np)
        do j = 1, n
            c = cell(j)
            do k = 1, 3
                i = corner(c, k)
                call put_f64(buf, pos, scaling * coords(1, i, c) + offset)
                call put_f64(buf, pos, scaling * coords(2, i, c) + offset)
                call put_f64(buf, pos, 0.0_c_double)
            end do
        end do
        call put_i32(buf, pos, 4 * np)
        do j = 0, np - 1
            call put_i32(buf, pos, j)
        end do
        call put_i32(buf, pos, 4 * n)
        do j = 1, n
            call put_i32(buf, pos, 3 * j)
        end do
        call put_i32(buf, pos, n)
        buf(pos + 1:pos + n) = 5_c_int8_t  ! VTK_TRIANGLE
        pos = pos + n
        call put_i32(buf, pos, 8 * n)
        do j = 1, n
            call put_f64(buf, pos, eta(cell(j)))
        end do
        call put_i32(buf, pos, 8 * n)
        do j = 1, n
            call put_f64(buf, pos, b(cell(j)))
        end do
        call put_i32(buf, pos, 24 * n)
        do j = 1, n
            call put_f64(buf, pos, hu(cell(j)))
            call put_f64(buf, pos, hv(cell(j)))
            call put_f64(buf, pos, 0.0_c_double)
        end do
        call put_i32(buf, pos, 4 * n)
        do j = 1, n
            call put_i32(buf, pos, 0_c_int64_t)
        end do
        call put_i32(buf, pos, 4 * n)
        do j = 1, n
            call put_i32(buf, pos, 1_c_int64_t)
        end do
        call put_i32(buf, pos, n)
        do j = 1, n
            buf(pos + j) = depth(cell(j))
        end do
        pos = pos + n
        call put_i32(buf, pos, n)
        do j = 1, n
            buf(pos + j) = refine(cell(j))
        end do
        pos = pos + n

        open(newunit=u, file=path, access='stream', form='unformatted', status='replace', iostat=ierr)
        if (ierr /= 0) return
        write(u, iostat=ierr) head, buf(1:pos), achar(10) // '  </AppendedData>' // achar(10) // &
            '</VTKFile>' // achar(10)
        close(u)

    contains

        ! cell written at position j of the output traversal
        pure integer(c_int64_t) function cell(j)
            integer(c_int64_t), intent(in) :: j
            if (forward) then
                cell = j
            else
                cell = n + 1 - j
            end if
        end function

        ! k-th corner written for cell c
        pure integer(c_int64_t) function corner(c, k)
            integer(c_int64_t), intent(in) :: c, k
            if (plotter(c) > 0) then
                corner = k
            else
                corner = 4 - k
            end if
        end function
    end subroutine

    function cell_array(typ, name, ncomp, s_nc, soff) result(s)
        character(*), intent(in) :: typ, name, s_nc, soff
        integer, intent(in) :: ncomp
        character(len=:), allocatable :: s
        character(len=8) :: sc
        write(sc, '(I0)') ncomp
        s = '        <DataArray type="' // typ // '" Name="' // name // '" NumberOfComponents="' // &
            trim(sc) // '" format="appended" NumberOfTuples="' // trim(s_nc) // '" offset="' // &
            trim(soff) // '">' // achar(10) // '        </DataArray>' // achar(10)
    end function

    subroutine put_i32(buf, pos, v)
        integer(c_int8_t), intent(inout) :: buf(:)
        integer(c_int64_t), intent(inout) :: pos
        integer(c_int64_t), intent(in) :: v
        integer :: k
        do k = 1, 4
            buf(pos + k) = int(iand(ishft(v, -8 * (4 - k)), 255_c_int64_t) - &
                               merge(256_c_int64_t, 0_c_int64_t, iand(ishft(v, -8 * (4 - k)), 255_c_int64_t) > 127), &
                               c_int8_t)
        end do
        pos = pos + 4
    end subroutine

    subroutine put_f64(buf, pos, x)
        integer(c_int8_t), intent(inout) :: buf(:)
        integer(c_int64_t), intent(inout) :: pos
        real(c_double), intent(in) :: x
        integer(c_int64_t) :: v, byte
        integer :: k
        v = transfer(x, 0_c_int64_t)
        do k = 1, 8
            byte = iand(ishft(v, -8 * (8 - k)), 255_c_int64_t)
            buf(pos + k) = int(byte - merge(256_c_int64_t, 0_c_int64_t, byte > 127), c_int8_t)
        end do
        pos = pos + 8
    end subroutine

    !> The reference's parallel master file for one piece (SWE_xml_output.f90
    !> post_traversal_grid_op), naming `piece` as its source.
    subroutine pvtu_write(path, piece, ierr)
        character(*), intent(in) :: path, piece
        integer, intent(out) :: ierr
        integer :: u
        character(len=1), parameter :: nl = achar(10)
        open(newunit=u, file=path, access='stream', form='unformatted', status='replace', iostat=ierr)
        if (ierr /= 0) return
        write(u, iostat=ierr) '<?xml version="1.0"?>' // nl // &
            '<VTKFile type="PUnstructuredGrid" version="0.1" byte_order="BigEndian">' // nl // &
            '  <PUnstructuredGrid GhostLevel="0">' // nl // &
            '    <PFieldData>' // nl // &
            '      <PDataArray type="Float64" Name="time" NumberOfComponents="1" />' // nl // &
            '    </PFieldData>' // nl // &
            '    <PPointData>' // nl // &
            '    </PPointData>' // nl // &
            '    <PCellData>' // nl // &
            '      <PDataArray type="Float64" Name="water height" NumberOfComponents="1" />' // nl // &
            '      <PDataArray type="Float64" Name="bathymetry" NumberOfComponents="1" />' // nl // &
            '      <PDataArray type="Float64" Name="momentum" NumberOfComponents="3" />' // nl // &
            '      <PDataArray type="Int32" Name="rank" NumberOfComponents="1" />' // nl // &
            '      <PDataArray type="Int32" Name="section index" NumberOfComponents="1" />' // nl // &
            '      <PDataArray type="Int8" Name="depth" NumberOfComponents="1" />' // nl // &
            '      <PDataArray type="Int8" Name="refinement flag" NumberOfComponents="1" />' // nl // &
            '    </PCellData>' // nl // &
            '    <PPoints>' // nl // &
            '      <PDataArray type="Float64" Name="Point" NumberOfComponents="3" />' // nl // &
            '    </PPoints>' // nl // &
            '    <Piece Source="' // piece // '" />' // nl // &
            '  </PUnstructuredGrid>' // nl // &
            '</VTKFile>' // nl
        close(u)
    end subroutine
end module swe_vtu
