! swe_scenarios.f90 -- initial conditions of the SWE runs.
!
!  * reference dam break: the reference's built-in non-ASAGI scenario,
!    eta = 1 for x < 0 else 0 at the cell centroid, b = 0, domain [-5,5]^2
!    (SWE_initialize.f90:349-370 and :185, SWE.f90:169-170);
!  * radial dam break on [0,1]^2: eta = 10 inside r < 0.1 around (0.5,0.5),
!    5 elsewhere, b = 0;
!  * tsunami on a 1000 km square: b = -4000 m plus a seeded Gaussian random
!    field (std 800 m, Gaussian covariance exp(-r^2/(2 l^2)), l = 50 km) with a
!    linear shelf over the last 150 km of the x = S edge rising to +200 m; eta
!    an elliptical Gaussian hump (amplitude A, semi-axes a and a/ratio) at a
!    seeded offshore position; dry cells get eta = min(b, 0), p = 0 (the
!    drying convention of SWE_euler_timestep.f90:280-283).
!
! Random numbers: MRG32k3a (P. L'Ecuyer, Operations Research 47(1), 1999)
! in 64-bit integer arithmetic without overflow, Box-Muller for normals.  The
! field is sampled by random Fourier features (Rahimi and Recht, NIPS 2007)
! evaluated on a regular grid as two matrix products, then interpolated
! bilinearly at the cell centroids.
module swe_scenarios
    use iso_c_binding
    implicit none
    private
    public :: rng_t, rng_seed, rng_uniform, rng_normal, tsunami_draw, tsunami_state, &
              grf_grid, centroid, tsunami_raster, raster_sample, raster_integrate, integrate_levels, &
              TSU_G, TSU_DX, TSU_SCALING

    integer(c_int64_t), parameter :: M1 = 4294967087_c_int64_t, M2 = 4294944443_c_int64_t
    real(c_double), parameter :: PI = 3.14159265358979323846_c_double
    real(c_double), parameter :: TSU_SCALING = 1.0e6_c_double
    integer, parameter :: TSU_G = 513
    real(c_double), parameter :: TSU_DX = TSU_SCALING / real(TSU_G - 1, c_double)

    type rng_t
        integer(c_int64_t) :: s1(3), s2(3)
    end type

contains

    subroutine rng_seed(r, seed)
        type(rng_t), intent(out)       :: r
        integer(c_int64_t), intent(in) :: seed
        integer :: i
        integer(c_int64_t) :: v
        v = mod(abs(seed), 2147483647_c_int64_t)
        r%s1 = [12345_c_int64_t + v, 23456_c_int64_t + 7 * v, 34567_c_int64_t + 13 * v]
        r%s2 = [45678_c_int64_t + 3 * v, 56789_c_int64_t + 11 * v, 67890_c_int64_t + 17 * v]
        r%s1 = mod(r%s1, M1)
        r%s2 = mod(r%s2, M2)
        do i = 1, 64
            v = int(rng_uniform(r) * 1.0e6_c_double, c_int64_t)
        end do
    end subroutine

    !> MRG32k3a, uniform on (0, 1).
    real(c_double) function rng_uniform(r)
        type(rng_t), intent(inout) :: r
        integer(c_int64_t) :: p1, p2
        p1 = mod(1403580_c_int64_t * r%s1(2) - 810728_c_int64_t * r%s1(1), M1)
        if (p1 < 0) p1 = p1 + M1
        r%s1(1) = r%s1(2); r%s1(2) = r%s1(3); r%s1(3) = p1
        p2 = mod(527612_c_int64_t * r%s2(3) - 1370589_c_int64_t * r%s2(1), M2)
        if (p2 < 0) p2 = p2 + M2
        r%s2(1) = r%s2(2); r%s2(2) = r%s2(3); r%s2(3) = p2
        if (p1 > p2) then
            rng_uniform = real(p1 - p2, c_double) / real(M1 + 1, c_double)
        else
            rng_uniform = real(p1 - p2 + M1, c_double) / real(M1 + 1, c_double)
        end if
    end function

    real(c_double) function rng_normal(r)
        type(rng_t), intent(inout) :: r
        real(c_double) :: u1, u2
        u1 = rng_uniform(r)
        u2 = rng_uniform(r)
        rng_normal = sqrt(-2.0_c_double * log(u1)) * cos(2.0_c_double * PI * u2)
    end function

    !> Centroid of each cell in unit-square coordinates: (x1 + x2 + x3) / 3.
    pure function centroid(coords) result(x)
        real(c_double), intent(in) :: coords(2, 3)
        real(c_double) :: x(2)
        x = (coords(:, 1) + coords(:, 2) + coords(:, 3)) / 3.0_c_double
    end function

    !> Gaussian random field on a g x g grid over [0, L]^2 (grid(i, j) at
    !> x = (i-1) L/(g-1), y = (j-1) L/(g-1)), zero mean, std sigma, covariance
    !> sigma^2 exp(-r^2 / (2 l^2)), from nfeat random Fourier features.
    subroutine grf_grid(r, g, L, sigma, ell, nfeat, grid)
        type(rng_t), intent(inout)  :: r
        integer, intent(in)         :: g, nfeat
        real(c_double), intent(in)  :: L, sigma, ell
        real(c_double), intent(out) :: grid(g, g)
        real(c_double), allocatable :: cx(:, :), sx(:, :), cy(:, :), sy(:, :)
        real(c_double) :: kx, ky, phi, x, amp
        integer :: m, i
        allocate(cx(nfeat, g), sx(nfeat, g), cy(nfeat, g), sy(nfeat, g))
        do m = 1, nfeat
            kx = rng_normal(r) / ell
            ky = rng_normal(r) / ell
            phi = 2.0_c_double * PI * rng_uniform(r)
            do i = 1, g
                x = real(i - 1, c_double) * L / real(g - 1, c_double)
                cx(m, i) = cos(kx * x + phi)
                sx(m, i) = sin(kx * x + phi)
                cy(m, i) = cos(ky * x)
                sy(m, i) = sin(ky * x)
            end do
        end do
        ! cos(kx x + phi + ky y) = cos(kx x + phi) cos(ky y) - sin(kx x + phi) sin(ky y)
        grid = matmul(transpose(cx), cy) - matmul(transpose(sx), sy)
        amp = sigma * sqrt(2.0_c_double / real(nfeat, c_double))
        grid = amp * grid
        deallocate(cx, sx, cy, sy)
    end subroutine

    !> Scenario parameters drawn from a seed: amplitude, axis ratio, source
    !> position (metres) and a uniform depth in [dlo, dhi].
    subroutine tsunami_draw(seed, amp, ratio, x0, y0, depth, dlo, dhi)
        integer(c_int64_t), intent(in) :: seed
        real(c_double), intent(out)    :: amp, ratio, x0, y0
        integer, intent(out)           :: depth
        integer, intent(in)            :: dlo, dhi
        type(rng_t) :: r
        call rng_seed(r, 1000003_c_int64_t * seed + 17_c_int64_t)
        amp = 0.5_c_double + 1.5_c_double * rng_uniform(r)
        ratio = 1.0_c_double + 2.0_c_double * rng_uniform(r)
        x0 = 2.0e5_c_double + 4.0e5_c_double * rng_uniform(r)
        y0 = 2.0e5_c_double + 6.0e5_c_double * rng_uniform(r)
        depth = dlo + min(dhi - dlo, int(real(dhi - dlo + 1, c_double) * rng_uniform(r)))
    end subroutine

    !> Bathymetry raster of the tsunami scenario: TSU_G x TSU_G nodes over
    !> [0, TSU_SCALING]^2 (spacing TSU_SCALING / (TSU_G - 1), x fastest):
    !> -4000 m plus the Gaussian random field, blended over the last 150 km of
    !> x towards a +200 m shelf.
    subroutine tsunami_raster(seed, grid)
        integer(c_int64_t), intent(in) :: seed
        real(c_double), intent(out)    :: grid(TSU_G, TSU_G)
        type(rng_t) :: r
        real(c_double) :: x, w, shelf0
        integer :: i
        call rng_seed(r, 7919_c_int64_t * seed + 3_c_int64_t)
        call grf_grid(r, TSU_G, TSU_SCALING, 800.0_c_double, 5.0e4_c_double, 256, grid)
        shelf0 = TSU_SCALING - 1.5e5_c_double
        do i = 1, TSU_G
            x = real(i - 1, c_double) * TSU_DX
            w = min(max((x - shelf0) / 1.5e5_c_double, 0.0_c_double), 1.0_c_double)
            grid(i, :) = (1.0_c_double - w) * (-4000.0_c_double + grid(i, :)) + w * 200.0_c_double
        end do
    end subroutine

    !> Bilinear sample of a raster of nx x ny nodes (x0 + (i-1) dx, y0 + (j-1) dy)
    !> at the world point p, clamped to the last raster cell.  The device's
    !> remesher evaluates the same expression for the cells it splits.
    pure real(c_double) function raster_sample(nx, ny, grid, x0, y0, dx, dy, p) result(v)
        integer, intent(in)        :: nx, ny
        real(c_double), intent(in) :: grid(nx, ny), x0, y0, dx, dy, p(2)
        real(c_double) :: fx, fy, wx, wy
        integer :: ix, iy
        ! outside the nodes' extent: the sea floor the reference assumes
        ! outside its data (SWE_initialize.f90:164-171)
        if (.not. (x0 <= p(1) .and. y0 <= p(2) .and. p(1) <= x0 + real(nx - 1, c_double) * dx .and. &
                   p(2) <= y0 + real(ny - 1, c_double) * dy)) then
            v = -5000.0_c_double
            return
        end if
        fx = (p(1) - x0) / dx
        fy = (p(2) - y0) / dy
        ix = min(max(int(fx), 0), nx - 2)
        iy = min(max(int(fy), 0), ny - 2)
        wx = fx - real(ix, c_double)
        wy = fy - real(iy, c_double)
        v = (1.0_c_double - wx) * (1.0_c_double - wy) * grid(ix + 1, iy + 1) &
            + wx * (1.0_c_double - wy) * grid(ix + 2, iy + 1) &
            + (1.0_c_double - wx) * wy * grid(ix + 1, iy + 2) + wx * wy * grid(ix + 2, iy + 2)
    end function

    !> get_bathymetry_at_element under _ADAPT_INTEGRATE for gridded data
    !> (SWE_initialize.f90:106-147): over `levels` halvings of the cell
    !> (x1, x2, x3) (unit square, x2 the right angle), 0.5 (value of
    !> (x1, m, x2) + value of (x2, m, x3)) with m = 0.5 (x1 + x3); at the
    !> bottom the raster at the world point scaling * (x1 + x2 + x3) / 3.
    pure recursive real(c_double) function raster_integrate(nx, ny, grid, x0, y0, dx, dy, scaling, &
                                                            x1, x2, x3, levels) result(v)
        integer, intent(in)        :: nx, ny, levels
        real(c_double), intent(in) :: grid(nx, ny), x0, y0, dx, dy, scaling, x1(2), x2(2), x3(2)
        real(c_double) :: m(2)
        if (levels > 0) then
            m = 0.5_c_double * (x1 + x3)
            v = 0.5_c_double * (raster_integrate(nx, ny, grid, x0, y0, dx, dy, scaling, x1, m, x2, &
                                                 levels - 1) &
                                + raster_integrate(nx, ny, grid, x0, y0, dx, dy, scaling, x2, m, x3, &
                                                   levels - 1))
        else
            v = raster_sample(nx, ny, grid, x0, y0, dx, dy, scaling * ((x1 + x2 + x3) / 3.0_c_double))
        end if
    end function

    !> Halvings of a cell at depth d for raster_integrate: the ASAGI branch of
    !> SWE_initialize.f90:127-133, min(16, min(dmax + 3, data_depth + 3) - d),
    !> with data_depth = nint(log(scaling**2 / (dx dy)) / log(2)).
    pure integer function integrate_levels(d, dmax, scaling, dx, dy) result(l)
        integer, intent(in)        :: d, dmax
        real(c_double), intent(in) :: scaling, dx, dy
        integer :: data_depth
        data_depth = nint(log(scaling ** 2 / (dx * dy)) / log(2.0_c_double))
        l = min(16, min(dmax + 3, data_depth + 3) - d)
    end function

    !> Tsunami state of n cells (unit-square corners, depths; dmax the mesh's
    !> maximum depth): b from the raster averaged over the cell as the
    !> reference's _ADAPT_INTEGRATE build does, eta the elliptical Gaussian
    !> source at the centroid (dry where it lies below b + drytol).
    subroutine tsunami_state(n, coords, depth, dmax, seed, amp, ratio, x0, y0, drytol, eta, hu, hv, b)
        integer(c_int64_t), intent(in) :: n, seed
        real(c_double), intent(in)     :: coords(2, 3, n), amp, ratio, x0, y0, drytol
        integer(c_int8_t), intent(in)  :: depth(n)
        integer, intent(in)            :: dmax
        real(c_double), intent(out)    :: eta(n), hu(n), hv(n), b(n)
        real(c_double), allocatable :: grid(:, :)
        real(c_double) :: x(2), a, c
        integer(c_int64_t) :: i
        allocate(grid(TSU_G, TSU_G))
        call tsunami_raster(seed, grid)
        a = 1.0e5_c_double
        c = a / ratio
        do i = 1, n
            x = TSU_SCALING * centroid(coords(:, :, i)) + 0.0_c_double
            b(i) = raster_integrate(TSU_G, TSU_G, grid, 0.0_c_double, 0.0_c_double, TSU_DX, TSU_DX, &
                                    TSU_SCALING, coords(:, 1, i), coords(:, 2, i), coords(:, 3, i), &
                                    integrate_levels(int(depth(i)), dmax, TSU_SCALING, TSU_DX, TSU_DX))
            eta(i) = amp * exp(-(((x(1) - x0) / a) ** 2 + ((x(2) - y0) / c) ** 2))
            hu(i) = 0.0_c_double
            hv(i) = 0.0_c_double
            if (eta(i) < b(i) + drytol) eta(i) = min(b(i), 0.0_c_double)
        end do
        deallocate(grid)
    end subroutine
end module swe_scenarios
