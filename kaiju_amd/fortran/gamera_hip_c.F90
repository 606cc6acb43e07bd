!> ISO_C_BINDING interfaces to libgamera_hip.so (include/gamera_hip.h).
!> Every argument is a plain C type: handles as c_ptr, arrays as the base
!> address of a contiguous column-major allocatable (assumed-size dummies).
module gamera_hip_c
    use, intrinsic :: iso_c_binding
    implicit none

    type, bind(C) :: gam_dims_t
        integer(c_int32_t) :: Nip, Njp, Nkp, nG
    end type gam_dims_t

    type, bind(C) :: gam_params_t
        real(c_double) :: gamma, CFL, pdmb, Vd0, cHogH, cHogM, Ca, dFloor, pFloor, tFin, limDT0, limCPR
        integer(c_int32_t) :: rmeth, doLFMLim, doBoris, doMHD, do25D, doCPR, fixedTimestep, precision
    end type gam_params_t

    type, bind(C) :: gam_clock_t
        real(c_double) :: t, dt, dt0, stateTime, oStateTime
        integer(c_int32_t) :: ts, status, dt0_after, ncpr
    end type gam_clock_t

    interface
        integer(c_int) function gam_create(dims, params, h) bind(C, name="gam_create")
            import :: c_int, c_ptr, gam_dims_t, gam_params_t
            type(gam_dims_t), intent(in) :: dims
            type(gam_params_t), intent(in) :: params
            type(c_ptr), intent(out) :: h
        end function gam_create

        integer(c_int) function gam_destroy(h) bind(C, name="gam_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function gam_destroy

        type(c_ptr) function gam_last_error(h) bind(C, name="gam_last_error")
            import :: c_ptr
            type(c_ptr), value :: h
        end function gam_last_error

        integer(c_int) function gam_sync(h) bind(C, name="gam_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function gam_sync

        integer(c_int) function gam_set_corners(h, xyz) bind(C, name="gam_set_corners")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: xyz(*)
        end function gam_set_corners

        integer(c_int) function gam_set_grid(h, xyz, xyzcc, xfc, face, Tf, edge, Te, Teb, volume, di, dj, dk) &
                bind(C, name="gam_set_grid")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: xyz(*), xyzcc(*), xfc(*), face(*), Tf(*), edge(*), Te(*), Teb(*)
            real(c_double), intent(in) :: volume(*), di(*), dj(*), dk(*)
        end function gam_set_grid

        integer(c_int) function gam_get_grid(h, xyz, xyzcc, xfc, face, Tf, edge, Te, Teb, volume, di, dj, dk) &
                bind(C, name="gam_get_grid")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(inout) :: xyz(*), xyzcc(*), xfc(*), face(*), Tf(*), edge(*), Te(*), Teb(*)
            real(c_double), intent(inout) :: volume(*), di(*), dj(*), dk(*)
        end function gam_get_grid

        integer(c_int) function gam_put_state(h, which, time, Gas, magFlux, Bxyz) bind(C, name="gam_put_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: which
            real(c_double), value :: time
            real(c_double), intent(in) :: Gas(*), magFlux(*), Bxyz(*)
        end function gam_put_state

        integer(c_int) function gam_get_state(h, which, time, Gas, magFlux, Bxyz) bind(C, name="gam_get_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: which
            real(c_double), intent(out) :: time
            real(c_double), intent(inout) :: Gas(*), magFlux(*), Bxyz(*)
        end function gam_get_state

        integer(c_int) function gam_set_clock(h, c) bind(C, name="gam_set_clock")
            import :: c_int, c_ptr, gam_clock_t
            type(c_ptr), value :: h
            type(gam_clock_t), intent(in) :: c
        end function gam_set_clock

        integer(c_int) function gam_get_clock(h, c) bind(C, name="gam_get_clock")
            import :: c_int, c_ptr, gam_clock_t
            type(c_ptr), value :: h
            type(gam_clock_t), intent(out) :: c
        end function gam_get_clock

        integer(c_int) function gam_ic_orszag_tang(h, rho0, p0, b0, kscale) bind(C, name="gam_ic_orszag_tang")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: rho0, p0, b0, kscale
        end function gam_ic_orszag_tang

        integer(c_int) function gam_ic_blast_wave(h, x0, y0, z0, D0, P0, pRat, dRat, rC, B0) &
                bind(C, name="gam_ic_blast_wave")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: x0, y0, z0, D0, P0, pRat, dRat, rC, B0
        end function gam_ic_blast_wave

        !> background field: Grid_T%fcB0/edgB0/B0/bFlux0/dpB0 (gamtypes.F90:231-238)
        integer(c_int) function gam_set_background(h, fcB0, edgB0, B0, bFlux0, dpB0) &
                bind(C, name="gam_set_background")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: fcB0(*), edgB0(*), B0(*), bFlux0(*), dpB0(*)
        end function gam_set_background

        !> AddB0 (background.F90:20-167) on the GPU for B0 = bu + dipole(x - xc)
        integer(c_int) function gam_background_dipole(h, M0, xc, bu) bind(C, name="gam_background_dipole")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: M0
            real(c_double), intent(in) :: xc(3), bu(3)
        end function gam_background_dipole

        !> boundary type per face (INI, OUTI, INJ, OUTJ, INK, OUTK)
        integer(c_int) function gam_set_bcs(h, types) bind(C, name="gam_set_bcs")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), intent(in) :: types(6)
        end function gam_set_bcs

        !> the constant solar wind of a WindBC outer i face (before gam_set_bcs)
        integer(c_int) function gam_set_wind(h, rho, p, v, b) bind(C, name="gam_set_wind")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: rho, p
            real(c_double), intent(in) :: v(3), b(3)
        end function gam_set_wind

        !> the solar wind as readWind's time series in code units: tW(nt), Q(:,DEN),
        !> Q(:,PRESSURE), Q(:,VELX:VELZ), B(nt,3), xyzBS(nt,3) and ByC, BzC, Bx0
        integer(c_int) function gam_set_wind_series(h, nt, t, rho, p, v, b, xyzBS, ByC, BzC, Bx0) &
                bind(C, name="gam_set_wind_series")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: nt
            real(c_double), intent(in) :: t(nt), rho(nt), p(nt), v(nt, 3), b(nt, 3), xyzBS(nt, 3)
            real(c_double), value :: ByC, BzC, Bx0
        end function gam_set_wind_series

        !> ring average about the LFM axis: nch(r) chunks in ring r (before gam_set_corners)
        integer(c_int) function gam_set_ring(h, nring, nch) bind(C, name="gam_set_ring")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: nring
            integer(c_int), intent(in) :: nch(*)
        end function gam_set_ring

        integer(c_int) function gam_hatch(h) bind(C, name="gam_hatch")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function gam_hatch

        integer(c_int) function gam_advance_mhd(h) bind(C, name="gam_advance_mhd")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function gam_advance_mhd

        integer(c_int) function gam_calc_dt(h, dt, status) bind(C, name="gam_calc_dt")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: dt
            integer(c_int), intent(out) :: status
        end function gam_calc_dt

        integer(c_int) function gam_enforce_bcs(h) bind(C, name="gam_enforce_bcs")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function gam_enforce_bcs

        integer(c_int) function gam_step(h, nsteps) bind(C, name="gam_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: nsteps
        end function gam_step

        integer(c_int) function gam_divb(h, tot) bind(C, name="gam_divb")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: tot
        end function gam_divb
    end interface

contains

    !> Text of gam_last_error() as a Fortran string.
    function gam_error_text(h) result(msg)
        type(c_ptr), intent(in) :: h
        character(len=512) :: msg
        character(kind=c_char), pointer :: cs(:)
        type(c_ptr) :: p
        integer :: n
        msg = ""
        p = gam_last_error(h)
        if (.not. c_associated(p)) return
        call c_f_pointer(p, cs, [512])
        do n = 1, 512
            if (cs(n) == c_null_char) exit
            msg(n:n) = cs(n)
        enddo
    end function gam_error_text

end module gamera_hip_c
