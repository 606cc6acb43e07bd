!> Host side of Hatch (reference src/gamera/init.F90:38-260) for the GPU step:
!> grid sizes, corners (genGridXML, init.F90:669-731, plus the warped grid of
!> the blast-wave benchmark), metric and initial condition on the GPU, then the
!> first BCs and dt (init.F90:204-244).
module gamprob_gpu
    use, intrinsic :: iso_c_binding
    use gamera_hip_c
    use gamtypes_gpu
    use gamapp_gpu
    implicit none

contains

    !> Corners of a Cartesian grid over xyzBds, optionally displaced by
    !> a*sin(2 pi xi) sin(2 pi eta) sin(2 pi zeta) (unit-scaled coordinates).
    subroutine GenCorners(Grid, xyzBds, warpAmp)
        type(Grid_T), intent(inout) :: Grid
        real(rp), intent(in) :: xyzBds(6), warpAmp
        integer :: i, j, k
        real(rp) :: dx, dy, dz, Lx, Ly, Lz, s
        dx = (xyzBds(2) - xyzBds(1))/Grid%Nip
        dy = (xyzBds(4) - xyzBds(3))/Grid%Njp
        dz = (xyzBds(6) - xyzBds(5))/Grid%Nkp
        Lx = xyzBds(2) - xyzBds(1); Ly = xyzBds(4) - xyzBds(3); Lz = xyzBds(6) - xyzBds(5)
        do k = Grid%ksg, Grid%keg + 1
            do j = Grid%jsg, Grid%jeg + 1
                do i = Grid%isg, Grid%ieg + 1
                    Grid%xyz(i, j, k, 1) = xyzBds(1) + (i - 1)*dx
                    Grid%xyz(i, j, k, 2) = xyzBds(3) + (j - 1)*dy
                    Grid%xyz(i, j, k, 3) = xyzBds(5) + (k - 1)*dz
                    if (warpAmp /= 0.0) then
                        s = warpAmp*sin(2*PI*((Grid%xyz(i, j, k, 1) - xyzBds(1))/Lx)) &
                                   *sin(2*PI*((Grid%xyz(i, j, k, 2) - xyzBds(3))/Ly)) &
                                   *sin(2*PI*((Grid%xyz(i, j, k, 3) - xyzBds(5))/Lz))
                        Grid%xyz(i, j, k, 1) = Grid%xyz(i, j, k, 1) + s*Lx
                        Grid%xyz(i, j, k, 2) = Grid%xyz(i, j, k, 2) + s*Ly
                        Grid%xyz(i, j, k, 3) = Grid%xyz(i, j, k, 3) + s*Lz
                    endif
                enddo
            enddo
        enddo
    end subroutine GenCorners

    !> Hatch on the GPU: problem "ot" (Orszag-Tang on [0,2pi]^2) or "bw" (warped blast wave)
    subroutine HatchGPU(App, problem, Nip, Njp, Nkp)
        type(gamApp_T), intent(inout) :: App
        character(len=*), intent(in) :: problem
        integer, intent(in) :: Nip, Njp, Nkp
        real(rp) :: bds(6), g
        type(gam_clock_t) :: c
        call SetGridSizes(App%Model, App%Grid, Nip, Njp, Nkp)
        call AllocGrid(App%Model, App%Grid)
        call AllocState(App%Model, App%Grid, App%State)
        call AllocState(App%Model, App%Grid, App%oState)
        select case (trim(problem))
        case ("ot")
            bds = [0.0_rp, 2*PI, 0.0_rp, 2*PI, 0.0_rp, Nkp*2*PI/Nip]
            call GenCorners(App%Grid, bds, 0.0_rp)
        case ("bw")
            bds = [0.0_rp, 1.0_rp, 0.0_rp, 1.0_rp, 0.0_rp, 1.0_rp]
            call GenCorners(App%Grid, bds, 0.05_rp)
        case default
            error stop "HatchGPU: unknown problem"
        end select
        call CreateGPU(App)
        call gpuCheck(App, gam_set_corners(App%gpu, App%Grid%xyz), "gam_set_corners")
        call gpuCheck(App, gam_get_grid(App%gpu, App%Grid%xyz, App%Grid%xyzcc, App%Grid%xfc, App%Grid%Face, &
                      App%Grid%Tf, App%Grid%Edge, App%Grid%Te, App%Grid%Teb, App%Grid%volume, App%Grid%di, &
                      App%Grid%dj, App%Grid%dk), "gam_get_grid")
        g = App%Model%gamma
        select case (trim(problem))
        case ("ot")
            call gpuCheck(App, gam_ic_orszag_tang(App%gpu, g*g, g, 1.0_rp, 1.0_rp), "ic_ot")
        case ("bw")
            call gpuCheck(App, gam_ic_blast_wave(App%gpu, 0.5_rp, 0.5_rp, 0.5_rp, 1.0_rp, 0.1_rp, 100.0_rp, &
                                                 1.0_rp, 0.1_rp, 10.0_rp), "ic_bw")
        end select
        ! clock: dt0 is captured by this host's consoleOutput
        call gpuCheck(App, gam_get_clock(App%gpu, c), "gam_get_clock")
        c%dt0_after = 0
        call gpuCheck(App, gam_set_clock(App%gpu, c), "gam_set_clock")
        call gpuCheck(App, gam_hatch(App%gpu), "gam_hatch")
        call PullClock(App)
        call PullState(App)
    end subroutine HatchGPU

end module gamprob_gpu
