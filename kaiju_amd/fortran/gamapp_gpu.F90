!> Per-step app routines over the GPU step, in the reference's order:
!> stepGamera = UpdateStateData (AdvanceMHD, ts += 1, t += dt) -> CalcDT ->
!> EnforceBCs + bFlux2Fld (reference src/gamera/gamapp.F90:73-116); the
!> gamApp_T bindings over them (AdvanceModel, the call the Voltron coupler and
!> gamera.x make) are in the submodule gamtypessub_gpu.
module gamapp_gpu
    use, intrinsic :: iso_c_binding
    use, intrinsic :: iso_fortran_env, only: output_unit, error_unit, int64
    use gamera_hip_c
    use gamtypes_gpu
    implicit none

    integer(int64), private :: clockStart = 0, clockRate = 1
    integer, private :: lastTs = 0

contains

    !> Fatal C-ABI failures stop here (the library itself never exits).
    subroutine gpuCheck(App, rc, what)
        type(gamApp_T), intent(in) :: App
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc < 0) then
            write(error_unit, '(a,a,i0,a,a)') trim(what), ': rc=', rc, ' ', trim(gam_error_text(App%gpu))
            error stop 1
        endif
    end subroutine gpuCheck

    subroutine CreateGPU(App)
        type(gamApp_T), intent(inout) :: App
        type(gam_dims_t) :: d
        type(gam_params_t) :: p
        integer(c_int) :: rc
        d%Nip = App%Grid%Nip; d%Njp = App%Grid%Njp; d%Nkp = App%Grid%Nkp; d%nG = App%Model%nG
        p = ModelParams(App%Model)
        rc = gam_create(d, p, App%gpu)
        if (rc /= 0) then
            write(error_unit, *) 'gam_create failed, rc = ', rc
            error stop 1
        endif
        if (any(App%Grid%bcType == 6)) then   !> WindBC (apps/wind.F90) with a constant wind
            call gpuCheck(App, gam_set_wind(App%gpu, real(App%Grid%wind(1), c_double), real(App%Grid%wind(2), c_double), &
                                            real(App%Grid%wind(3:5), c_double), real(App%Grid%wind(6:8), c_double)), &
                          "gam_set_wind")
        endif
        if (any(App%Grid%bcType /= 0)) then
            call gpuCheck(App, gam_set_bcs(App%gpu, int(App%Grid%bcType, c_int)), "gam_set_bcs")
        endif
        if (App%Model%doRing) then   !> InitRings + RingGridFix (init.F90:562-567,968)
            call gpuCheck(App, gam_set_ring(App%gpu, int(App%Model%Ring%NumR, c_int), &
                                            int(App%Model%Ring%Nch, c_int)), "gam_set_ring")
        endif
    end subroutine CreateGPU

    subroutine DestroyGPU(App)
        type(gamApp_T), intent(inout) :: App
        integer(c_int) :: rc
        if (c_associated(App%gpu)) rc = gam_destroy(App%gpu)
        App%gpu = c_null_ptr
    end subroutine DestroyGPU

    !> Model/State times from the device clock
    subroutine PullClock(App)
        type(gamApp_T), intent(inout) :: App
        type(gam_clock_t) :: c
        call gpuCheck(App, gam_get_clock(App%gpu, c), "gam_get_clock")
        App%Model%t = c%t; App%Model%dt = c%dt; App%Model%dt0 = c%dt0; App%Model%ts = c%ts
        App%Model%status = c%status
        App%State%time = c%stateTime; App%oState%time = c%oStateTime
    end subroutine PullClock

    subroutine PushClock(App)
        type(gamApp_T), intent(inout) :: App
        type(gam_clock_t) :: c
        c%t = App%Model%t; c%dt = App%Model%dt; c%dt0 = App%Model%dt0; c%ts = App%Model%ts
        c%stateTime = App%State%time; c%oStateTime = App%oState%time
        c%status = 0; c%dt0_after = 0; c%ncpr = 0
        call gpuCheck(App, gam_set_clock(App%gpu, c), "gam_set_clock")
    end subroutine PushClock

    !> Host copies of State and oState (needed before coupling reads, IO, restart)
    subroutine PullState(App)
        type(gamApp_T), intent(inout) :: App
        real(rp) :: t
        call gpuCheck(App, gam_get_state(App%gpu, 0, t, App%State%Gas, App%State%magFlux, App%State%Bxyz), "get State")
        App%State%time = t
        call gpuCheck(App, gam_get_state(App%gpu, 1, t, App%oState%Gas, App%oState%magFlux, App%oState%Bxyz), "get oState")
        App%oState%time = t
    end subroutine PullState

    !> Device copies from the host arrays (after a coupling write or a restart read)
    subroutine PushState(App)
        type(gamApp_T), intent(inout) :: App
        call gpuCheck(App, gam_put_state(App%gpu, 0, App%State%time, App%State%Gas, App%State%magFlux, &
                                         App%State%Bxyz), "put State")
        call gpuCheck(App, gam_put_state(App%gpu, 1, App%oState%time, App%oState%Gas, App%oState%magFlux, &
                                         App%oState%Bxyz), "put oState")
        call PushClock(App)
    end subroutine PushState

    !> gamapp.F90:96-116
    subroutine UpdateStateData(App)
        type(gamApp_T), intent(inout) :: App
        call gpuCheck(App, gam_advance_mhd(App%gpu), "gam_advance_mhd")
        call PullClock(App)
    end subroutine UpdateStateData

    !> gamapp.F90:73-94
    subroutine stepGamera(App)
        type(gamApp_T), intent(inout) :: App
        real(c_double) :: dt
        integer(c_int) :: status, rc
        call UpdateStateData(App)
        rc = gam_calc_dt(App%gpu, dt, status)
        call gpuCheck(App, rc, "gam_calc_dt")
        App%Model%dt = dt
        App%Model%status = status
        if (status /= 0) then
            ! step.F90:111-156: the reference writes a black box and stops; this
            ! host reports and leaves the decision to its caller
            write(error_unit, '(a,i0,a,es12.4)') '<CalcDT status ', status, '> dt = ', dt
            return
        endif
        call gpuCheck(App, gam_enforce_bcs(App%gpu), "gam_enforce_bcs")
    end subroutine stepGamera

    !> output.F90:47-85: dt0 capture at the first output after ts > 0, kZCs
    subroutine consoleOutput(App)
        type(gamApp_T), intent(inout) :: App
        integer(int64) :: now
        real(rp) :: wall
        if (App%Model%ts > 0 .and. App%Model%dt0 < TINY) then
            App%Model%dt0 = App%Model%dt
            call PushClock(App)
        endif
        call system_clock(now, clockRate)
        if (clockStart > 0 .and. App%Model%ts > lastTs) then
            wall = real(now - clockStart, rp)/real(clockRate, rp)
            if (wall > 0) App%Model%kzcsMHD = (App%Model%ts - lastTs)*real(App%Grid%Nip, rp)*App%Grid%Njp*App%Grid%Nkp/wall*1.0e-3
        endif
        clockStart = now
        lastTs = App%Model%ts
    end subroutine consoleOutput

end module gamapp_gpu
