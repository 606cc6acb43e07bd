!> gamera.x-style driver on the GPU step (reference src/drivers/gamerax.F90:30-62):
!>   gamerax_gpu <ot|bw> <Nip> <Njp> <Nkp> <nsteps>
!> Each loop iteration advances the model by its current dt through the
!> type-bound AdvanceModel, called on a class(BaseApp_T) dummy as the Voltron
!> coupler calls it (reference voltron/modelInterfaces/gamCplTypessub.F90:129,146),
!> then writes console output (dt0 capture, kZCs).  The last
!> lines are the run's summary: time, step count, dt, totals of the State
!> arrays over the active cells and div.B.
program gamerax_gpu
    use, intrinsic :: iso_c_binding
    use, intrinsic :: iso_fortran_env, only: output_unit
    use gamera_hip_c
    use basetypes_gpu
    use gamtypes_gpu
    use gamapp_gpu
    use gamprob_gpu
    implicit none

    type(gamApp_T) :: gApp
    character(len=32) :: arg, problem
    integer :: Nip, Njp, Nkp, nsteps, n, v
    real(c_double) :: divb
    integer(c_int) :: rc

    call get_command_argument(1, problem)
    call get_command_argument(2, arg); read(arg, *) Nip
    call get_command_argument(3, arg); read(arg, *) Njp
    call get_command_argument(4, arg); read(arg, *) Nkp
    call get_command_argument(5, arg); read(arg, *) nsteps

    gApp%Model%tFin = 1.0e30_rp
    call HatchGPU(gApp, problem, Nip, Njp, Nkp)
    write(output_unit, '(a,a,3(1x,i0),a,es24.16)') 'hatched ', trim(problem), Nip, Njp, Nkp, ' dt = ', gApp%Model%dt

    do n = 1, nsteps
        call coupledAdvance(gApp, gApp%Model%dt)
        if (gApp%Model%status /= 0) exit
        call gApp%WriteConsoleOutput()
    enddo

    call PullState(gApp)
    rc = gam_divb(gApp%gpu, divb)
    call gpuCheck(gApp, rc, "gam_divb")
    write(output_unit, '(a,es24.16)') 'RESULT t ', gApp%Model%t
    write(output_unit, '(a,i0)') 'RESULT ts ', gApp%Model%ts
    write(output_unit, '(a,es24.16)') 'RESULT dt ', gApp%Model%dt
    write(output_unit, '(a,es24.16)') 'RESULT dt0 ', gApp%Model%dt0
    associate(G => gApp%Grid)
    do v = 1, NVAR
        write(output_unit, '(a,i0,1x,es24.16)') 'RESULT gas ', v, &
            sum(gApp%State%Gas(G%is:G%ie, G%js:G%je, G%ks:G%ke, v, BLK))
    enddo
    write(output_unit, '(a,es24.16)') 'RESULT magflux ', sum(abs(gApp%State%magFlux))
    end associate
    write(output_unit, '(a,es24.16)') 'RESULT divb ', divb
    write(output_unit, '(a,es12.4)') 'RESULT kzcs ', gApp%Model%kzcsMHD
    call gApp%Cleanup()

contains

    !> a coupler's view of the model: any BaseApp_T, advanced through its binding
    subroutine coupledAdvance(App, dt)
        class(BaseApp_T), intent(inout) :: App
        real(rp), intent(in) :: dt
        call App%AdvanceModel(dt)
    end subroutine coupledAdvance

end program gamerax_gpu
