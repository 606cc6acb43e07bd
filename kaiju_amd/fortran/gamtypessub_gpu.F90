!> Bodies of gamApp_T's type-bound procedures (reference
!> src/gamera/gamtypessub.F90:78-100), over the GPU step routines.
submodule (gamtypes_gpu) gamtypessub_gpu
    use gamapp_gpu, only: stepGamera, consoleOutput, DestroyGPU
    implicit none

contains

    !> gamtypessub.F90:78-94.  One difference: the reference's CalcDT stops the
    !> program on a bad or disastrous timestep (step.F90:143-152); here the
    !> library returns the status, stepGamera leaves it in Model%status, and the
    !> loop stops at it so the caller can decide (CPR, black box, stop).
    module subroutine gamAdvanceModel(App, dt)
        class(gamApp_T), intent(inout) :: App
        real(rp), intent(in) :: dt
        real(rp) :: targetSimT
        targetSimT = App%Model%t + dt
        call stepGamera(App)
        do while (App%Model%t < targetSimT .and. App%Model%status == 0)
            call stepGamera(App)
        enddo
    end subroutine gamAdvanceModel

    module subroutine gamWriteConsoleOutput(App)
        class(gamApp_T), intent(inout) :: App
        call consoleOutput(App)
    end subroutine gamWriteConsoleOutput

    module subroutine gamCleanup(App)
        class(gamApp_T), intent(inout) :: App
        call DestroyGPU(App)
    end subroutine gamCleanup

end submodule gamtypessub_gpu
