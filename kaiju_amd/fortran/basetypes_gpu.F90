!> The abstract app every model extends, as the reference's BaseApp_T
!> (reference src/base/types/basetypes.F90:12-29,87-97): a caller that holds a
!> class(BaseApp_T) -- the Voltron coupler does, gamCplTypessub.F90:129,146 --
!> advances any model with call App%AdvanceModel(dt).  Only the bindings this
!> host implements are deferred here: AdvanceModel, WriteConsoleOutput and
!> Cleanup (the reference's InitModel/InitIO take its XML deck, and its
!> restart and file-output bindings its HDF5 writer; this host has neither).
module basetypes_gpu
    use, intrinsic :: iso_c_binding, only: c_double
    implicit none

    integer, parameter :: rp = c_double

    type, abstract :: BaseApp_T
    contains
        procedure(AdvanceModel_interface), deferred :: AdvanceModel
        procedure(WriteConsoleOutput_interface), deferred :: WriteConsoleOutput
        procedure(Cleanup_interface), deferred :: Cleanup
    end type BaseApp_T

    abstract interface
        subroutine AdvanceModel_interface(App, dt)
            import BaseApp_T, rp
            class(BaseApp_T), intent(inout) :: App
            real(rp), intent(in) :: dt
        end subroutine AdvanceModel_interface

        subroutine WriteConsoleOutput_interface(App)
            import BaseApp_T
            class(BaseApp_T), intent(inout) :: App
        end subroutine WriteConsoleOutput_interface

        subroutine Cleanup_interface(App)
            import BaseApp_T
            class(BaseApp_T), intent(inout) :: App
        end subroutine Cleanup_interface
    end interface

end module basetypes_gpu
