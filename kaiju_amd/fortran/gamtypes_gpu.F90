!> The Gamera app types the host keeps, restated with the reference's field
!> names and index bounds (reference src/base/types/gamtypes.F90:71-314):
!> Model_T, Grid_T, State_T, gamApp_T.  gamApp_T additionally owns the handle of
!> the GPU step; State arrays on the host are current only after a sync
!> (gamApp_T%PullState), the device copy is authoritative while stepping.
module gamtypes_gpu
    use, intrinsic :: iso_c_binding
    use gamera_hip_c
    use basetypes_gpu
    implicit none
    integer, parameter :: NVAR = 5, NDIM = 3, BLK = 0
    real(rp), parameter :: TINY = 1.0D-12, HUGE = 1.0D+12
    real(rp), parameter :: PI = 4.0D0*atan(1.0D0)
    enum, bind(C)
        enumerator :: DEN = 1, MOMX, MOMY, MOMZ, ENERGY
    endenum
    enum, bind(C)
        enumerator :: IDIR = 1, JDIR, KDIR
    endenum

    !> gamtypes.F90:30-50 Ring_T, the LFM-grid ring average ("lfm" only)
    type :: Ring_T
        integer :: NumR = 0            !> rings per pole
        integer :: Nch(32) = 0         !> chunks in ring 1..NumR
    end type Ring_T

    !> gamtypes.F90:71-146 (the step's knobs; XML/IO members left out)
    type :: Model_T
        character(len=64) :: RunID = "Sim"
        integer :: nSpc = 0, nDim = 3, nG = 4
        real(rp) :: CFL = 0.3_rp, gamma = 5.0_rp/3.0_rp
        real(rp) :: Vd0 = 0.5_rp
        real(rp) :: t = 0.0, tFin = 1.0, dt = 0.0
        real(rp) :: dt0 = 0.0
        logical :: fixedTimestep = .false.
        integer :: ts = 0
        real(rp) :: kzcsMHD = 0.0
        logical :: doMHD = .true., do25D = .false., doAB3 = .false.
        logical :: doBoris = .false., doHogs = .true.
        real(rp) :: cHogH = 0.25_rp, cHogM = 0.25_rp
        real(rp) :: Ca = HUGE
        logical :: doLFMLim = .true.
        real(rp) :: limDT0 = 1.0e-3_rp
        logical :: doCPR = .false.
        real(rp) :: limCPR = 1.0e-1_rp
        !> recon's module pdmb and gdefs' floors live here in this host
        real(rp) :: pdmb = 0.75_rp, dFloor = TINY, pFloor = TINY
        character(len=8) :: rmeth = "7UP"
        integer :: status = 0      !> last CalcDT outcome (0 ok, 1 bad, 2 disaster)
        !> FP32-bulk step (FP64 kept for magFlux/CT, the cell update and dt);
        !> not a reference knob, the reference is FP64 throughout (kdefs.F90:14)
        logical :: doFP32Bulk = .false.
        !> ring average about the LFM axis (Model%doRing, ring/gid = "lfm")
        logical :: doRing = .false.
        type(Ring_T) :: Ring
    end type Model_T

    !> gamtypes.F90:163-245 (sizes, index ranges and the geometry arrays)
    type :: Grid_T
        integer :: Nip, Njp, Nkp
        integer :: Ni, Nj, Nk
        integer :: is, ie, js, je, ks, ke
        integer :: isg, ieg, jsg, jeg, ksg, keg
        integer :: isDT, jsDT, ksDT, ieDT, jeDT, keDT
        real(rp), dimension(:,:,:,:), allocatable :: xyz, xyzcc
        real(rp), dimension(:,:,:,:,:), allocatable :: xfc, Tf, Te, Teb
        real(rp), dimension(:,:,:), allocatable :: di, dj, dk, volume
        real(rp), dimension(:,:,:,:), allocatable :: Face, Edge
        !> the reference's Grid%externalBCs, as one type per face INI, OUTI, INJ,
        !> OUTJ, INK, OUTK: 0 periodic, 1 zeroGrad, 2 zeroExten, 3 cartReflect, 4 SphIn/SphOut,
        !> 5 lfmIn/lfmOut (j faces), 6 WindBC (outer i face)
        integer :: bcType(6) = 0
        !> the WindBC's solar wind, constant in time: rho, P, V(3), B(3) (code units)
        real(rp) :: wind(8) = 0.0_rp
    end type Grid_T

    !> gamtypes.F90:248-264
    type :: State_T
        real(rp) :: time = -HUGE
        real(rp), dimension(:,:,:,:,:), allocatable :: Gas
        real(rp), dimension(:,:,:,:), allocatable :: magFlux
        real(rp), dimension(:,:,:,:), allocatable :: Efld
        real(rp), dimension(:,:,:,:), allocatable :: Bxyz
    end type State_T

    !> gamtypes.F90:294-314; the solver work arrays live on the GPU.  The
    !> bindings' bodies are in the submodule gamtypessub_gpu (they call the step
    !> routines of gamapp_gpu, which uses this module).
    type, extends(BaseApp_T) :: gamApp_T
        type(Model_T) :: Model
        type(Grid_T)  :: Grid
        type(State_T) :: State, oState, ooState
        type(c_ptr)   :: gpu = c_null_ptr
    contains
        procedure :: AdvanceModel => gamAdvanceModel
        procedure :: WriteConsoleOutput => gamWriteConsoleOutput
        procedure :: Cleanup => gamCleanup
    end type gamApp_T

    interface
        !> gamtypessub.F90:78-94: stepGamera at least once, then until t >= t0 + dt
        module subroutine gamAdvanceModel(App, dt)
            class(gamApp_T), intent(inout) :: App
            real(rp), intent(in) :: dt
        end subroutine gamAdvanceModel
        !> output.F90:47-85 consoleOutput
        module subroutine gamWriteConsoleOutput(App)
            class(gamApp_T), intent(inout) :: App
        end subroutine gamWriteConsoleOutput
        !> releases the GPU step (gamtypessub.F90 gamCleanup has nothing to free)
        module subroutine gamCleanup(App)
            class(gamApp_T), intent(inout) :: App
        end subroutine gamCleanup
    end interface

contains

    !> init.F90:648-663 index ranges and SetDomain (init.F90:519-538)
    subroutine SetGridSizes(Model, Grid, Nip, Njp, Nkp)
        type(Model_T), intent(inout) :: Model
        type(Grid_T), intent(inout) :: Grid
        integer, intent(in) :: Nip, Njp, Nkp
        Grid%Nip = Nip; Grid%Njp = Njp; Grid%Nkp = Nkp
        Model%do25D = (Nkp == 1)
        Grid%Ni = Nip + 2*Model%nG; Grid%Nj = Njp + 2*Model%nG; Grid%Nk = Nkp + 2*Model%nG
        Grid%is = 1; Grid%ie = Nip; Grid%js = 1; Grid%je = Njp; Grid%ks = 1; Grid%ke = Nkp
        Grid%isg = Grid%is - Model%nG; Grid%ieg = Grid%ie + Model%nG
        Grid%jsg = Grid%js - Model%nG; Grid%jeg = Grid%je + Model%nG
        Grid%ksg = Grid%ks - Model%nG; Grid%keg = Grid%ke + Model%nG
        Grid%isDT = Grid%is - 1; Grid%ieDT = Grid%ie + 1
        Grid%jsDT = Grid%js - 1; Grid%jeDT = Grid%je + 1
        Grid%ksDT = Grid%ks - 1; Grid%keDT = Grid%ke + 1
    end subroutine SetGridSizes

    !> gamutils.F90:169-201 allocState / allocGrid with the reference's bounds
    subroutine AllocState(Model, Grid, State)
        type(Model_T), intent(in) :: Model
        type(Grid_T), intent(in) :: Grid
        type(State_T), intent(inout) :: State
        associate(isg=>Grid%isg, ieg=>Grid%ieg, jsg=>Grid%jsg, jeg=>Grid%jeg, ksg=>Grid%ksg, keg=>Grid%keg)
        if (.not. allocated(State%Gas)) allocate(State%Gas(isg:ieg, jsg:jeg, ksg:keg, 1:NVAR, 0:Model%nSpc))
        if (.not. allocated(State%magFlux)) allocate(State%magFlux(isg:ieg+1, jsg:jeg+1, ksg:keg+1, 1:NDIM))
        if (.not. allocated(State%Efld)) allocate(State%Efld(isg:ieg, jsg:jeg, ksg:keg, 1:NDIM))
        if (.not. allocated(State%Bxyz)) allocate(State%Bxyz(isg:ieg, jsg:jeg, ksg:keg, 1:NDIM))
        end associate
        State%Gas = 0.0; State%magFlux = 0.0; State%Efld = 0.0; State%Bxyz = 0.0
        State%time = Model%t
    end subroutine AllocState

    subroutine AllocGrid(Model, Grid)
        type(Model_T), intent(in) :: Model
        type(Grid_T), intent(inout) :: Grid
        associate(isg=>Grid%isg, ieg=>Grid%ieg, jsg=>Grid%jsg, jeg=>Grid%jeg, ksg=>Grid%ksg, keg=>Grid%keg)
        allocate(Grid%xyz(isg:ieg+1, jsg:jeg+1, ksg:keg+1, NDIM))
        allocate(Grid%xyzcc(isg:ieg, jsg:jeg, ksg:keg, NDIM))
        allocate(Grid%xfc(isg:ieg+1, jsg:jeg+1, ksg:keg+1, NDIM, NDIM))
        allocate(Grid%Tf(isg:ieg+1, jsg:jeg+1, ksg:keg+1, NDIM*NDIM, NDIM))
        allocate(Grid%Te(isg:ieg+1, jsg:jeg+1, ksg:keg+1, NDIM*NDIM, NDIM))
        allocate(Grid%Teb(isg:ieg+1, jsg:jeg+1, ksg:keg+1, 4, NDIM))
        allocate(Grid%Face(isg:ieg+1, jsg:jeg+1, ksg:keg+1, NDIM))
        allocate(Grid%Edge(isg:ieg+1, jsg:jeg+1, ksg:keg+1, NDIM))
        allocate(Grid%volume(isg:ieg, jsg:jeg, ksg:keg))
        allocate(Grid%di(isg:ieg, jsg:jeg, ksg:keg))
        allocate(Grid%dj(isg:ieg, jsg:jeg, ksg:keg))
        allocate(Grid%dk(isg:ieg, jsg:jeg, ksg:keg))
        end associate
        Grid%xyz = 0.0
    end subroutine AllocGrid

    !> Model_T -> the C-ABI parameter block
    function ModelParams(Model) result(p)
        type(Model_T), intent(in) :: Model
        type(gam_params_t) :: p
        p%gamma = Model%gamma; p%CFL = Model%CFL; p%pdmb = Model%pdmb; p%Vd0 = Model%Vd0
        p%cHogH = merge(Model%cHogH, 0.0_rp, Model%doHogs)
        p%cHogM = merge(Model%cHogM, 0.0_rp, Model%doHogs)
        p%Ca = Model%Ca; p%dFloor = Model%dFloor; p%pFloor = Model%pFloor; p%tFin = Model%tFin
        p%limDT0 = Model%limDT0; p%limCPR = Model%limCPR
        p%rmeth = merge(1, 0, trim(Model%rmeth) == "7UP")
        p%doLFMLim = merge(1, 0, Model%doLFMLim)
        p%doBoris = merge(1, 0, Model%doBoris)
        p%doMHD = merge(1, 0, Model%doMHD)
        p%do25D = merge(1, 0, Model%do25D)
        p%doCPR = merge(1, 0, Model%doCPR)
        p%fixedTimestep = merge(1, 0, Model%fixedTimestep)
        p%precision = merge(1, 0, Model%doFP32Bulk)
    end function ModelParams

end module gamtypes_gpu
