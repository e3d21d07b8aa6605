!> randomField_gpu -- Fortran calling surface of the gfx950 spectral random-field
!> path (librf_hip.so), for hosts written against randomField.
!>
!> Replace `use calls_RF` (or `use randomFieldND`) by `use randomField_gpu`:
!>   * type(RF) / type(MESH) carry the components the generation path reads
!>     and writes (type_RF.f90:11-55, type_MESH.f90:9-54);
!>   * createRandomField(RDF, MSH) is the generic of create_RF_Unstruct_Init
!>     (calls_RF.f90:20-22, randomFieldND.f90:27-132);
!>   * createStandardGaussianFieldUnstruct is the README's front door
!>     (README.md:58-88), which the reference documents but does not define;
!>   * multiVariateTransformation (randomFieldND.f90:139-178),
!>     calculate_random_seed / init_random_seed (math_RF.f90:475-566);
!>   * init_RF / init_MESH with the reference's argument lists (type(IPT_RF),
!>     type_RF.f90:62, type_MESH.f90:61) or the short ones of this module;
!>   * the localization step: round_basic_inputs, set_global_extremes
!>     (topography_RF.f90:19-144), prepare_Localization
!>     (localization_RF.f90:127-192), generateUnityPartition_Matrix
!>     (localization_RF.f90:705-857), localization_blend (add_RF_to_Group of
!>     every field plus the boundary division, localization_RF.f90:199-263,
!>     calls_RF.f90:209-369, in one GPU launch) and normalize_randField
!>     (randomFieldND.f90:820-926).
!>
!> The uniforms are drawn here, on the host, with the compiler's random_number
!> after random_seed(PUT=RDF%seed), in the reference's order (wavevector
!> draws of method 3 first, spectra_RF.f90:120; then per realization:
!> randomFieldND.f90:278 for methods 2/3, :424-426 / :460-463 for method 1),
!> so a given seed gives the reference's phases with the same compiler.
!> The sums, the marginal map and everything O(N) run on the GPU.
!> Failures stop with the library's message, as the reference stops.
module randomField_gpu
    use iso_c_binding
    implicit none
    private

    ! constants_RF.f90:5-21
    double precision, parameter, public :: PI = 3.1415926535898d0
    integer, parameter, public :: ISOTROPIC = 1, SHINOZUKA = 2, RANDOMIZATION = 3, FFT = 4
    integer, parameter, public :: cm_GAUSSIAN = 1, cm_COS = 2
    integer, parameter, public :: fom_GAUSSIAN = 1, fom_LOGNORMAL = 2
    ! GPU path controls (no counterpart in the reference)
    integer, parameter, public :: RF_REF_COMPAT = 0, RF_FIXED = 1
    integer, parameter, public :: RF_FORM_FP64 = 0, RF_FORM_FP32 = 1, &
                                  RF_FORM_MFMA_F16 = 2, RF_FORM_MFMA_BF16 = 3
    integer, parameter, public :: RF_KSAMP_BOX = 0, RF_KSAMP_DENSITY = 1

    type, bind(C), public :: rf_params_c
        integer(c_int32_t) :: nDim = 0, method = 0, corrMod = 0, mode = 0
        integer(c_int32_t) :: form = 0, kSampling = 0, Nmc = 0, margiFirst = 0
        integer(c_int64_t) :: nKDensity = 0, xSlab = 0
        real(c_double)     :: corrL(3) = 1d0, xRange(3) = 1d0
        real(c_double)     :: fieldAvg = 0d0, fieldVar = 1d0
        integer(c_int64_t) :: seed = 0
        integer(c_int32_t) :: isoShared = 0
        integer(c_int32_t) :: gridN(3) = 0
        integer(c_int32_t) :: reserved(4) = 0
    end type rf_params_c

    type, bind(C), public :: rf_spectrum_info_c
        integer(c_int64_t) :: kNTotal, kN_unsigned, nKUniform, nPhaseUniform
        integer(c_int32_t) :: kNStep(3), nSignBlocks, nVarIso, pad
        real(c_double)     :: kMax(3), kDelta(3), ampMult, isoStep
    end type rf_spectrum_info_c

    !> type_RF.f90:11-55 (generation components)
    type, public :: RF
        ! MPI variables of the reference (kept for hosts that set them; one GPU)
        integer :: comm = -1, rang = -1, nb_procs = -1, log_ID = -1
        logical :: validProc = .true.
        integer :: nDim = -1
        integer(kind=8) :: kNTotal = -1
        integer, dimension(:), allocatable :: seed
        integer :: seedStart = -1
        double precision :: fieldAvg = -1, fieldVar = -1
        integer :: corrMod = -1
        integer :: margiFirst = -1
        integer :: method = -1
        integer :: Nmc = -1
        logical :: init = .false.
        double precision, dimension(:), allocatable :: corrL, kMax, kDelta
        double precision, dimension(:, :), allocatable :: kPoints
        double precision, dimension(:), allocatable :: SkVec
        double precision, dimension(:), allocatable :: xRange
        integer, dimension(:), allocatable :: kNStep
        double precision, pointer :: xPoints(:,:) => null()
        double precision, pointer :: randField(:,:) => null()
        double precision, dimension(:,:), pointer :: RF_2D => null()
        double precision, dimension(:,:,:), pointer :: RF_3D => null()
        logical, dimension(:), allocatable :: calculate
        double precision, dimension(:, :), allocatable :: xPoints_Local, randField_Local
        integer, dimension(:), allocatable :: kExtent
        integer, dimension(:, :), allocatable :: neighSeed
        double precision, dimension(:, :), allocatable :: neighRange
        ! GPU path controls
        integer :: compatMode = RF_REF_COMPAT   !< RF_REF_COMPAT or RF_FIXED
        integer :: form = RF_FORM_FP64          !< precision form of the sum
        integer :: kSampling = RF_KSAMP_BOX     !< method 3 wavevector sampling
        integer(kind=8) :: nKDensity = 0        !< N_k for RF_KSAMP_DENSITY
        logical :: isoShared = .false.          !< method 1: shared directions
        logical :: deviceDraws = .false.        !< device generator, not random_number
        integer(kind=8) :: deviceSeed = 0
        integer :: device = 0
    end type RF

    !> type_MESH.f90:9-54 (the components gen_Std_Gauss reads and rescales)
    type, public :: MESH
        integer :: comm = -1, rang = -1, nb_procs = -1, log_ID = -1
        logical :: validProc = .true.
        integer :: meshMod = -1, method = -1
        integer :: nDim = -1
        integer(kind=8) :: xNTotal = -1
        integer, dimension(:), allocatable :: xNStep, procPerDim, coords, neigh, op_neigh
        double precision, dimension(:), allocatable :: xStep, xMinGlob, xMaxGlob, corrL
        double precision, dimension(:), allocatable :: procStart, procExtent
        integer, dimension(:), allocatable :: pointsPerCorrL, origin
        double precision, dimension(:), allocatable :: xMinInt, xMaxInt, xMinExt, xMaxExt
        double precision, dimension(:), allocatable :: xMinBound, xMaxBound
        double precision, dimension(:), allocatable :: overlap   !< in correlation lengths
        integer, dimension(:, :), allocatable :: neighShift
        logical :: init = .false.
    end type MESH

    !> type_inputRF.f90:10-70: the input-record components that init_RF,
    !> init_MESH and the localization step read
    type, public :: IPT_RF
        integer :: log_ID = -1, comm = -1, rang = -1, nb_procs = -1
        logical :: init = .false.
        integer :: nDim_mesh = -1, meshMod = -1
        double precision, dimension(:), allocatable :: xMaxGlob, xMinGlob
        integer, dimension(:), allocatable :: pointsPerCorrL, procPerDim, coords
        integer :: nDim_gen = -1
        double precision :: fieldAvg = -1, fieldVar = -1
        integer :: corrMod = -1, margiFirst = -1, method = -1, Nmc = -1, seedStart = 0
        integer, dimension(:), allocatable :: nFields
        integer :: localizationLevel = 1, nTotalFields = -1
        logical :: global = .false., extLoc = .false.
        integer :: nDim = -1
        double precision, dimension(:), allocatable :: corrL, overlap, xStep, stepProc, procExtent
        integer :: loc_comm = -1
    end type IPT_RF

    type, bind(C) :: rf_layout_params_c
        integer(c_int32_t) :: nDim = 0, nFields(3) = 1, pointsPerCorrL(3) = 1, pad = 0
        real(c_double) :: xMinGlob(3) = 0d0, xMaxGlob(3) = 0d0, corrL(3) = 1d0, overlap(3) = 0d0
    end type rf_layout_params_c

    type, bind(C) :: rf_layout_info_c
        integer(c_int32_t) :: nDim, nSub, groupN(3), subN(3), nOvl(3), nOvlRaw(3)
        real(c_double) :: xStep(3), overlap(3), xMinGlob(3), xMaxGlob(3), procExtent(3), stepProc(3), &
                          xRangeSub(3)
    end type rf_layout_info_c

    interface init_RF
        module procedure init_RF_ipt, init_RF_args
    end interface init_RF

    interface init_MESH
        module procedure init_MESH_ipt, init_MESH_args
    end interface init_MESH

    interface createRandomField
        module procedure create_RF_Unstruct_Init
    end interface createRandomField

    interface createStandardGaussianFieldUnstruct
        module procedure createStdGaussUnstruct_str, createStdGaussUnstruct_int
    end interface createStandardGaussianFieldUnstruct

    public :: createRandomField, create_RF_Unstruct_Init, createStandardGaussianFieldUnstruct
    public :: multiVariateTransformation, calculate_random_seed, init_random_seed
    public :: init_RF, finalize_RF, init_MESH, finalize_MESH, rf_draw_sequence
    public :: round_basic_inputs, set_global_extremes, prepare_Localization
    public :: generateUnityPartition_Matrix, localization_blend, normalize_randField
    public :: rf_c_params, rf_check, rf_ctx

    interface
        integer(c_int) function c_rf_context_create(ctx, device) bind(C, name="rf_context_create")
            import :: c_int, c_ptr
            type(c_ptr) :: ctx
            integer(c_int), value :: device
        end function
        function c_rf_last_error(ctx) bind(C, name="rf_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
            type(c_ptr) :: c_rf_last_error
        end function
        integer(c_int) function c_rf_plan_spectrum(p, info) bind(C, name="rf_plan_spectrum")
            import :: c_int, rf_params_c, rf_spectrum_info_c
            type(rf_params_c), intent(in) :: p
            type(rf_spectrum_info_c), intent(out) :: info
        end function
        integer(c_int) function c_rf_spectrum_tables(p, kU, kPts, Sk) bind(C, name="rf_spectrum_tables")
            import :: c_int, c_ptr, rf_params_c
            type(rf_params_c), intent(in) :: p
            type(c_ptr), value :: kU, kPts, Sk
        end function
        integer(c_int) function c_rf_generate(ctx, p, x, N, kU, phU, calc, out, dtype) &
                bind(C, name="rf_generate")
            import :: c_int, c_int32_t, c_int64_t, c_ptr, rf_params_c
            type(c_ptr), value :: ctx
            type(rf_params_c), intent(in) :: p
            type(c_ptr), value :: x, kU, phU, calc, out
            integer(c_int64_t), value :: N
            integer(c_int32_t), value :: dtype
        end function
        integer(c_int) function c_rf_generate_grid(ctx, p, origin, step, nstep, phU, calc, out, dtype) &
                bind(C, name="rf_generate_grid")
            import :: c_int, c_int32_t, c_double, c_ptr, rf_params_c
            type(c_ptr), value :: ctx
            type(rf_params_c), intent(in) :: p
            real(c_double), intent(in) :: origin(*), step(*)
            integer(c_int32_t), intent(in) :: nstep(*)
            type(c_ptr), value :: phU, calc, out
            integer(c_int32_t), value :: dtype
        end function
        integer(c_int) function c_rf_localization_layout(p, info, origins, sizes, masks, xMinSub) &
                bind(C, name="rf_localization_layout")
            import :: c_int, c_ptr, rf_layout_params_c, rf_layout_info_c
            type(rf_layout_params_c), intent(in) :: p
            type(rf_layout_info_c), intent(out) :: info
            type(c_ptr), value :: origins, sizes, masks, xMinSub
        end function
        integer(c_int) function c_rf_unity_partition(nDim, xNStep, nOvl, faceMask, mode, pu) &
                bind(C, name="rf_unity_partition")
            import :: c_int, c_int32_t, c_ptr
            integer(c_int32_t), value :: nDim, faceMask, mode
            integer(c_int32_t), intent(in) :: xNStep(*), nOvl(*)
            type(c_ptr), value :: pu
        end function
        integer(c_int) function c_rf_blend_host(ctx, nDim, groupN, Nmc, nSub, origins, sizes, nOvl, masks, &
                                                mode, fields, group) bind(C, name="rf_blend_host")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nDim, Nmc, nSub, mode
            integer(c_int32_t), intent(in) :: groupN(*), origins(*), sizes(*), nOvl(*), masks(*)
            type(c_ptr), intent(in) :: fields(*)
            type(c_ptr), value :: group
        end function
        integer(c_int) function c_rf_normalize_window_host(ctx, field, nDim, gridN, minPos, maxPos, Nmc) &
                bind(C, name="rf_normalize_window_host")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx, field
            integer(c_int32_t), value :: nDim, Nmc
            integer(c_int32_t), intent(in) :: gridN(*), minPos(*), maxPos(*)
        end function
        integer(c_int) function c_rf_marginal_host(ctx, margi, avg, var, field, count) &
                bind(C, name="rf_marginal_host")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: margi
            real(c_double), value :: avg, var
            type(c_ptr), value :: field
            integer(c_int64_t), value :: count
        end function
    end interface

    type(c_ptr), save :: g_ctx = c_null_ptr
    integer, save :: g_ctx_device = -1

contains

    !> The process-wide context on `device` (created on first use).
    function rf_ctx(device) result(ctx)
        integer, intent(in) :: device
        type(c_ptr) :: ctx
        integer(c_int) :: rc
        if (.not. c_associated(g_ctx) .or. g_ctx_device /= device) then
            rc = c_rf_context_create(g_ctx, int(device, c_int))
            call rf_check(rc, "rf_context_create")
            g_ctx_device = device
        end if
        ctx = g_ctx
    end function rf_ctx

    !> Stops with the library's message on a nonzero status.
    subroutine rf_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        type(c_ptr) :: msg
        character(kind=c_char), pointer :: chars(:)
        character(len=512) :: text
        integer :: i
        if (rc == 0) return
        msg = c_rf_last_error(g_ctx)
        text = ""
        if (c_associated(msg)) then
            call c_f_pointer(msg, chars, [512])
            do i = 1, 512
                if (chars(i) == c_null_char) exit
                text(i:i) = chars(i)
            end do
        end if
        write(*,*) "ERROR in ", trim(what), " (status ", rc, "): ", trim(text)
        error stop 1
    end subroutine rf_check

    !> init_RF (type_RF.f90:62-126) with the reference's arguments: the
    !> generation settings come from the input record, comm / nb_procs / rang
    !> are kept in RDF (one GPU runs every subdomain).
    subroutine init_RF_ipt(RF_a, IPT, comm, nb_procs, rang)
        type(RF), intent(inout) :: RF_a
        type(IPT_RF), intent(in) :: IPT
        integer, intent(in) :: comm, nb_procs, rang
        integer :: n
        if (.not. allocated(IPT%corrL)) then
            write(*,*) "ERROR in init_RF: IPT%corrL is not allocated"
            error stop 1
        end if
        call init_RF_args(RF_a, IPT%nDim_gen, IPT%Nmc, IPT%method, IPT%corrMod, IPT%margiFirst, IPT%corrL, &
                          IPT%fieldAvg, IPT%fieldVar, IPT%seedStart)
        RF_a%log_ID = IPT%log_ID
        RF_a%comm = comm
        RF_a%nb_procs = nb_procs
        RF_a%rang = rang
        RF_a%validProc = .true.
        n = size(RF_a%seed)
        allocate(RF_a%kExtent(IPT%nDim_gen), RF_a%neighSeed(n, 3**IPT%nDim_gen - 1))
        allocate(RF_a%neighRange(IPT%nDim_gen, 3**IPT%nDim_gen - 1))
        RF_a%kExtent = -1
        RF_a%neighSeed = -1
    end subroutine init_RF_ipt

    !> init_RF (type_RF.f90:62-126) without the IPT_RF input record.
    subroutine init_RF_args(RDF, nDim, Nmc, method, corrMod, margiFirst, corrL, fieldAvg, fieldVar, seedStart)
        type(RF), intent(inout) :: RDF
        integer, intent(in) :: nDim, Nmc, method, corrMod, margiFirst
        double precision, intent(in) :: corrL(:), fieldAvg, fieldVar
        integer, intent(in), optional :: seedStart
        integer :: n
        call finalize_RF(RDF)
        allocate(RDF%corrL(nDim), RDF%kMax(nDim), RDF%kNStep(nDim), RDF%kDelta(nDim))
        allocate(RDF%calculate(Nmc), RDF%xRange(nDim))
        call random_seed(size = n)
        allocate(RDF%seed(n))
        RDF%nDim = nDim
        RDF%Nmc = Nmc
        RDF%method = method
        RDF%corrMod = corrMod
        RDF%margiFirst = margiFirst
        RDF%corrL = corrL(1:nDim)
        RDF%fieldAvg = fieldAvg
        RDF%fieldVar = fieldVar
        RDF%seedStart = -1
        if (present(seedStart)) RDF%seedStart = seedStart
        RDF%xRange = -1.0D0
        RDF%kMax = -1
        RDF%kDelta = -1
        RDF%seed = -1
        RDF%kNStep = -1
        RDF%calculate(:) = .true.
        RDF%init = .true.
        nullify(RDF%xPoints, RDF%randField, RDF%RF_2D, RDF%RF_3D)
    end subroutine init_RF_args

    subroutine finalize_RF(RDF)
        type(RF), intent(inout) :: RDF
        if (allocated(RDF%corrL)) deallocate(RDF%corrL)
        if (allocated(RDF%kMax)) deallocate(RDF%kMax)
        if (allocated(RDF%kNStep)) deallocate(RDF%kNStep)
        if (allocated(RDF%kDelta)) deallocate(RDF%kDelta)
        if (allocated(RDF%calculate)) deallocate(RDF%calculate)
        if (allocated(RDF%xRange)) deallocate(RDF%xRange)
        if (allocated(RDF%seed)) deallocate(RDF%seed)
        if (allocated(RDF%kPoints)) deallocate(RDF%kPoints)
        if (allocated(RDF%SkVec)) deallocate(RDF%SkVec)
        if (allocated(RDF%xPoints_Local)) deallocate(RDF%xPoints_Local)
        if (allocated(RDF%randField_Local)) deallocate(RDF%randField_Local)
        if (allocated(RDF%kExtent)) deallocate(RDF%kExtent)
        if (allocated(RDF%neighSeed)) deallocate(RDF%neighSeed)
        if (allocated(RDF%neighRange)) deallocate(RDF%neighRange)
        nullify(RDF%xPoints, RDF%randField, RDF%RF_2D, RDF%RF_3D)
        RDF%init = .false.
    end subroutine finalize_RF

    !> init_MESH (type_MESH.f90:61-142) with the reference's arguments: the
    !> mesh settings of the input record; procPerDim = IPT%nFields unless
    !> newNbProcs is given (then -1, set by the caller).
    subroutine init_MESH_ipt(MESH_a, IPT, comm, rang, newNbProcs)
        type(MESH), intent(inout) :: MESH_a
        type(IPT_RF), intent(in) :: IPT
        integer, intent(in) :: comm, rang
        integer, intent(in), optional :: newNbProcs
        integer :: d
        d = IPT%nDim_mesh
        call init_MESH_args(MESH_a, d)
        MESH_a%log_ID = IPT%log_ID
        MESH_a%comm = comm
        MESH_a%rang = rang
        MESH_a%meshMod = IPT%meshMod
        MESH_a%method = IPT%method
        if (present(newNbProcs)) then
            MESH_a%nb_procs = newNbProcs
            MESH_a%procPerDim = -1
        else
            MESH_a%nb_procs = product(IPT%nFields)
            MESH_a%procPerDim = IPT%nFields(1:d)
        end if
        if (allocated(IPT%xMaxGlob)) MESH_a%xMaxGlob = IPT%xMaxGlob(1:d)
        if (allocated(IPT%xMinGlob)) MESH_a%xMinGlob = IPT%xMinGlob(1:d)
        if (allocated(IPT%corrL)) MESH_a%corrL = IPT%corrL(1:d)
        if (allocated(IPT%pointsPerCorrL)) MESH_a%pointsPerCorrL = IPT%pointsPerCorrL(1:d)
        if (allocated(IPT%overlap)) MESH_a%overlap = IPT%overlap(1:d)
        if (allocated(IPT%xStep)) MESH_a%xStep = IPT%xStep(1:d)
        if (allocated(IPT%procExtent)) MESH_a%procExtent = IPT%procExtent(1:d)
    end subroutine init_MESH_ipt

    !> init_MESH (type_MESH.f90:61-142) without the IPT_RF input record.
    subroutine init_MESH_args(MSH, nDim)
        type(MESH), intent(inout) :: MSH
        integer, intent(in) :: nDim
        call finalize_MESH(MSH)
        allocate(MSH%xNStep(nDim), MSH%xStep(nDim), MSH%xMinGlob(nDim), MSH%xMaxGlob(nDim))
        allocate(MSH%xMinInt(nDim), MSH%xMaxInt(nDim), MSH%xMinExt(nDim), MSH%xMaxExt(nDim))
        allocate(MSH%xMinBound(nDim), MSH%xMaxBound(nDim))
        allocate(MSH%procPerDim(nDim), MSH%coords(nDim), MSH%corrL(nDim), MSH%procStart(nDim))
        allocate(MSH%procExtent(nDim), MSH%pointsPerCorrL(nDim), MSH%origin(nDim), MSH%overlap(nDim))
        allocate(MSH%neigh(3**nDim - 1), MSH%op_neigh(3**nDim - 1), MSH%neighShift(nDim, 3**nDim - 1))
        MSH%procPerDim = 1
        MSH%coords = 0
        MSH%corrL = 1.0D0
        MSH%procStart = 0.0D0
        MSH%procExtent = 0.0D0
        MSH%pointsPerCorrL = 1
        MSH%origin = 1
        MSH%overlap = 0.0D0
        MSH%neigh = -2
        MSH%op_neigh = -2
        MSH%neighShift = 0
        MSH%nDim = nDim
        MSH%xNStep = -1
        MSH%xStep = 1.0D0
        MSH%xMinGlob = 0.0D0; MSH%xMaxGlob = 0.0D0
        MSH%xMinInt = -1; MSH%xMaxInt = -1; MSH%xMinExt = -1; MSH%xMaxExt = -1
        MSH%xMinBound = -1; MSH%xMaxBound = -1
        MSH%init = .true.
    end subroutine init_MESH_args

    subroutine finalize_MESH(MSH)
        type(MESH), intent(inout) :: MSH
        if (allocated(MSH%xNStep)) deallocate(MSH%xNStep)
        if (allocated(MSH%xStep)) deallocate(MSH%xStep)
        if (allocated(MSH%xMinGlob)) deallocate(MSH%xMinGlob)
        if (allocated(MSH%xMaxGlob)) deallocate(MSH%xMaxGlob)
        if (allocated(MSH%xMinInt)) deallocate(MSH%xMinInt)
        if (allocated(MSH%xMaxInt)) deallocate(MSH%xMaxInt)
        if (allocated(MSH%xMinExt)) deallocate(MSH%xMinExt)
        if (allocated(MSH%xMaxExt)) deallocate(MSH%xMaxExt)
        if (allocated(MSH%xMinBound)) deallocate(MSH%xMinBound)
        if (allocated(MSH%xMaxBound)) deallocate(MSH%xMaxBound)
        if (allocated(MSH%procPerDim)) deallocate(MSH%procPerDim)
        if (allocated(MSH%coords)) deallocate(MSH%coords)
        if (allocated(MSH%corrL)) deallocate(MSH%corrL)
        if (allocated(MSH%procStart)) deallocate(MSH%procStart)
        if (allocated(MSH%procExtent)) deallocate(MSH%procExtent)
        if (allocated(MSH%pointsPerCorrL)) deallocate(MSH%pointsPerCorrL)
        if (allocated(MSH%origin)) deallocate(MSH%origin)
        if (allocated(MSH%overlap)) deallocate(MSH%overlap)
        if (allocated(MSH%neigh)) deallocate(MSH%neigh)
        if (allocated(MSH%op_neigh)) deallocate(MSH%op_neigh)
        if (allocated(MSH%neighShift)) deallocate(MSH%neighShift)
        MSH%init = .false.
    end subroutine finalize_MESH

    !> calculate_random_seed (math_RF.f90:538-566) for a deterministic start:
    !> seed(k) = 72 + 18*seedStart + 37*(k-1); a negative start takes the
    !> system clock as the reference does.
    subroutine calculate_random_seed(seed, seedStart)
        integer, dimension(:), allocatable, intent(out) :: seed
        integer, optional, intent(in) :: seedStart
        integer :: i, n, clock
        call random_seed(size = n)
        allocate(seed(n))
        call system_clock(count = clock)
        if (present(seedStart)) then
            if (seedStart >= 0) then
                seed = 72 + seedStart*18 + 37*(/ (i - 1, i = 1, n) /)
                return
            end if
        end if
        seed = clock + 37*(/ (i - 1, i = 1, n) /)
    end subroutine calculate_random_seed

    !> init_random_seed (math_RF.f90:475-497)
    subroutine init_random_seed(seedIn, seedStart)
        integer, dimension(1:), optional, intent(in) :: seedIn
        integer, optional, intent(in) :: seedStart
        integer, dimension(:), allocatable :: seed
        if (present(seedIn)) then
            call random_seed(put = seedIn)
        else
            call calculate_random_seed(seed, seedStart)
            call random_seed(put = seed)
            deallocate(seed)
        end if
    end subroutine init_random_seed

    !> The C parameter block of RDF (coordinates normalised inside the library).
    function rf_c_params(RDF) result(p)
        type(RF), intent(in) :: RDF
        type(rf_params_c) :: p
        integer :: i
        p%nDim = RDF%nDim
        p%method = RDF%method
        p%corrMod = RDF%corrMod
        p%mode = RDF%compatMode
        p%form = RDF%form
        p%kSampling = RDF%kSampling
        p%Nmc = RDF%Nmc
        p%margiFirst = 0
        p%nKDensity = RDF%nKDensity
        p%xSlab = 0
        p%corrL = 1d0
        p%xRange = 1d0
        do i = 1, RDF%nDim
            p%corrL(i) = RDF%corrL(i)
            p%xRange(i) = RDF%xRange(i)
        end do
        p%fieldAvg = RDF%fieldAvg
        p%fieldVar = RDF%fieldVar
        p%seed = RDF%deviceSeed
        p%isoShared = merge(1, 0, RDF%isoShared)
    end function rf_c_params

    !> The uniforms of one generation call, drawn in the reference's order
    !> after random_seed(PUT=seed): kU (method 3 wavevectors, :120) then
    !> phU(:, n) per realization (methods 2/3 always, :276-279; method 1 only
    !> when calculate(n), :420-426 / :456-463).
    subroutine rf_draw_sequence(p, info, seed, calc, kU, phU)
        type(rf_params_c), intent(in) :: p
        type(rf_spectrum_info_c), intent(in) :: info
        integer, dimension(:), intent(in) :: seed
        logical, dimension(:), intent(in) :: calc
        double precision, dimension(:), intent(out) :: kU
        double precision, dimension(:, :), intent(out) :: phU
        integer :: n, v, nv
        integer(kind=8) :: kN
        call random_seed(put = seed)
        if (info%nKUniform > 0) call random_number(kU(1:info%nKUniform))
        kN = info%kNTotal
        do n = 1, p%Nmc
            if (p%method == ISOTROPIC) then
                if (.not. calc(n)) then
                    phU(:, n) = 0d0
                    cycle
                end if
                nv = info%nVarIso
                do v = 1, nv
                    call random_number(phU((v - 1)*kN + 1:v*kN, n))
                end do
            else if (p%method == FFT) then
                ! gen_Std_Gauss_FFT on one rank (randomFieldND.f90:670-681):
                ! gammaK (kN), kN-1 discarded draws, phiK (kN), in one stream
                call random_number(phU(1:info%nPhaseUniform, n))
            else
                call random_number(phU(1:kN, n))
            end if
        end do
    end subroutine rf_draw_sequence

    !> C entry for test harnesses: the same draw sequence as
    !> create_RF_Unstruct_Init for parameter block p and seed vector seed.
    subroutine rf_fortran_draws(p, seed, nseed, calc, kU, phU) bind(C, name="rf_fortran_draws")
        type(rf_params_c), intent(in) :: p
        integer(c_int), value :: nseed
        integer(c_int), intent(in) :: seed(nseed)
        integer(c_int8_t), intent(in) :: calc(*)
        real(c_double), intent(out) :: kU(*)
        real(c_double), intent(out) :: phU(*)
        type(rf_spectrum_info_c) :: info
        integer(c_int) :: rc
        integer :: n, ns
        integer, allocatable :: s(:)
        logical, allocatable :: lc(:)
        double precision, allocatable :: ku_l(:), ph_l(:,:)
        rc = c_rf_plan_spectrum(p, info)
        call rf_check(rc, "rf_plan_spectrum")
        call random_seed(size = ns)
        allocate(s(ns))
        s = seed(1)
        do n = 1, min(ns, int(nseed))
            s(n) = seed(n)
        end do
        allocate(lc(p%Nmc))
        do n = 1, p%Nmc
            lc(n) = calc(n) /= 0
        end do
        allocate(ku_l(max(1_8, info%nKUniform)), ph_l(info%nPhaseUniform, p%Nmc))
        call rf_draw_sequence(p, info, s, lc, ku_l, ph_l)
        if (info%nKUniform > 0) kU(1:info%nKUniform) = ku_l(1:info%nKUniform)
        phU(1:info%nPhaseUniform*p%Nmc) = reshape(ph_l, [info%nPhaseUniform*p%Nmc])
    end subroutine rf_fortran_draws

    !> Number of integers random_seed takes (the compiler's seed size).
    integer(c_int) function rf_fortran_seed_size() bind(C, name="rf_fortran_seed_size")
        integer :: n
        call random_seed(size = n)
        rf_fortran_seed_size = n
    end function rf_fortran_seed_size

    !> create_RF_Unstruct_Init (randomFieldND.f90:27-132) on the GPU.
    !> Preconditions as in the reference: RDF initialised, RDF%xPoints =>
    !> nDim x N points, RDF%randField => N x Nmc storage, RDF%xRange set,
    !> product(MSH%xNStep) = N (randomFieldND.f90:209).
    subroutine create_RF_Unstruct_Init(RDF, MSH)
        type(RF), intent(inout), target :: RDF
        type(MESH), intent(inout) :: MSH
        type(rf_params_c) :: p
        type(rf_spectrum_info_c) :: info
        integer(kind=8) :: N
        integer(c_int) :: rc
        integer :: i
        double precision, allocatable, target :: kU(:), phU(:,:)
        integer(c_int8_t), allocatable, target :: calc(:)
        type(c_ptr) :: kU_p, phU_p

        if (.not. RDF%init) call fail("RDF is not initialised (init_RF)")
        if (.not. associated(RDF%xPoints)) call fail("RDF%xPoints is not associated")
        if (.not. associated(RDF%randField)) call fail("RDF%randField is not associated")
        if (size(RDF%xPoints, 1) /= RDF%nDim) call fail("size(xPoints,1) /= nDim")
        N = size(RDF%xPoints, 2, kind=8)
        if (.not. allocated(MSH%xNStep)) call fail("MSH%xNStep is not allocated (init_MESH)")
        if (product(int(MSH%xNStep, 8)) /= N) call fail("product(MSH%xNStep) /= number of points")
        if (size(RDF%randField, 1, kind=8) /= N .or. size(RDF%randField, 2) /= RDF%Nmc) &
            call fail("randField must be N x Nmc")
        do i = 1, RDF%nDim
            if (.not. (RDF%xRange(i) > 0d0)) call fail("RDF%xRange must be set and > 0")
        end do
        MSH%xNTotal = N

        p = rf_c_params(RDF)
        if (RDF%method == FFT) then
            do i = 1, RDF%nDim
                p%gridN(i) = MSH%xNStep(i)
            end do
        end if
        rc = c_rf_plan_spectrum(p, info)
        call rf_check(rc, "rf_plan_spectrum")

        allocate(calc(RDF%Nmc))
        do i = 1, RDF%Nmc
            calc(i) = merge(1_c_int8_t, 0_c_int8_t, RDF%calculate(i))
        end do
        allocate(kU(max(1_8, info%nKUniform)), phU(info%nPhaseUniform, RDF%Nmc))
        kU_p = c_null_ptr
        phU_p = c_null_ptr
        if (.not. RDF%deviceDraws) then
            if (.not. allocated(RDF%seed)) call fail("RDF%seed is not allocated")
            call rf_draw_sequence(p, info, RDF%seed, RDF%calculate, kU, phU)
            if (info%nKUniform > 0) kU_p = c_loc(kU)
            phU_p = c_loc(phU)
        end if

        if (RDF%method == FFT) then
            ! gen_Std_Gauss_FFT (randomFieldND.f90:520-763): the whole structured
            ! mesh MSH%xNStep from MSH%xMinGlob at MSH%xStep, x fastest
            if (.not. allocated(MSH%xStep) .or. .not. allocated(MSH%xMinGlob)) &
                call fail("method 4 needs MSH%xStep and MSH%xMinGlob")
            block
                real(c_double) :: org(3), stp(3)
                integer(c_int32_t) :: ns(3)
                double precision, allocatable, target :: fs(:,:)
                org = 0d0
                stp = 1d0
                ns = 1
                do i = 1, RDF%nDim
                    org(i) = MSH%xMinGlob(i)
                    stp(i) = MSH%xStep(i)
                    ns(i) = MSH%xNStep(i)
                end do
                allocate(fs(N, RDF%Nmc))
                rc = c_rf_generate_grid(rf_ctx(RDF%device), p, org, stp, ns, phU_p, c_loc(calc), c_loc(fs), &
                                        0_c_int32_t)
                if (rc == 0) RDF%randField = fs
            end block
            call rf_check(rc, "rf_generate_grid")
        else if (is_contiguous(RDF%xPoints) .and. is_contiguous(RDF%randField)) then
            rc = c_rf_generate(rf_ctx(RDF%device), p, c_loc(RDF%xPoints), int(N, c_int64_t), &
                               kU_p, phU_p, c_loc(calc), c_loc(RDF%randField), 0_c_int32_t)
        else
            ! pointer sections: stage through contiguous copies
            block
                double precision, allocatable, target :: xs(:,:), fs(:,:)
                allocate(xs(RDF%nDim, N), fs(N, RDF%Nmc))
                xs = RDF%xPoints
                rc = c_rf_generate(rf_ctx(RDF%device), p, c_loc(xs), int(N, c_int64_t), &
                                   kU_p, phU_p, c_loc(calc), c_loc(fs), 0_c_int32_t)
                if (rc == 0) RDF%randField = fs
            end block
        end if
        if (RDF%method /= FFT) call rf_check(rc, "rf_generate")

        ! spectral state the reference leaves in RDF (spectra_RF.f90:68-133)
        RDF%kNTotal = info%kNTotal
        RDF%kMax = info%kMax(1:RDF%nDim)
        RDF%kDelta = info%kDelta(1:RDF%nDim)
        RDF%kNStep = info%kNStep(1:RDF%nDim)
        if (allocated(RDF%kPoints)) deallocate(RDF%kPoints)
        if (allocated(RDF%SkVec)) deallocate(RDF%SkVec)
        if (RDF%method == ISOTROPIC) then
            allocate(RDF%kPoints(1, info%kNTotal))
        else
            allocate(RDF%kPoints(RDF%nDim, info%kNTotal))
        end if
        allocate(RDF%SkVec(info%kNTotal))
        if (.not. RDF%deviceDraws .or. info%nKUniform == 0) then
            rc = c_rf_spectrum_tables(p, merge(c_loc(kU), c_null_ptr, info%nKUniform > 0), &
                                      c_loc(RDF%kPoints), c_loc(RDF%SkVec))
            call rf_check(rc, "rf_spectrum_tables")
        end if
    contains
        subroutine fail(msg)
            character(len=*), intent(in) :: msg
            write(*,*) "ERROR in create_RF_Unstruct_Init: ", msg
            error stop 1
        end subroutine fail
    end subroutine create_RF_Unstruct_Init

    !> multiVariateTransformation (randomFieldND.f90:139-178), on the GPU.
    subroutine multiVariateTransformation(margiFirst, fieldAvg, fieldVar, randField)
        integer, intent(in) :: margiFirst
        double precision, intent(in) :: fieldAvg, fieldVar
        double precision, dimension(1:, 1:), intent(inout), target, contiguous :: randField
        integer(c_int) :: rc
        rc = c_rf_marginal_host(rf_ctx(0), int(margiFirst, c_int32_t), fieldAvg, fieldVar, &
                                c_loc(randField), size(randField, kind=c_int64_t))
        call rf_check(rc, "multiVariateTransformation")
    end subroutine multiVariateTransformation

    !> README.md:58-88 with string model names ("gaussian"; "gaussian" or
    !> "lognormal").  The box extent of the points sets the wavevector lattice
    !> (xRange); seedStart (default 0) gives the reference seed vector
    !> 72 + 18*seedStart + 37*(k-1).
    subroutine createStdGaussUnstruct_str(xPoints, corrMod, margiFirst, corrL, fieldAvg, fieldVar, &
                                          Nmc, method, randField, seedStart, mode, form)
        double precision, dimension(:, :), intent(in), target, contiguous :: xPoints
        character(len=*), intent(in) :: corrMod, margiFirst
        double precision, dimension(:), intent(in) :: corrL
        double precision, intent(in) :: fieldAvg, fieldVar
        integer, intent(in) :: Nmc, method
        double precision, dimension(:, :), intent(inout), target, contiguous :: randField
        integer, intent(in), optional :: seedStart, mode, form
        integer :: cm, mf
        select case (trim(adjustl(corrMod)))
            case ("gaussian"); cm = cm_GAUSSIAN
            case ("cos", "cosinus"); cm = cm_COS
            case default
                write(*,*) "ERROR: unknown corrMod '", trim(corrMod), "'"
                error stop 1
        end select
        select case (trim(adjustl(margiFirst)))
            case ("gaussian"); mf = fom_GAUSSIAN
            case ("lognormal"); mf = fom_LOGNORMAL
            case default
                write(*,*) "ERROR: unknown margiFirst '", trim(margiFirst), "'"
                error stop 1
        end select
        call createStdGaussUnstruct_int(xPoints, cm, mf, corrL, fieldAvg, fieldVar, Nmc, method, &
                                        randField, seedStart, mode, form)
    end subroutine createStdGaussUnstruct_str

    subroutine createStdGaussUnstruct_int(xPoints, corrMod, margiFirst, corrL, fieldAvg, fieldVar, &
                                          Nmc, method, randField, seedStart, mode, form)
        double precision, dimension(:, :), intent(in), target, contiguous :: xPoints
        integer, intent(in) :: corrMod, margiFirst
        double precision, dimension(:), intent(in) :: corrL
        double precision, intent(in) :: fieldAvg, fieldVar
        integer, intent(in) :: Nmc, method
        double precision, dimension(:, :), intent(inout), target, contiguous :: randField
        integer, intent(in), optional :: seedStart, mode, form
        type(RF) :: RDF
        type(MESH) :: MSH
        integer :: nDim, i, s
        integer, allocatable :: seed(:)
        nDim = size(xPoints, 1)
        call init_RF(RDF, nDim, Nmc, method, corrMod, margiFirst, corrL, fieldAvg, fieldVar)
        call init_MESH(MSH, nDim)
        MSH%xNStep = 1
        MSH%xNStep(1) = size(xPoints, 2)
        do i = 1, nDim
            RDF%xRange(i) = maxval(xPoints(i, :)) - minval(xPoints(i, :))
            if (.not. (RDF%xRange(i) > 0d0)) RDF%xRange(i) = corrL(i)
        end do
        s = 0
        if (present(seedStart)) s = seedStart
        call calculate_random_seed(seed, s)
        RDF%seed = seed
        if (present(mode)) RDF%compatMode = mode
        if (present(form)) RDF%form = form
        RDF%xPoints => xPoints
        RDF%randField => randField
        call create_RF_Unstruct_Init(RDF, MSH)
        call multiVariateTransformation(margiFirst, fieldAvg, fieldVar, randField)
        call finalize_RF(RDF)
        call finalize_MESH(MSH)
    end subroutine createStdGaussUnstruct_int

    ! ------------------------------------------------------------ localization

    !> The library's localization geometry for the settings of MSH.
    subroutine layout_of_mesh(MSH, overlapIn, info, what)
        type(MESH), intent(in) :: MSH
        double precision, dimension(:), intent(in) :: overlapIn
        type(rf_layout_info_c), intent(out) :: info
        character(len=*), intent(in) :: what
        type(rf_layout_params_c) :: p
        integer :: d
        integer(c_int) :: rc
        p%nDim = MSH%nDim
        do d = 1, MSH%nDim
            p%nFields(d) = MSH%procPerDim(d)
            p%pointsPerCorrL(d) = MSH%pointsPerCorrL(d)
            p%xMinGlob(d) = MSH%xMinGlob(d)
            p%xMaxGlob(d) = MSH%xMaxGlob(d)
            p%corrL(d) = MSH%corrL(d)
            p%overlap(d) = overlapIn(d)
        end do
        rc = c_rf_localization_layout(p, info, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr)
        call rf_check(rc, what)
    end subroutine layout_of_mesh

    !> round_basic_inputs (topography_RF.f90:19-58): xStep = corrL /
    !> pointsPerCorrL; overlap (in correlation lengths) rounded to whole steps,
    !> at least two steps where MSH%procPerDim > 1, zero where it is 1.
    !> xStep / overlap may be MSH%xStep / MSH%overlap themselves.
    subroutine round_basic_inputs(MSH, xStep, overlap)
        type(MESH), intent(in) :: MSH
        double precision, dimension(:), intent(inout) :: xStep, overlap
        type(rf_layout_info_c) :: info
        double precision :: ovl(MSH%nDim)
        ovl = overlap(1:MSH%nDim)
        call layout_of_mesh(MSH, ovl, info, "round_basic_inputs")
        xStep(1:MSH%nDim) = info%xStep(1:MSH%nDim)
        overlap(1:MSH%nDim) = info%overlap(1:MSH%nDim)
    end subroutine round_basic_inputs

    !> set_global_extremes (topography_RF.f90:63-144): the global box rounded
    !> to whole steps and grown so that procPerDim subdomains of equal extent
    !> overlap by MSH%overlap; procStart = xMinGlob + stepProc * MSH%coords.
    subroutine set_global_extremes(MSH, xMaxGlob, xMinGlob, procExtent, procStart, stepProc_out)
        type(MESH), intent(in) :: MSH
        double precision, dimension(:), intent(inout) :: xMinGlob, xMaxGlob
        double precision, dimension(:), intent(out) :: procExtent, procStart
        double precision, dimension(:), intent(out), optional :: stepProc_out
        type(rf_layout_info_c) :: info
        integer :: d, n
        n = MSH%nDim
        call layout_of_mesh(MSH, MSH%overlap, info, "set_global_extremes")
        do d = 1, n
            xMinGlob(d) = info%xMinGlob(d)
            xMaxGlob(d) = info%xMaxGlob(d)
            procExtent(d) = info%procExtent(d)
            procStart(d) = info%xMinGlob(d) + info%stepProc(d) * dble(MSH%coords(d))
            if (present(stepProc_out)) stepProc_out(d) = info%stepProc(d)
        end do
    end subroutine set_global_extremes

    !> prepare_Localization (localization_RF.f90:127-192) for one group
    !> holding every field: subdivisionCoords(:, i) = IPT%xMinGlob +
    !> IPT%stepProc * c with the LAST axis fastest over nFields**level
    !> subdivisions (setGrid inverse), subdivisionId = c, the files' boxes, the
    !> group box and its point count and origin (1-based point of the global
    !> grid).
    subroutine prepare_Localization(IPT, subdivisionCoords, subdivisionId, xMinFiles, xMaxFiles, &
                                    xMin_Group, xMax_Group, gen_GroupRange, xNStep_Group, origin_Group, &
                                    xNTotal_Group)
        type(IPT_RF), intent(in) :: IPT
        double precision, dimension(:, :), intent(out) :: subdivisionCoords, xMinFiles, xMaxFiles
        integer, dimension(:, :), intent(out) :: subdivisionId
        double precision, dimension(:), intent(out) :: xMin_Group, xMax_Group, gen_GroupRange
        integer, dimension(:), intent(out) :: xNStep_Group, origin_Group
        integer(kind=8), intent(out) :: xNTotal_Group
        integer :: d, i, n, rem, nTot
        integer :: locStep(IPT%nDim)
        n = IPT%nDim
        locStep = IPT%nFields(1:n)**IPT%localizationLevel
        nTot = product(locStep)
        if (size(subdivisionCoords, 2) < nTot .or. size(xMinFiles, 2) < nTot) then
            write(*,*) "ERROR in prepare_Localization: arrays hold fewer than ", nTot, " fields"
            error stop 1
        end if
        do i = 1, nTot
            rem = i - 1
            do d = n, 1, -1
                subdivisionId(d, i) = modulo(rem, locStep(d))
                rem = rem / locStep(d)
            end do
            do d = 1, n
                subdivisionCoords(d, i) = IPT%xMinGlob(d) + IPT%stepProc(d) * dble(subdivisionId(d, i))
            end do
            xMinFiles(1:n, i) = subdivisionCoords(1:n, i)
            xMaxFiles(1:n, i) = subdivisionCoords(1:n, i) + IPT%procExtent(1:n)
        end do
        do d = 1, n
            xMin_Group(d) = minval(xMinFiles(d, 1:nTot))
            xMax_Group(d) = maxval(xMaxFiles(d, 1:nTot))
            gen_GroupRange(d) = xMax_Group(d) - xMin_Group(d)
            xNStep_Group(d) = 1 + nint(gen_GroupRange(d) / IPT%xStep(d))
            origin_Group(d) = 1 + nint((xMin_Group(d) - IPT%xMinGlob(d)) / IPT%xStep(d))
        end do
        xNTotal_Group = product(int(xNStep_Group(1:n), 8))
    end subroutine prepare_Localization

    !> generateUnityPartition_Matrix (localization_RF.f90:705-857): the
    !> partition-of-unity weights of one subdomain of xNStep points, tapered
    !> over nint((overlap corrL - 2 xStep) / xStep) + 1 points on every face,
    !> or, with neigh / neighShift, on the faces that have a neighbour (the
    !> others with reverse).  partitionType 1 (the cosine taper; the only one
    !> the reference calls).  mode RF_REF_COMPAT (default): the reference's
    !> weights; RF_FIXED: their square-root exact form.
    subroutine generateUnityPartition_Matrix(xNStep, overlap, corrL, xStep, partitionType, unityPartition, &
                                             nDim, neigh, neighShift, reverse, mode)
        integer, dimension(:), intent(in) :: xNStep
        double precision, dimension(:), intent(in) :: overlap, corrL, xStep
        integer, intent(in) :: partitionType, nDim
        double precision, dimension(:), intent(out), target, contiguous :: unityPartition
        integer, dimension(:), intent(in), optional :: neigh
        integer, dimension(:, :), intent(in), optional :: neighShift
        logical, intent(in), optional :: reverse
        integer, intent(in), optional :: mode
        integer(c_int32_t) :: ns(3), no(3), mask, md
        integer :: i, f
        integer(c_int) :: rc
        logical :: has
        if (partitionType /= 1) then
            write(*,*) "ERROR in generateUnityPartition_Matrix: partitionType ", partitionType, &
                       " (only the cosine taper, 1, is implemented)"
            error stop 1
        end if
        ns = 1
        no = 0
        mask = 0
        do i = 1, nDim
            ns(i) = xNStep(i)
            no(i) = nint((overlap(i) * corrL(i) - 2 * xStep(i)) / xStep(i)) + 1
            do f = 0, 1
                has = .true.
                if (present(neigh) .and. present(neighShift)) then
                    has = any(neighShift(i, :) == 2 * f - 1 .and. neigh >= 0)
                    if (present(reverse)) then
                        if (reverse) has = .not. has
                    end if
                end if
                if (has) mask = ior(mask, ishft(1_c_int32_t, 2 * (i - 1) + f))
            end do
        end do
        md = RF_REF_COMPAT
        if (present(mode)) md = mode
        if (size(unityPartition, kind=8) < product(int(ns, 8))) then
            write(*,*) "ERROR in generateUnityPartition_Matrix: unityPartition holds fewer than product(xNStep)"
            error stop 1
        end if
        rc = c_rf_unity_partition(int(nDim, c_int32_t), ns, no, mask, md, c_loc(unityPartition))
        call rf_check(rc, "generateUnityPartition_Matrix")
    end subroutine generateUnityPartition_Matrix

    !> The localization step of build_random_field (calls_RF.f90:209-369) for
    !> every field of the group at once, on the GPU: field s (its points
    !> xNStep_Proc from xMinFiles(:, s), x fastest, Nmc realizations) weighted
    !> by its partition of unity and added into the group grid xNStep_Group
    !> from xMin_Group (add_RF_to_Group, localization_RF.f90:199-263); in
    !> RF_REF_COMPAT mode (default) every face is tapered and realization 1 of
    !> the group is then divided by the taper of the group's own faces
    !> (calls_RF.f90:324-369); RF_FIXED tapers the faces shared with another
    !> field with the square-root weights.  fields(:, :, s), randField_Group
    !> (xNTotal_Group x Nmc) is overwritten.
    subroutine localization_blend(nDim, Nmc, xStep, overlap, corrL, xNStep_Proc, xMinFiles, xMin_Group, &
                                  xNStep_Group, fields, randField_Group, mode)
        integer, intent(in) :: nDim, Nmc
        double precision, dimension(:), intent(in) :: xStep, overlap, corrL, xMin_Group
        integer, dimension(:), intent(in) :: xNStep_Proc, xNStep_Group
        double precision, dimension(:, :), intent(in) :: xMinFiles
        double precision, dimension(:, :, :), intent(in), target, contiguous :: fields
        double precision, dimension(:, :), intent(inout), target, contiguous :: randField_Group
        integer, intent(in), optional :: mode
        integer(c_int32_t), allocatable :: org(:), siz(:), masks(:)
        integer(c_int32_t) :: gn(3), no(3), md
        type(c_ptr), allocatable :: ptrs(:)
        integer :: nSub, s, d
        integer(c_int) :: rc
        nSub = size(fields, 3)
        md = RF_REF_COMPAT
        if (present(mode)) md = mode
        if (size(fields, 1, kind=8) /= product(int(xNStep_Proc(1:nDim), 8)) .or. size(fields, 2) < Nmc) then
            write(*,*) "ERROR in localization_blend: fields must be product(xNStep_Proc) x Nmc x nFields"
            error stop 1
        end if
        if (size(randField_Group, 1, kind=8) /= product(int(xNStep_Group(1:nDim), 8)) .or. &
            size(randField_Group, 2) /= Nmc) then
            write(*,*) "ERROR in localization_blend: randField_Group must be product(xNStep_Group) x Nmc"
            error stop 1
        end if
        allocate(org(nDim * nSub), siz(nDim * nSub), masks(nSub), ptrs(nSub))
        gn = 1
        no = 0
        do d = 1, nDim
            gn(d) = xNStep_Group(d)
            no(d) = max(0, nint((overlap(d) * corrL(d) - 2 * xStep(d)) / xStep(d)) + 1)
        end do
        do s = 1, nSub
            masks(s) = 0
            do d = 1, nDim
                org((s - 1) * nDim + d) = nint((xMinFiles(d, s) - xMin_Group(d)) / xStep(d))
                siz((s - 1) * nDim + d) = xNStep_Proc(d)
                if (md == RF_REF_COMPAT .or. org((s - 1) * nDim + d) > 0) &
                    masks(s) = ior(masks(s), ishft(1_c_int32_t, 2 * (d - 1)))
                if (md == RF_REF_COMPAT .or. org((s - 1) * nDim + d) + xNStep_Proc(d) < xNStep_Group(d)) &
                    masks(s) = ior(masks(s), ishft(1_c_int32_t, 2 * (d - 1) + 1))
            end do
            ptrs(s) = c_loc(fields(1, 1, s))
        end do
        rc = c_rf_blend_host(rf_ctx(0), int(nDim, c_int32_t), gn, int(Nmc, c_int32_t), int(nSub, c_int32_t), &
                             org, siz, no, masks, md, ptrs, c_loc(randField_Group))
        call rf_check(rc, "localization_blend")
    end subroutine localization_blend

    !> normalize_randField (randomFieldND.f90:820-926), on the GPU: the mean
    !> and then the root mean square of the window minPos..maxPos of RF_2D /
    !> RF_3D (the grid view of realization 1) are subtracted from / divided
    !> into realizations 1..Nmc of randField.  loc_Comm is not used (one GPU
    !> holds the whole window).
    subroutine normalize_randField(randField, minPos, maxPos, nDim, Nmc, loc_Comm, RF_2D, RF_3D)
        double precision, dimension(:, :), intent(inout), target, contiguous :: randField
        integer, dimension(:), intent(in) :: minPos, maxPos
        integer, intent(in) :: nDim, loc_Comm, Nmc
        double precision, dimension(:, :), intent(in), optional :: RF_2D
        double precision, dimension(:, :, :), intent(in), optional :: RF_3D
        integer(c_int32_t) :: gn(3), lo(3), hi(3)
        integer(c_int) :: rc
        integer :: d
        if (loc_Comm < -huge(0)) continue
        gn = 1
        if (nDim == 2 .and. present(RF_2D)) then
            gn(1:2) = shape(RF_2D)
        else if (nDim == 3 .and. present(RF_3D)) then
            gn(1:3) = shape(RF_3D)
        else
            write(*,*) "ERROR in normalize_randField: nDim = ", nDim, " needs RF_2D / RF_3D"
            error stop 1
        end if
        if (size(randField, 1, kind=8) /= product(int(gn, 8)) .or. size(randField, 2) < Nmc) then
            write(*,*) "ERROR in normalize_randField: randField must be product(shape(RF_nD)) x >= Nmc"
            error stop 1
        end if
        lo = 1
        hi = 1
        do d = 1, nDim
            lo(d) = minPos(d)
            hi(d) = maxPos(d)
        end do
        rc = c_rf_normalize_window_host(rf_ctx(0), c_loc(randField), int(nDim, c_int32_t), gn, lo, hi, &
                                        int(Nmc, c_int32_t))
        call rf_check(rc, "normalize_randField")
    end subroutine normalize_randField

end module randomField_gpu
