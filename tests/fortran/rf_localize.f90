!> Fortran host of the localization step written against the reference's
!> calling surface: an input record (type(IPT_RF)), init_MESH / init_RF with
!> the reference's argument lists, round_basic_inputs and
!> set_global_extremes on the global mesh (sample_RF.f90:207-216),
!> prepare_Localization (calls_RF.f90:195-199), one createRandomField per
!> field on its own grid with seed seedStart + field number
!> (sample_RF.f90:280-348), the blend of every field (calls_RF.f90:209-369),
!> normalize_randField over the window of non-repeated points and the
!> marginal map (sample_RF.f90:820-847).
!>
!> usage: rf_localize <mode> <out.bin>
!>   mode 1: geometry and partition of unity only (no GPU)
!>   mode 2: the whole step, REF_COMPAT (needs the GPU)
!>   mode 3: the whole step, FIXED blend weights (needs the GPU)
!> Settings: 3-D box [0,6]^3 (x) with corrL (1, 1, 1), 2 points per corrL,
!> overlap 2 corrL, nFields (2, 2, 2), Nmc 2, seedStart 100, method 2,
!> gaussian marginal avg 0 var 1.
!> out.bin (stream): int32 nDim, nFields(3), Nmc; float64 xStep(3),
!>   overlap(3), xMinGlob(3), xMaxGlob(3), procExtent(3), procStart(3),
!>   stepProc(3); int32 xNStep_Proc(3), xNStep_Group(3), origin_Group(3);
!>   float64 subdivisionCoords(3, 8), xMinFiles(3, 8), xMin_Group(3),
!>   unityPartition(prod xNStep_Proc); then for modes 2/3: float64
!>   fields(prod xNStep_Proc, Nmc, 8), group(prod xNStep_Group, Nmc) after
!>   the blend, group after normalize_randField, int32 minP(3), maxP(3)
program rf_localize
    use randomField_gpu
    implicit none
    character(len=256) :: arg, outPath
    integer :: imode, u, i, d, j, nTot, Nmc
    type(IPT_RF) :: IPT
    type(MESH) :: globMSH, MSH
    type(RF) :: RDF
    double precision :: stepProc(3), xMin_Group(3), xMax_Group(3), gen_GroupRange(3)
    double precision, allocatable :: subdivisionCoords(:,:), xMinFiles(:,:), xMaxFiles(:,:)
    integer, allocatable :: subdivisionId(:,:)
    integer :: xNStep_Group(3), origin_Group(3), xNStep_Proc(3), minP(3), maxP(3), nOvl(3)
    integer(kind=8) :: xNTotal_Group, xNTotal_Proc
    double precision, allocatable, target :: pu(:), fields(:,:,:), group(:,:), x(:,:), f(:,:)
    double precision, pointer :: RF_3D(:,:,:)
    double precision, allocatable, target :: groupNorm(:,:)
    integer, allocatable :: seed(:)

    call get_command_argument(1, arg)
    read(arg, *) imode
    call get_command_argument(2, outPath)
    Nmc = 2

    ! the input record (type_inputRF.f90:10-70)
    IPT%nDim = 3; IPT%nDim_mesh = 3; IPT%nDim_gen = 3
    allocate(IPT%xMinGlob(3), IPT%xMaxGlob(3), IPT%pointsPerCorrL(3), IPT%nFields(3), IPT%corrL(3))
    allocate(IPT%overlap(3), IPT%xStep(3), IPT%stepProc(3), IPT%procExtent(3))
    IPT%xMinGlob = 0d0
    IPT%xMaxGlob = 6d0
    IPT%pointsPerCorrL = 2
    IPT%nFields = 2
    IPT%corrL = 1d0
    IPT%overlap = 2d0
    IPT%xStep = 0.5d0
    IPT%procExtent = 0d0
    IPT%localizationLevel = 1
    IPT%Nmc = Nmc
    IPT%method = SHINOZUKA
    IPT%corrMod = cm_GAUSSIAN
    IPT%margiFirst = fom_GAUSSIAN
    IPT%fieldAvg = 0d0
    IPT%fieldVar = 1d0
    IPT%seedStart = 100

    ! process_input (sample_RF.f90:207-226)
    call init_MESH(globMSH, IPT, 0, 0)
    call round_basic_inputs(globMSH, globMSH%xStep, globMSH%overlap)
    call set_global_extremes(globMSH, globMSH%xMaxGlob, globMSH%xMinGlob, globMSH%procExtent, &
                             globMSH%procStart, stepProc)
    IPT%xStep = globMSH%xStep
    IPT%overlap = globMSH%overlap
    IPT%xMinGlob = globMSH%xMinGlob
    IPT%xMaxGlob = globMSH%xMaxGlob
    IPT%procExtent = globMSH%procExtent
    IPT%stepProc = stepProc
    nTot = product(IPT%nFields)
    IPT%nTotalFields = nTot

    ! build_random_field (calls_RF.f90:186-199)
    allocate(subdivisionCoords(3, nTot), xMinFiles(3, nTot), xMaxFiles(3, nTot), subdivisionId(3, nTot))
    call prepare_Localization(IPT, subdivisionCoords, subdivisionId, xMinFiles, xMaxFiles, xMin_Group, &
                              xMax_Group, gen_GroupRange, xNStep_Group, origin_Group, xNTotal_Group)
    xNStep_Proc = 1 + nint(IPT%procExtent / IPT%xStep)      ! find_nSamples_in_proc
    xNTotal_Proc = product(int(xNStep_Proc, 8))
    allocate(pu(xNTotal_Proc))
    call generateUnityPartition_Matrix(xNStep_Proc, IPT%overlap, IPT%corrL, IPT%xStep, 1, pu, 3)

    open(newunit=u, file=trim(outPath), access="stream", form="unformatted", status="replace")
    write(u) int(3, 4), int(IPT%nFields, 4), int(Nmc, 4)
    write(u) globMSH%xStep, globMSH%overlap, globMSH%xMinGlob, globMSH%xMaxGlob, globMSH%procExtent, &
             globMSH%procStart, stepProc
    write(u) int(xNStep_Proc, 4), int(xNStep_Group, 4), int(origin_Group, 4)
    write(u) subdivisionCoords, xMinFiles, xMin_Group, pu

    if (imode >= 2) then
        ! one field per subdivision (sample_RF.f90:280-348), on its own grid
        allocate(fields(xNTotal_Proc, Nmc, nTot), x(3, xNTotal_Proc), f(xNTotal_Proc, Nmc))
        do i = 1, nTot
            call init_RF(RDF, IPT, 0, 1, 0)
            call init_MESH(MSH, IPT, 0, 0, 1)
            RDF%compatMode = merge(RF_REF_COMPAT, RF_FIXED, imode == 2)
            MSH%xNStep = xNStep_Proc
            MSH%xStep = IPT%xStep
            MSH%xMinGlob = xMinFiles(:, i)
            RDF%xRange = IPT%procExtent
            do j = 1, int(xNTotal_Proc)
                x(1, j) = xMinFiles(1, i) + IPT%xStep(1) * modulo(j - 1, xNStep_Proc(1))
                x(2, j) = xMinFiles(2, i) + IPT%xStep(2) * modulo((j - 1) / xNStep_Proc(1), xNStep_Proc(2))
                x(3, j) = xMinFiles(3, i) + IPT%xStep(3) * ((j - 1) / (xNStep_Proc(1) * xNStep_Proc(2)))
            end do
            call calculate_random_seed(seed, IPT%seedStart + i)
            RDF%seed = seed
            RDF%xPoints => x
            RDF%randField => f
            call createRandomField(RDF, MSH)
            fields(:, :, i) = f
            call finalize_RF(RDF)
            call finalize_MESH(MSH)
        end do
        allocate(group(xNTotal_Group, Nmc))
        call localization_blend(3, Nmc, IPT%xStep, IPT%overlap, IPT%corrL, xNStep_Proc, xMinFiles, xMin_Group, &
                                xNStep_Group, fields, group, mode = merge(RF_REF_COMPAT, RF_FIXED, imode == 2))
        ! transform_and_write_output (sample_RF.f90:820-843) on one rank
        allocate(groupNorm(xNTotal_Group, Nmc))
        groupNorm = group
        nOvl = nint((IPT%overlap * IPT%corrL - 2 * IPT%xStep) / IPT%xStep) + 1
        minP = 1
        maxP = xNStep_Group - nOvl
        RF_3D(1:xNStep_Group(1), 1:xNStep_Group(2), 1:xNStep_Group(3)) => groupNorm(:, 1)
        call normalize_randField(groupNorm, minP, maxP, 3, Nmc, 0, RF_3D=RF_3D)
        call multiVariateTransformation(IPT%margiFirst, IPT%fieldAvg, IPT%fieldVar, groupNorm)
        write(u) fields, group, groupNorm, int(minP, 4), int(maxP, 4)
        do d = 1, 3
            if (maxP(d) < minP(d)) error stop 2
        end do
    end if
    close(u)
    call finalize_MESH(globMSH)
end program rf_localize
