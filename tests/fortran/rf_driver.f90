!> Fortran host of the GPU path, written the way a randomField host is:
!> init_RF / init_MESH, points and field storage owned by the caller,
!> RDF%seed from calculate_random_seed, then createRandomField(RDF, MSH).
!>
!> usage: rf_driver <case> <out.bin>
!>   case 1: Shinozuka, 3-D, REF_COMPAT, FP64 form, 500 points, Nmc 3, seed 1
!>   case 2: same points, FIXED, FP32 form
!>   case 3: randomization (box), 2-D, REF_COMPAT, FP64, 400 points, Nmc 2,
!>           calculate(2) = .false.
!>   case 4: isotropic, 3-D, FIXED, FP64, 300 points, Nmc 2, seed 7
!>   case 5: createStandardGaussianFieldUnstruct, lognormal, FIXED FP64
!>   case 6: method 4 (FFT), 3-D grid 12 x 10 x 8 at steps (0.25, 0.3, 0.4)
!>           from (1, -2, 0.5), corrL (1, 1.5, 0.8), Nmc 2, seed 5
!> out.bin (stream): int32 N, int32 Nmc, int32 nDim, int64 kNTotal,
!>   xPoints(nDim,N), randField(N,Nmc), kPoints(rows,kNTotal), SkVec(kNTotal),
!>   int32 seed(1)
program rf_driver
    use randomField_gpu
    implicit none
    character(len=256) :: arg, outPath
    integer :: icase, N, nDim, Nmc, i, u, rows
    double precision, allocatable, target :: x(:,:), field(:,:)
    type(RF) :: RDF
    type(MESH) :: MSH
    integer, allocatable :: seed(:)

    call get_command_argument(1, arg)
    read(arg, *) icase
    call get_command_argument(2, outPath)

    select case (icase)
        case (1, 2, 5); nDim = 3; N = 500; Nmc = 3
        case (3);       nDim = 2; N = 400; Nmc = 2
        case (6);       nDim = 3; N = 12 * 10 * 8; Nmc = 2
        case default;   nDim = 3; N = 300; Nmc = 2
    end select
    allocate(x(nDim, N), field(N, Nmc))
    ! deterministic points in [0,10)^nDim (a fixed low-discrepancy walk)
    do i = 1, N
        x(1, i) = 10d0 * modulo(0.7548776662466927d0 * i, 1d0)
        x(2, i) = 10d0 * modulo(0.5698402909980532d0 * i, 1d0)
        if (nDim == 3) x(3, i) = 10d0 * modulo(0.3819660112501051d0 * i, 1d0)
    end do

    if (icase == 6) then
        ! the grid points themselves, x fastest
        do i = 1, N
            x(1, i) = 1d0 + 0.25d0 * modulo(i - 1, 12)
            x(2, i) = -2d0 + 0.3d0 * modulo((i - 1) / 12, 10)
            x(3, i) = 0.5d0 + 0.4d0 * ((i - 1) / 120)
        end do
    end if

    if (icase == 5) then
        call createStandardGaussianFieldUnstruct(x, "gaussian", "lognormal", [1d0, 1.5d0, 0.8d0], &
                                                 1d0, 0.25d0, Nmc, SHINOZUKA, field, &
                                                 seedStart = 3, mode = RF_FIXED)
        open(newunit=u, file=trim(outPath), access="stream", form="unformatted", status="replace")
        write(u) int(N, 4), int(Nmc, 4), int(nDim, 4), int(0, 8), x, field
        close(u)
        stop
    end if

    select case (icase)
        case (1, 2)
            call init_RF(RDF, nDim, Nmc, SHINOZUKA, cm_GAUSSIAN, fom_GAUSSIAN, [1d0, 1d0, 1d0], 0d0, 1d0)
            call calculate_random_seed(seed, 1)
            RDF%compatMode = merge(RF_REF_COMPAT, RF_FIXED, icase == 1)
            RDF%form = merge(RF_FORM_FP64, RF_FORM_FP32, icase == 1)
        case (3)
            call init_RF(RDF, nDim, Nmc, RANDOMIZATION, cm_GAUSSIAN, fom_GAUSSIAN, [1d0, 2d0], 0d0, 1d0)
            call calculate_random_seed(seed, 11)
            RDF%calculate(2) = .false.
        case (6)
            call init_RF(RDF, nDim, Nmc, FFT, cm_GAUSSIAN, fom_GAUSSIAN, [1d0, 1.5d0, 0.8d0], 0d0, 1d0)
            call calculate_random_seed(seed, 5)
            RDF%compatMode = RF_FIXED
        case default
            call init_RF(RDF, nDim, Nmc, ISOTROPIC, cm_GAUSSIAN, fom_GAUSSIAN, [0.5d0, 0.5d0, 0.5d0], 0d0, 1d0)
            call calculate_random_seed(seed, 7)
            RDF%compatMode = RF_FIXED
    end select
    RDF%seed = seed
    call init_MESH(MSH, nDim)
    MSH%xNStep = 1
    MSH%xNStep(1) = N
    RDF%xRange = 10d0
    if (icase == 6) then
        MSH%xNStep = [12, 10, 8]
        MSH%xStep = [0.25d0, 0.3d0, 0.4d0]
        MSH%xMinGlob = [1d0, -2d0, 0.5d0]
        RDF%xRange = (MSH%xNStep - 1) * MSH%xStep
    end if
    RDF%xPoints => x
    RDF%randField => field
    call createRandomField(RDF, MSH)

    rows = size(RDF%kPoints, 1)
    open(newunit=u, file=trim(outPath), access="stream", form="unformatted", status="replace")
    write(u) int(N, 4), int(Nmc, 4), int(nDim, 4), int(RDF%kNTotal, 8), x, field, RDF%kPoints, RDF%SkVec, &
             int(RDF%seed(1), 4)
    close(u)
    call finalize_RF(RDF)
    call finalize_MESH(MSH)
end program rf_driver
