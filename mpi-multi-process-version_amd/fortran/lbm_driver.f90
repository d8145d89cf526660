! lbm_driver.f90 -- Fortran driver of the HIP lattice-Boltzmann solver.
!
! Plays the part of the reference's `program main`
! (natural-convection_2dim_mono.f90:283-345, natural-convection_3dim_mono.f90:347-412)
! with the compile-time `parameter`s of `commondata` turned into namelist
! inputs, and the solver behind the C ABI (module lbm_capi):
!   mesh + initial          -> lbm_create + lbm_init_rest
!   loop body               -> lbm_step
!   check every 2000 steps  -> lbm_check (same stopping rule and print line)
!   gather_data             -> lbm_get_fields
!   output_ASCII            -> write_tecplot (same file name, header and formats)
! plus two additions: the mean Nusselt numbers of the heated walls, and an
! optional FP64 binary dump of the fields.
!
! Usage:  lbm_driver [namelist-file]      (default file: lbm.nml, if present)
!
!   &lbm
!     dim = 2, nx = 129, Ra = 1e4, Pr = 0.71, Ma = 0.1, Thot = 1, Tcold = 0,
!     constA = 3.2, grid_kind = 0, precision = 0, epsU = 1e-6, epsT = 1e-6,
!     check_every = 2000, itc_max = 20000, device = 0,
!     output = 'tecplot', outdir = '.'
!   /
!
! switch_eps (default 0: off): with precision = 0, once both residuals are
! at or below switch_eps, the run continues in FP64 from the FP32 state
! (lbm_copy_state), so that it can reach residuals below the FP32 floor
! (about 1e-6); 1e-4 reaches the reference's 1e-8 in as many steps as FP64
! throughout.
!
! ny and nz default to nx (the reference's cavities are square/cubic).
! Unset nx, Ra and itc_max take the reference's values for the chosen
! dimension (2D: 513, 1e5, 20000; 3D: 257, 1e6, 2000).  precision is
! 0 (FP32 shifted populations) or 1 (FP64, the reference's arithmetic).
! output is one of 'tecplot', 'binary', 'both', 'none'.
!
! Binary dump (stream access, little-endian as the host writes it), file
! lbm_fields-<itc>.bin:
!   int32 dim, nx, ny, nz; int64 itc; float64 nu_hot, nu_cold;
!   float64 x(nx), y(ny), [z(nz) in 3D];
!   float64 rho, u, v, [w in 3D], T, each nx*ny(*nz) values, i fastest.
program lbm_driver
    use, intrinsic :: iso_c_binding
    use, intrinsic :: ieee_arithmetic, only: ieee_is_finite
    use lbm_capi
    implicit none

    integer :: dim = 2, nx = 0, ny = 0, nz = 0
    real(8) :: Ra = 0d0, Pr = 0.71d0, Ma = 0.1d0, Thot = 1d0, Tcold = 0d0, constA = 3.2d0
    integer :: grid_kind = 0, precision = 0
    real(8) :: epsU = 1d-6, epsT = 1d-6, switch_eps = 0d0
    integer :: check_every = 2000, device = 0
    integer(8) :: itc_max = 0
    character(len=16) :: output = 'tecplot'
    character(len=512) :: outdir = '.'
    namelist /lbm/ dim, nx, ny, nz, Ra, Pr, Ma, Thot, Tcold, constA, grid_kind, precision, &
        epsU, epsT, check_every, itc_max, device, output, outdir, switch_eps

    type(lbm_params) :: p
    type(c_ptr) :: h, h64
    real(c_double) :: info(LBM_INFO_COUNT)
    real(c_double), allocatable, target :: x(:), y(:), z(:)
    real(c_double), allocatable, target :: rho(:), u(:), v(:), w(:), t(:)
    real(c_double) :: errU, errT, nu_hot, nu_cold
    integer(c_int64_t) :: itc, nstep
    integer(8) :: clk0, clk1, rate
    real(8) :: elapsed, dx0, nodes
    character(len=512) :: nmlfile
    integer :: ios, unit
    logical :: exists

    ! ---- inputs ----------------------------------------------------------
    nmlfile = 'lbm.nml'
    if (command_argument_count() >= 1) call get_command_argument(1, nmlfile)
    inquire(file=trim(nmlfile), exist=exists)
    if (exists) then
        open(newunit=unit, file=trim(nmlfile), status='old', action='read')
        read(unit, nml=lbm, iostat=ios)
        close(unit)
        if (ios /= 0) then
            write(*,*) 'cannot read namelist &lbm from ', trim(nmlfile), ' (iostat ', ios, ')'
            error stop 2
        end if
    else if (command_argument_count() >= 1) then
        write(*,*) 'namelist file not found: ', trim(nmlfile)
        error stop 2
    end if
    if (dim /= 2 .and. dim /= 3) then
        write(*,*) 'dim must be 2 or 3'
        error stop 2
    end if
    if (nx == 0) nx = merge(513, 257, dim == 2)
    if (ny == 0) ny = nx
    if (dim == 3 .and. nz == 0) nz = nx
    if (dim == 2) nz = 1
    if (Ra == 0d0) Ra = merge(1d5, 1d6, dim == 2)
    if (itc_max == 0) itc_max = merge(20000_8, 2000_8, dim == 2)
    if (check_every <= 0) then
        write(*,*) 'check_every must be positive'
        error stop 2
    end if
    select case (trim(output))
    case ('tecplot', 'binary', 'both', 'none')
    case default
        write(*,*) "output must be 'tecplot', 'binary', 'both' or 'none'"
        error stop 2
    end select

    p = lbm_params(dim=dim, nx=nx, ny=ny, nz=nz, Ra=Ra, Pr=Pr, Ma=Ma, Thot=Thot, Tcold=Tcold, &
                   stretch=constA, grid_kind=grid_kind, precision=precision, epsU=epsU, &
                   epsT=epsT, check_every=check_every, device=device, itc_max=itc_max)

    ! ---- mesh + initial (the reference's timer starts before mesh) -------
    call system_clock(clk0, rate)
    call lbm_require(lbm_create(p, h), 'lbm_create')
    allocate(x(nx), y(ny), z(nz))
    call lbm_require(lbm_get_grid(h, c_loc(x), c_loc(y), merge(c_loc(z), c_null_ptr, dim == 3)), &
                     'lbm_get_grid')
    call lbm_require(lbm_get_info(h, info), 'lbm_get_info')
    dx0 = 1d0 / info(LBM_INFO_LLB)
    call print_parameters()
    call lbm_require(lbm_init_rest(h), 'lbm_init_rest')

    ! ---- time loop: the reference's stopping rule ------------------------
    errU = 100d0
    errT = 100d0
    itc = 0
    do while ((errU > epsU .or. errT > epsT) .and. itc < itc_max)
        nstep = min(int(check_every, c_int64_t) - mod(itc, int(check_every, c_int64_t)), itc_max - itc)
        call lbm_require(lbm_step(h, nstep), 'lbm_step')
        itc = itc + nstep
        if (mod(itc, int(check_every, c_int64_t)) == 0) then
            call lbm_require(lbm_check(h, errU, errT), 'lbm_check')
            write(*,*) int(itc, 4), ' ', errU, ' ', errT
            if (.not. (ieee_is_finite(errU) .and. ieee_is_finite(errT))) then
                write(*,*) 'diverged: non-finite residual at itc =', int(itc, 4)
                error stop 3
            end if
            if (p%precision == 0 .and. switch_eps > 0d0 .and. errU <= switch_eps .and. &
                errT <= switch_eps .and. (errU > epsU .or. errT > epsT)) then
                ! FP32 has done the transient: FP64 from its state onwards
                p%precision = 1
                call lbm_require(lbm_create(p, h64), 'lbm_create (FP64)')
                call lbm_require(lbm_copy_state(h64, h), 'lbm_copy_state')
                call lbm_require(lbm_destroy(h), 'lbm_destroy')
                h = h64
                write(*,*) 'switched to FP64 at itc =', int(itc, 4)
            end if
        end if
    end do
    call lbm_require(lbm_synchronize(h), 'lbm_synchronize')
    call system_clock(clk1)
    elapsed = dble(clk1 - clk0) / dble(rate)
    nodes = dble(nx) * dble(ny) * dble(nz)
    if (dim == 2) then
        write(*,*) "Time=", elapsed, "s"
    else
        write(*,*) "Time (CPU) =", real(elapsed), "s"
    end if
    write(*,*) "MLUPS = ", real(nodes / 1e6 * dble(itc) / elapsed)

    ! ---- gather_data + outputs -------------------------------------------
    allocate(rho(nx*ny*nz), u(nx*ny*nz), v(nx*ny*nz), w(nx*ny*nz), t(nx*ny*nz))
    call lbm_require(lbm_get_fields(h, c_loc(rho), c_loc(u), c_loc(v), c_loc(w), c_loc(t)), &
                     'lbm_get_fields')
    call lbm_require(lbm_nusselt(h, nu_hot, nu_cold), 'lbm_nusselt')
    write(*,*) "Nu (hot wall) =", nu_hot, ", Nu (cold wall) =", nu_cold
    if (trim(output) == 'tecplot' .or. trim(output) == 'both') call write_tecplot()
    if (trim(output) == 'binary' .or. trim(output) == 'both') call write_binary()
    call lbm_require(lbm_destroy(h), 'lbm_destroy')
    write(*,*) "Successfully: DNS completed!"

contains

    ! The parameter printout of mesh and initial
    ! (natural-convection_2dim_mono.f90:96-106, :166-173;
    !  natural-convection_3dim_mono.f90:149-158, :214-221).
    subroutine print_parameters()
        real(8) :: length_LB
        length_LB = info(LBM_INFO_LLB)
        if (dim == 2) then
            write(*,*) "nx =", nx, ", ny =", ny
        else
            write(*,*) "nx =", nx, ", ny =", ny, ", nz =", nz
        end if
        write(*,*) "Ra =", real(Ra)
        write(*,*) "  "
        write(*,*) "---in 0 system---"
        write(*,*) "deltaX = ", dx0
        write(*,*) "deltaT = ", dx0
        write(*,*) "length0 = ", real(1d0 - dx0)
        write(*,*) "  "
        if (dim == 2) then
            write(*,*) "---in system LB"
            write(*,*) "deltaX = ", x(1)
            write(*,*) "deltaT = ", info(LBM_INFO_DT)
        else
            write(*,*) "---in LB unit---"
        end if
        write(*,*) "characteristic length   =", real(length_LB), "l.u."
        write(*,*) "viscosity_LB =", real(info(LBM_INFO_NU)), "l.u.^2/t.s."
        write(*,*) "timeStep ratio for (uniform) / (non-uniform) : ", real(length_LB / dble(nx))
        if (dim == 2) then
            write(*,*) "tauf =", real(info(LBM_INFO_TAUF))
            write(*,*) "    "
        else
            write(*,*) "    "
            write(*,*) "tauf =", real(info(LBM_INFO_TAUF))
        end if
    end subroutine print_parameters

    function path(name) result(s)
        character(len=*), intent(in) :: name
        character(len=:), allocatable :: s
        s = trim(outdir) // '/' // name
    end function path

    ! output_ASCII (natural-convection_2dim_mono.f90:818-843;
    ! natural-convection_3dim_mono.f90:1097-1121): same file names, headers,
    ! POINT order and edit descriptors.
    subroutine write_tecplot()
        character(len=100) :: tag
        integer :: i, j, k, n, iu
        if (dim == 2) then
            write(tag,*) int(itc, 4)
            tag = adjustl(tag)
            open(newunit=iu, file=path('MRTcavity-'//trim(tag)//'.dat'), status='unknown')
            write(iu,*) 'TITLE="thermal convective flows"'
            write(iu,*) 'VARIABLES="X" "Y" "U" "V" "T" "rho"'
            write(iu,101) nx, ny
            do j = 1, ny
                do i = 1, nx
                    n = i + nx*(j-1)
                    write(iu,100) x(i), y(j), u(n), v(n), t(n), rho(n)
                end do
            end do
        else
            write(tag,*) int(nx)
            tag = adjustl(tag)
            open(newunit=iu, file=path("side-heated convection"//trim(tag)//'.dat'), status='unknown')
            write(iu,*) 'TITLE="side-heated convection "'
            write(iu,*) 'VARIABLES="X" "Y" "Z" "U" "V" "W" "RHO" "T" '
            write(iu,102) nx, ny, nz
            do k = 1, nz
                do j = 1, ny
                    do i = 1, nx
                        n = i + nx*(j-1) + nx*ny*(k-1)
                        write(iu,100) x(i), y(j), z(k), u(n), v(n), w(n), rho(n), t(n)
                    end do
                end do
            end do
        end if
        close(iu)
100     format(1x,2(e11.4,' '),10(e13.6,' '))
101     format('ZONE',1x,'I=',1x,i5,2x,'J=',1x,i5,1x,'F=POINT')
102     format('ZONE',1x,'I=',1x,i5,2x,'J=',1x,i5,1x,'K=',1x,i5,1x,'F=POINT')
    end subroutine write_tecplot

    subroutine write_binary()
        character(len=100) :: tag
        integer :: iu
        write(tag,*) int(itc, 4)
        tag = adjustl(tag)
        open(newunit=iu, file=path('lbm_fields-'//trim(tag)//'.bin'), access='stream', &
             form='unformatted', status='replace')
        write(iu) int(dim, 4), int(nx, 4), int(ny, 4), int(nz, 4), int(itc, 8), nu_hot, nu_cold
        write(iu) x, y
        if (dim == 3) write(iu) z
        write(iu) rho, u, v
        if (dim == 3) write(iu) w
        write(iu) t
        close(iu)
    end subroutine write_binary

end program lbm_driver
