!> Host side of the device path in Fortran: the INI registry of the hot path's
!> keys, the readers of ICE's input files (Tecplot BLOCK initial condition,
!> ATLAS bc.txt, constant property table), the block metrics, and the Tecplot
!> solution writer.  The block and boundary data are held in the layouts of
!> ICE's derived types (src/lib/base/Advanced_Types_m.f90:13-153), so that the
!> arrays go to the C-ABI (ice_gpu_m) as they are.
module ice_host_m
  use, intrinsic :: iso_fortran_env, only: I4 => int32, R8 => real64, error_unit
  use, intrinsic :: ieee_arithmetic, only: ieee_value, ieee_quiet_nan
  use ice_gpu_m
  implicit none
  private

  integer, parameter, public :: GC = 2

  !> One block: its cells, nodes, metrics and the primitive state of its family,
  !> prim(nc, 1-gc:ni+gc, 1-gc:nj+gc, 1-gc:nk+gc) (Mod_Allocate_Data.f90:216-238)
  type, public :: host_block_t
    integer(I4) :: dim(3) = 0
    real(R8), allocatable :: node(:,:,:,:)          ! (3, 0:ni, 0:nj, 0:nk)
    real(R8), allocatable :: prim(:,:,:,:)          ! (nc, -1:ni+2, ...)
    real(R8), allocatable :: vol(:,:,:), dl(:,:,:,:), m(:,:,:,:,:)
    real(R8), allocatable :: fi(:,:,:,:), fj(:,:,:,:), fk(:,:,:,:)
  end type

  !> The registry: every key the path reads, with its section and default
  !> (Register_Numerics.f90:28-80, Register_Physics.f90, Read_Ini.f90)
  !> ([ICE-Multigrid] levels and level<m>-iter, Register_Numerics.f90:90-115: up to
  !> MAXLEV levels here)
  integer, parameter, public :: MAXLEV = 4
  integer, parameter :: NKEY = 43
  character(len=24), parameter :: key_sec(NKEY) = [character(len=24) :: &
    'ICE-Parameters', 'ICE-Numerics', 'ICE-Numerics', 'ICE-Numerics', 'ICE-Numerics', 'ICE-Numerics', &
    'ICE-Numerics', 'ICE-Numerics', 'ICE-Numerics', 'ICE-Numerics', 'ICE-Numerics', 'ICE-Numerics', &
    'ICE-Numerics', 'ICE-Physics', 'ICE-Physics', 'ICE-IO', 'ICE-IO', 'ICE-IO', 'ICE-Device', 'ICE-Family', &
    'ICE-Device', 'ICE-Parameters', 'ICE-Multigrid', 'ICE-Multigrid', 'ICE-Multigrid', 'ICE-Multigrid', &
    'ICE-Multigrid', &
    'ICE-Physics', 'ICE-Physics', 'ICE-Physics', 'ICE-Physics', 'ICE-Physics', 'ICE-Physics', 'ICE-Physics', &
    'ICE-Physics', 'ICE-Physics', 'ICE-Physics', 'ICE-Physics', 'ICE-Physics', 'ICE-Physics', 'ICE-Physics', &
    'ICE-Physics', 'ICE-IO']
  character(len=24), parameter :: key_name(NKEY) = [character(len=24) :: &
    'iter-threshold', 'time-scheme', 'cfl', 'dt-max', 'tau-factor', 'cfl-rise-threshold', 'time-accurate', &
    'irs', 'irs-beta', 'space-reconstruction', 'flux-limiter', 'shock-detector', 'riemann-solver', &
    'drag', 'heat-transfer', 'ic-format', 'sol-format', 'shell-diter', 'mode', 'closure', 'autotune', &
    'res-threshold', 'levels', 'level1-iter', 'level2-iter', 'level3-iter', 'level4-iter', &
    'evaporation', 'evaporation-interface', 'evaporation-blowing', 'density', 'specific-heat', 'latent-heat', &
    'emissivity', 'vapour-molar-mass', 'boiling-temperature', 'Tboil', 'boiling-pressure', 'vapour-specific-heat', &
    'lewis-number', 'vapour-mass-fraction', 'evaporation-coefficient', 'gas-path']
  !> (Register_Physics.f90:26-126, Register_IO.f90:46-48)
  character(len=24), parameter :: key_default(NKEY) = [character(len=24) :: &
    '1000000000', 'RK2', '0.5', '1e-4', '1.0', '0', '.true.', '.false.', '0.5', 'first-order', 'none', 'none', &
    '', 'none', 'none', 'tecplot ascii', 'tecplot ascii', '10', 'plain', 'IG', '.false.', &
    '0.0', '1', '1000000000', '1000000000', '1000000000', '1000000000', &
    'none', 'VLE', 'none', '2700.0', '1598.0', '1.08e7', '0.0', '26.98', '2792.0', '2792.0', '101325.0', '0.0', &
    '1.0', '0.0', '1.0', 'INPUT/']

  type, public :: registry_t
    character(len=64) :: value(NKEY)
    logical :: set(NKEY) = .false.      !< given in input.ini
    integer :: nfam = 0
    character(len=8) :: closure(14) = 'IG'
  end type

  public :: read_registry, reg_str, reg_real, reg_int, reg_bool, reg_set, config_from_registry
  public :: read_ic, read_bc, read_material, block_metrics, write_solution, refuse, coarse_blocks

contains

  subroutine refuse(routine, msg)
    character(len=*), intent(in) :: routine, msg
    write(error_unit, '(A)') ' [ERROR] [ICE::'//trim(routine)//'] '//trim(msg)
    error stop 1
  end subroutine

  ! ------------------------------------------------------------------ INI
  !> Reads input.ini; an unknown section or key stops (Backend_INI.f90:196-288).
  subroutine read_registry(path, reg)
    character(len=*), intent(in)  :: path
    type(registry_t), intent(out) :: reg
    character(len=512) :: line, sec, key, val
    integer :: u, ios, q, n, eq, nline, p
    logical :: found
    reg%value = key_default
    sec = ''
    nline = 0
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call refuse('Read_Inifile', 'cannot open '//trim(path))
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      nline = nline + 1
      q = scan(line, ';#')
      if (q > 0) line = line(:q-1)
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      if (line(1:1) == '[') then
        sec = line(2:index(line, ']') - 1)
        if (sec(1:10) == 'ICE-Family') then
          read(sec(11:), *, iostat=ios) p
          if (ios /= 0 .or. p < 1 .or. p > 14) call refuse('Read_Inifile', 'bad section ['//trim(sec)//']')
          reg%nfam = max(reg%nfam, p)
        end if
        cycle
      end if
      eq = index(line, '=')
      if (eq == 0) call refuse('Read_Inifile', trim(path)//': a line that is neither [section] nor key = value')
      key = adjustl(line(:eq-1))
      val = adjustl(line(eq+1:))
      found = .false.
      if (sec(1:10) == 'ICE-Family' .and. trim(key) == 'closure') then
        read(sec(11:), *) p
        reg%closure(p) = trim(val)
        found = .true.
      else
        do n = 1, NKEY
          if (trim(key_sec(n)) == trim(sec) .and. trim(key_name(n)) == trim(key)) then
            reg%value(n) = trim(val)
            reg%set(n) = .true.
            found = .true.
          end if
        end do
      end if
      if (.not. found) call refuse('Read_Inifile', trim(path)//': unknown key "'//trim(key)//'" in ['//trim(sec)//']')
    end do
    close(u)
    if (reg%nfam == 0) call refuse('Read_Inifile', 'no [ICE-Family1] section')
  end subroutine

  function reg_str(reg, key) result(s)
    type(registry_t), intent(in) :: reg
    character(len=*), intent(in) :: key
    character(len=64) :: s
    integer :: n
    do n = 1, NKEY
      if (trim(key_name(n)) == key) then
        s = reg%value(n)
        return
      end if
    end do
    call refuse('registry', 'no key '//key)
  end function

  real(R8) function reg_real(reg, key)
    type(registry_t), intent(in) :: reg
    character(len=*), intent(in) :: key
    integer :: ios
    read(reg_str(reg, key), *, iostat=ios) reg_real
    if (ios /= 0) call refuse('Read_Inifile', key//' is not a number')
  end function

  integer function reg_int(reg, key)
    type(registry_t), intent(in) :: reg
    character(len=*), intent(in) :: key
    integer :: ios
    read(reg_str(reg, key), *, iostat=ios) reg_int
    if (ios /= 0) call refuse('Read_Inifile', key//' is not an integer')
  end function

  logical function reg_set(reg, key)
    type(registry_t), intent(in) :: reg
    character(len=*), intent(in) :: key
    integer :: n
    reg_set = .false.
    do n = 1, NKEY
      if (trim(key_name(n)) == key) reg_set = reg_set .or. reg%set(n)
    end do
  end function

  logical function reg_bool(reg, key)
    type(registry_t), intent(in) :: reg
    character(len=*), intent(in) :: key
    character(len=64) :: v
    v = reg_str(reg, key)
    reg_bool = any(trim(v) == [character(len=8) :: '.true.', 'true', 'T', '1', 'yes', '.TRUE.'])
  end function

  !> ice_config_t and the families from the registry (Assign_Setup.f90:52-117):
  !> coupled when the gas field is present (owcoupled, Assign_Setup.f90:50-54),
  !> with the exchange models of [ICE-Physics] and the material and vapour keys
  !> packed as Setup_Materials does (Setup_Materials.f90:140-170).
  subroutine config_from_registry(reg, ndir, rho_mat, cs, coupled, cfg, fams)
    type(registry_t), intent(in) :: reg
    integer, intent(in) :: ndir
    real(R8), intent(in) :: rho_mat, cs
    logical, intent(in) :: coupled
    type(ice_config_t), intent(out) :: cfg
    type(ice_family_t), allocatable, intent(out) :: fams(:)
    character(len=64) :: w
    integer :: p, evap, intf, blow
    real(R8) :: ep(8), tboil, pboil
    real(R8), parameter :: RU = 8314.46_R8, PATM = 101325._R8   ! Lib_Evaporation.f90:22, Parameters_m
    character(len=10), parameter :: lims(0:10) = [character(len=10) :: 'none', 'minmod', 'vanalbada', &
      'vanleer', 'ospre', 'umist', 'osher', 'sweby', 'mc', 'koren', 'superbee']
    character(len=16), parameter :: drags(13) = [character(len=16) :: 'Newton', 'Stokes', 'Schlichting', &
      'Schiller-Naumann', 'Wen-Yu', 'Putnam', 'Clift-Gauvin', 'Morsi-Alexander', 'Carlson-Hoglund', 'Henderson', &
      'Crowe', 'Hermsen', 'NoDrag']
    character(len=16), parameter :: heats(8) = [character(len=16) :: 'Stokes', 'JAXA1', 'JAXA2', 'JAXA3', &
      'JAXA4', 'Ranz-Marshall', 'Kavanau-Drake', 'NoHeat']
    character(len=8), parameter :: evaps(5) = [character(len=8) :: 'd2-law', 'CEM', 'CEM-B', 'ASM', 'TC']
    cfg%coupled = merge(1, 0, coupled)
    evap = 0; intf = 0; blow = 0
    ep = 0._R8
    if (coupled) then
      ! assign_drag / assign_heat (Lib_Drag.f90:17-80, Lib_Heat.f90:18-62)
      w = reg_str(reg, 'drag')
      if (trim(w) == 'none') call refuse('assign_drag', 'drag is not set: [ICE-Physics] drag is required for a '// &
                                         'coupled run')
      cfg%drag_select = 0
      do p = 1, 13
        if (trim(w) == trim(drags(p))) cfg%drag_select = p
      end do
      if (cfg%drag_select == 0) call refuse('assign_drag', 'Wrong drag input ---> '//trim(w))
      w = reg_str(reg, 'heat-transfer')
      if (trim(w) == 'none') call refuse('assign_heat', 'heat-transfer is not set: [ICE-Physics] heat-transfer is '// &
                                         'required for a coupled run')
      cfg%heat_select = 0
      do p = 1, 8
        if (trim(w) == trim(heats(p))) cfg%heat_select = p
      end do
      if (cfg%heat_select == 0) call refuse('assign_heat', 'Wrong heat input ---> '//trim(w))
      w = reg_str(reg, 'evaporation')
      if (trim(w) /= 'none') then
        do p = 1, 5
          if (trim(w) == trim(evaps(p))) evap = p
        end do
        if (evap == 0) call refuse('assign_evaporation', 'unknown evaporation model '//trim(w))
      end if
      w = reg_str(reg, 'evaporation-interface')
      if (trim(w) == 'LK') then
        intf = 1
      else if (trim(w) /= 'VLE') then
        call refuse('assign_interface', 'evaporation-interface is VLE or LK, not '//trim(w))
      end if
      w = reg_str(reg, 'evaporation-blowing')
      if (trim(w) == 'LK') then
        blow = 1
      else if (trim(w) /= 'none') then
        call refuse('assign_blowing', 'evaporation-blowing is LK or none, not '//trim(w))
      end if
      if (evap == 0) then
        intf = 0
        blow = 0
      end if
      if (evap > 0) then
        ! Setup_Materials.f90:140-170: the keys an evaporating material needs, then ep
        if (.not. reg_set(reg, 'vapour-specific-heat')) call refuse('Setup_Materials', &
          'evaporation requires vapour-specific-heat (0 for the gas cp): give it in [ICE-Physics]')
        if (.not. (reg_set(reg, 'boiling-temperature') .or. reg_set(reg, 'Tboil'))) call refuse('Setup_Materials', &
          'evaporation requires boiling-temperature (alias Tboil) in [ICE-Physics]')
        if (reg_set(reg, 'boiling-temperature') .and. reg_set(reg, 'Tboil')) call refuse('Setup_Materials', &
          'boiling-temperature and its alias Tboil are both given: give one of them')
        if (.not. reg_set(reg, 'lewis-number')) call refuse('Setup_Materials', &
          'evaporation requires lewis-number in [ICE-Physics]')
        if (.not. reg_set(reg, 'vapour-mass-fraction')) call refuse('Setup_Materials', &
          'evaporation requires vapour-mass-fraction in [ICE-Physics]')
        if (.not. reg_real(reg, 'vapour-molar-mass') > 0._R8) call refuse('Setup_Materials', &
          'vapour-molar-mass = 0 needs an Mv column in the property table')
        tboil = reg_real(reg, 'boiling-temperature')
        if (reg_set(reg, 'Tboil')) tboil = reg_real(reg, 'Tboil')
        pboil = reg_real(reg, 'boiling-pressure')
        ep(1) = reg_real(reg, 'vapour-molar-mass')
        ep(2) = reg_real(reg, 'latent-heat')
        ep(3) = reg_real(reg, 'vapour-specific-heat')
        ep(4) = reg_real(reg, 'lewis-number')
        ep(5) = reg_real(reg, 'vapour-mass-fraction')
        ep(6) = ep(2) * ep(1) / RU
        ep(7) = 1._R8 / tboil
        ep(8) = reg_real(reg, 'evaporation-coefficient')
        ! boiling-temperature given at boiling-pressure: the line anchored at one atmosphere (invTboil_at)
        if (pboil /= PATM) then
          ep(7) = 1._R8 / tboil + log(pboil / PATM) / ep(6)
          if (.not. ep(7) > 0._R8) call refuse('Setup_Materials', 'boiling-pressure and boiling-temperature give '// &
            'no one-atmosphere boiling point on the Clausius-Clapeyron line')
        end if
      end if
    end if
    cfg%blow_select = blow
    w = reg_str(reg, 'time-scheme')
    select case (trim(w))
    case ('euler'); cfg%nrk = 1
    case ('RK2');   cfg%nrk = 2
    case ('RK3');   cfg%nrk = 3
    case default;   call refuse('Assign_Setup', 'time-scheme '//trim(w)//' is not euler, RK2 or RK3')
    end select
    cfg%ngroups = reg%nfam
    cfg%cfl = reg_real(reg, 'cfl')
    cfg%dt_max = reg_real(reg, 'dt-max')
    cfg%tau_factor = reg_real(reg, 'tau-factor')
    cfg%cfl_ramp_iter = reg_int(reg, 'cfl-rise-threshold')
    cfg%time_accurate = merge(1, 0, reg_bool(reg, 'time-accurate'))
    cfg%irs = merge(1, 0, reg_bool(reg, 'irs'))
    if (cfg%irs == 1 .and. cfg%time_accurate == 1) &
      call refuse('Assign_Setup', 'implicit residual smoothing needs time-accurate = false')
    cfg%irs_beta = reg_real(reg, 'irs-beta')
    cfg%ndir = ndir
    cfg%limiter = 0
    if (trim(reg_str(reg, 'space-reconstruction')) == 'MUSCL') then
      w = reg_str(reg, 'flux-limiter')
      cfg%limiter = -1
      do p = 0, 10
        if (trim(w) == trim(lims(p))) cfg%limiter = p
      end do
      if (cfg%limiter < 0) call refuse('Assign_Setup', 'unknown flux-limiter '//trim(w))
    end if
    cfg%shock_detector = merge(1, 0, trim(reg_str(reg, 'shock-detector')) == 'Jameson')
    cfg%mode = merge(ICE_MODE_BATCHED, ICE_MODE_PLAIN, trim(reg_str(reg, 'mode')) == 'batched')
    allocate(fams(reg%nfam))
    do p = 1, reg%nfam
      select case (trim(reg%closure(p)))
      case ('MK'); fams(p)%closure = ICE_MK; fams(p)%riemann = ICE_SAUREL
      case ('IG'); fams(p)%closure = ICE_IG; fams(p)%riemann = ICE_RUSANOV
      case ('AG'); fams(p)%closure = ICE_AG; fams(p)%riemann = ICE_RUSANOV
      case default; call refuse('Assign_Setup', 'closure '//trim(reg%closure(p))//' is not MK, IG or AG')
      end select
      w = reg_str(reg, 'riemann-solver')
      if (trim(w) == 'Saurel') fams(p)%riemann = ICE_SAUREL
      if (trim(w) == 'Rusanov') fams(p)%riemann = ICE_RUSANOV
      if (trim(w) == 'HLLE') fams(p)%riemann = ICE_HLLE
      fams(p)%rho_mat = rho_mat
      fams(p)%cs = cs
      fams(p)%lv = reg_real(reg, 'latent-heat')
      fams(p)%emiss = min(reg_real(reg, 'emissivity'), 1._R8)
      fams(p)%evap_select = evap
      fams(p)%intf_select = intf
      fams(p)%ep = ep
    end do
  end subroutine

  ! ------------------------------------------------------------------ Tecplot
  integer function zone_key(line, key)
    character(len=*), intent(in) :: line, key
    integer :: q, r, ios
    zone_key = -1
    q = index(line, ' '//key//'=')
    if (q == 0) q = index(line, ','//key//'=')
    if (q == 0) return
    r = q + len(key) + 2
    read(line(r:), *, iostat=ios) zone_key
  end function

  !> The initial condition: one zone per block, nodal x y z and nc cell-centred
  !> primitive variables, BLOCK packing (read_ic, IO_Solution.f90:36-38).
  subroutine read_ic(path, nc, blocks)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nc
    type(host_block_t), allocatable, intent(out) :: blocks(:)
    character(len=4096) :: line
    character(len=64) :: up
    integer :: u, ios, nz, b, nn, ncell, nv, i0, i1, n, ni, nj, nk, i, j, k, v
    integer :: dims(3, 64)
    real(R8), allocatable :: vals(:)
    type(host_block_t), allocatable :: tmp(:)
    real(R8) :: nan
    nan = ieee_value(nan, ieee_quiet_nan)
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call refuse('read_ic', 'cannot open '//trim(path))
    ! pass 1: zone sizes
    nz = 0
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      up = adjustl(line(1:64))
      if (up(1:4) == 'ZONE') then
        nz = nz + 1
        if (nz > 64) call refuse('read_ic', 'more than 64 zones')
        dims(1, nz) = zone_key(line, 'I')
        dims(2, nz) = zone_key(line, 'J')
        dims(3, nz) = zone_key(line, 'K')
        if (any(dims(:, nz) < 1)) call refuse('read_ic', 'a zone header without I, J, K: '//trim(line(1:80)))
      end if
    end do
    if (nz == 0) call refuse('read_ic', trim(path)//' holds no zone')
    rewind(u)
    allocate(tmp(nz))
    b = 0
    nv = 0
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      up = adjustl(line(1:64))
      if (up(1:4) == 'ZONE') then
        b = b + 1
        nn = product(dims(:, b))
        ncell = (dims(1, b) - 1) * (dims(2, b) - 1) * max(dims(3, b) - 1, 1)
        if (allocated(vals)) deallocate(vals)
        allocate(vals(3 * nn + nc * ncell))
        nv = 0
        cycle
      end if
      if (b == 0) cycle
      if (up(1:11) == 'DATASETAUXD' .or. up(1:7) == 'AUXDATA') cycle
      ! the numbers of the line, blank separated
      i0 = 1
      do while (i0 <= len_trim(line))
        do while (i0 <= len_trim(line) .and. line(i0:i0) == ' ')
          i0 = i0 + 1
        end do
        if (i0 > len_trim(line)) exit
        i1 = index(line(i0:), ' ')
        if (i1 == 0) then
          i1 = len_trim(line)
        else
          i1 = i0 + i1 - 2
        end if
        nv = nv + 1
        if (nv > size(vals)) call refuse('read_ic', 'zone '//itoa(b)//' holds more values than '// &
                                         '3 nodal and '//itoa(nc)//' cell variables')
        read(line(i0:i1), *, iostat=ios) vals(nv)
        if (ios /= 0) call refuse('read_ic', 'not a number: '//line(i0:i1))
        i0 = i1 + 1
      end do
      if (nv == size(vals)) call store(b)
    end do
    close(u)
    allocate(blocks(nz))
    do b = 1, nz
      if (.not. allocated(tmp(b)%prim)) call refuse('read_ic', 'zone '//itoa(b)//' is cut short')
      call move_alloc(tmp(b)%node, blocks(b)%node)
      call move_alloc(tmp(b)%prim, blocks(b)%prim)
      blocks(b)%dim = tmp(b)%dim
    end do

  contains

    subroutine store(b)
      integer, intent(in) :: b
      ni = dims(1, b) - 1; nj = dims(2, b) - 1; nk = max(dims(3, b) - 1, 1)
      if (dims(3, b) == 1) call refuse('read_ic', 'a zone with K = 1: ICE blocks have one cell at least in k')
      tmp(b)%dim = [ni, nj, nk]
      allocate(tmp(b)%node(3, 0:ni, 0:nj, 0:nk))
      n = 0
      do v = 1, 3
        do k = 0, nk
          do j = 0, nj
            do i = 0, ni
              n = n + 1
              tmp(b)%node(v, i, j, k) = vals(n)
            end do
          end do
        end do
      end do
      allocate(tmp(b)%prim(nc, 1-GC:ni+GC, 1-GC:nj+GC, 1-GC:nk+GC))
      tmp(b)%prim = nan
      do v = 1, nc
        do k = 1, nk
          do j = 1, nj
            do i = 1, ni
              n = n + 1
              tmp(b)%prim(v, i, j, k) = vals(n)
            end do
          end do
        end do
      end do
    end subroutine
  end subroutine

  !> A constant property table (Temperature, Cp, Density, Enthalpy rows); the
  !> device path holds one density and heat capacity per material.
  !> The property table INPUT/part-properties.dat (Load_Table.f90:57-147): the
  !> VARIABLES header names the columns (Temperature first, then Cp, Density and
  !> Enthalpy or Enthalpy_abs in any order, other columns ignored), one POINT row per
  !> integer kelvin.  Returns the first row's density and cp, and the device columns
  !> of ice_gpu_set_properties: density, cp, e = h - hOff (table_h_offset, :535-543),
  !> with the rho_varies / cs_varies flags.
  subroutine read_material(path, rho_mat, cs, t_min, cols, flags)
    character(len=*), intent(in) :: path
    real(R8), intent(out) :: rho_mat, cs
    integer, intent(out) :: t_min, flags
    real(R8), allocatable, intent(out) :: cols(:,:)
    character(len=1024) :: line
    character(len=64) :: names(32)
    real(R8), allocatable :: rows(:,:), tmp(:,:)
    real(R8) :: h_off, T1
    integer :: u, ios, n, nv, it, icp, irho, ih, m
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call refuse('Load_Table', 'cannot open '//trim(path))
    nv = 0
    n = 0
    allocate(rows(32, 64))
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      if (index(line, 'VARIABLES') > 0) then
        call header_names(line, names, nv)
        cycle
      end if
      if (nv == 0) cycle
      read(line, *, iostat=ios) rows(1:nv, min(n + 1, size(rows, 2)))
      if (ios /= 0) cycle
      n = n + 1
      if (n == size(rows, 2)) then
        allocate(tmp(32, 2 * n))
        tmp(:, 1:n) = rows
        call move_alloc(tmp, rows)
      end if
    end do
    close(u)
    if (nv == 0) call refuse('Load_Table', trim(path)//': no VARIABLES header')
    if (n < 2) call refuse('Load_Table', trim(path)//': a property table needs at least two rows')
    it = column('Temperature'); icp = column('Cp'); irho = column('Density'); ih = column('Enthalpy')
    if (ih == 0) ih = column('Enthalpy_abs')
    if (it /= 1) call refuse('Load_Table', trim(path)//': Temperature must be the first column')
    if (icp == 0 .or. irho == 0 .or. ih == 0) call refuse('Load_Table', trim(path)//': Cp, Density and Enthalpy needed')
    t_min = nint(rows(1, 1))
    if (t_min < 0) call refuse('Load_Table', trim(path)//': a temperature below 0 K')
    do m = 1, n
      if (abs(rows(1, m) - real(t_min + m - 1, R8)) > 1.e-6_R8) &
        call refuse('Load_Table', trim(path)//': rows not on consecutive integer kelvins')
      if (rows(irho, m) <= 0._R8 .or. rows(icp, m) <= 0._R8) &
        call refuse('Load_Table', trim(path)//': a density or cp that is not positive')
      if (m > 1) then
        if (rows(ih, m) <= rows(ih, m - 1)) call refuse('Load_Table', trim(path)//': an enthalpy that does not increase')
      end if
    end do
    T1 = rows(1, 1)
    if (all(rows(icp, 1:n) == rows(icp, 1))) then
      h_off = rows(ih, 1) - rows(icp, 1) * T1
    else
      h_off = rows(ih, 1) - T1 * (rows(ih, 2) - rows(ih, 1))
    end if
    allocate(cols(n, 0:ICE_PROP_NCOL - 1))
    cols = 0._R8
    cols(:, ICE_PROP_RHO) = rows(irho, 1:n)
    cols(:, ICE_PROP_CP) = rows(icp, 1:n)
    cols(:, ICE_PROP_E) = rows(ih, 1:n) - h_off
    flags = 0
    if (any(rows(irho, 1:n) /= rows(irho, 1))) flags = ior(flags, ICE_PROP_RHO_VARIES)
    if (any(rows(icp, 1:n) /= rows(icp, 1))) flags = ior(flags, ICE_PROP_CS_VARIES)
    rho_mat = rows(irho, 1)
    cs = rows(icp, 1)
  contains
    integer function column(name)
      character(len=*), intent(in) :: name
      integer :: k
      column = 0
      do k = 1, nv
        if (trim(names(k)) == name) column = k
      end do
    end function
  end subroutine

  !> The quoted names of a VARIABLES line
  subroutine header_names(line, names, nv)
    character(len=*), intent(in) :: line
    character(len=64), intent(out) :: names(:)
    integer, intent(out) :: nv
    integer :: a, b
    nv = 0
    a = index(line, '"')
    do while (a > 0 .and. nv < size(names))
      b = index(line(a + 1:), '"')
      if (b == 0) exit
      nv = nv + 1
      names(nv) = line(a + 1:a + b - 1)
      a = a + b + 1
      if (a > len_trim(line)) exit
      b = index(line(a:), '"')
      if (b == 0) exit
      a = a + b - 1
    end do
  end subroutine

  !> The next coarser level's blocks (Setup_Multigrid, Mod_Multigrid.f90:37-70, and
  !> Check_Multigrid, Lib_Multigrid.f90:12-35): half the cells in each direction, a
  !> flat block keeping nk = 1, its nodes every other node of the finer level
  !> (Coarse_Grid, :194-213), with their metrics
  subroutine coarse_blocks(fine, coarse)
    type(host_block_t), intent(in) :: fine(:)
    type(host_block_t), allocatable, intent(out) :: coarse(:)
    integer :: b, step3
    allocate(coarse(size(fine)))
    do b = 1, size(fine)
      if (mod(fine(b)%dim(1), 2) /= 0 .or. mod(fine(b)%dim(2), 2) /= 0 .or. &
          (fine(b)%dim(3) /= 1 .and. mod(fine(b)%dim(3), 2) /= 0)) &
        call refuse('Check_Multigrid', 'block '//itoa(b)//' not divisible by 2^(MGL-1)')
      coarse(b)%dim(1:2) = fine(b)%dim(1:2) / 2
      coarse(b)%dim(3) = max(1, fine(b)%dim(3) / 2)
      step3 = 2
      if (fine(b)%dim(3) == 1) step3 = 1
      allocate(coarse(b)%node(3, 0:coarse(b)%dim(1), 0:coarse(b)%dim(2), 0:coarse(b)%dim(3)))
      coarse(b)%node = fine(b)%node(:, 0:fine(b)%dim(1):2, 0:fine(b)%dim(2):2, 0:fine(b)%dim(3):step3)
      call block_metrics(coarse(b))
    end do
  end subroutine

  ! ------------------------------------------------------------------ BC
  !> The ATLAS boundary file (Setup_BC, IO_BC.f90:65-344) into the C-ABI's record
  !> arrays: reci(ICE_REC_NI, nrec), recd(ICE_REC_ND, nrec), record idx =
  !> (n-1)*ngroups + p.  The file holds one copy of the table for all families, or
  !> one copy per family, block by block with every copy of a block together over
  !> the same cells (copy c is family c); any other number of copies is refused.
  subroutine read_bc(path, blocks, ngroups, reci, recd, donor, weight)
    character(len=*), intent(in) :: path
    type(host_block_t), intent(in) :: blocks(:)
    integer, intent(in) :: ngroups
    integer(I4), allocatable, intent(out) :: reci(:,:), donor(:,:)
    real(R8), allocatable, intent(out) :: recd(:,:), weight(:)
    integer :: u, ios, nline, ncell, n, b, i, j, k, f, code, p, s, nchi(2), ndon, cs(9)
    integer :: nrec, ncopy, c, r, pos, base, nbb
    logical :: eof
    integer, allocatable :: ri(:,:), dtmp(:,:)
    real(R8), allocatable :: rd(:,:), wtmp(:)
    character(len=1024) :: line
    character(len=64) :: tok(10)
    real(R8) :: ang(2)
    logical :: nrm(2), onf
    ncell = 0
    do b = 1, size(blocks)
      ncell = ncell + 2 * (blocks(b)%dim(2) * blocks(b)%dim(3) + blocks(b)%dim(1) * blocks(b)%dim(3) + &
                           blocks(b)%dim(1) * blocks(b)%dim(2))
    end do
    allocate(ri(0:ICE_REC_NI-1, ncell * max(ngroups, 1)), rd(0:ICE_REC_ND-1, ncell * max(ngroups, 1)), &
             dtmp(4, 1024), wtmp(1024))
    ri = 0
    rd = 0
    ndon = 0
    nline = 0
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call refuse('Setup_BC', 'cannot open '//trim(path))
    ! every record of the file, in file order (Count_BC_Records walks the same rules)
    nrec = 0
    do
      call next_line(eof)
      if (eof) exit
      n = nrec + 1
      if (n > size(ri, 2)) call line_error('more records than one copy of the boundary table per family')
      nrec = n
      read(line, *, iostat=ios) b, i, j, k, f, code
      if (ios /= 0) call line_error('a header needs b i j k f code')
      if (code /= 0 .and. (code < 100 .or. code > 999)) call line_error('bad BC code '//itoa(code))
      if (b < 1 .or. b > size(blocks)) call line_error('block '//itoa(b)//' is not in the mesh')
      onf = (f == 1 .and. i == 1) .or. (f == 2 .and. i == blocks(b)%dim(1)) .or. (f == 3 .and. j == 1) .or. &
            (f == 4 .and. j == blocks(b)%dim(2)) .or. (f == 5 .and. k == 1) .or. (f == 6 .and. k == blocks(b)%dim(3))
      if (.not. onf .or. i < 1 .or. j < 1 .or. k < 1 .or. i > blocks(b)%dim(1) .or. j > blocks(b)%dim(2) .or. &
          k > blocks(b)%dim(3)) call line_error('the cell is not on the face it names')
      ri(R_B, n) = b; ri(R_I, n) = i; ri(R_J, n) = j; ri(R_K, n) = k; ri(R_F, n) = f; ri(R_TYPE, n) = code
      select case (code)
      case (101, 201)
        call next_line()
        read(line, *, iostat=ios) cs
        if (ios /= 0) call line_error('a connection payload needs bs is js ks fs d11 d12 d21 d22')
        ri(R_BS, n) = cs(1); ri(R_IS, n) = cs(2); ri(R_JS, n) = cs(3); ri(R_KS, n) = cs(4); ri(R_FS, n) = cs(5)
      case (102)
        call next_line()
        read(line, *, iostat=ios) nchi
        if (ios /= 0) call line_error('a chimera payload needs nchi_g1 nchi_g2')
        ri(R_NCHI1, n) = nchi(1); ri(R_NCHI2, n) = nchi(2); ri(R_DONOR0, n) = ndon
        do s = 1, nchi(1) + nchi(2)
          call next_line()
          if (ndon == size(wtmp)) call grow()
          ndon = ndon + 1
          read(line, *, iostat=ios) dtmp(:, ndon), wtmp(ndon)
          if (ios /= 0) call line_error('a donor line needs b i j k weight')
        end do
        if (sum(wtmp(ndon - nchi(1) - nchi(2) + 1:ndon - nchi(2))) < 0.5_R8 .or. &
            sum(wtmp(ndon - nchi(2) + 1:ndon)) < 0.5_R8) &
          call line_error('a chimera ghost layer whose donor weights sum to ~0')
      case (401:403)
        call next_line()
        tok = ''
        read(line, *, iostat=ios) tok(1:6)
        if (ios /= 0) call line_error('an inlet payload needs gp v alpha beta T radius')
        do s = 1, 2
          nrm(s) = trim(tok(2 + s)) == 'normal'
          ang(s) = 0
          if (.not. nrm(s)) read(tok(2 + s), *, iostat=ios) ang(s)
          if (ios /= 0) call line_error('a direction that is neither a number nor normal')
        end do
        if (nrm(1) .neqv. nrm(2)) call line_error('alpha and beta are both normal or both numbers')
        read(tok(1), *) rd(RD_MASSFLUX, n)
        read(tok(2), *) rd(RD_VELOCITY, n)
        read(tok(5), *) rd(RD_TEMPERATURE, n)
        read(tok(6), *) rd(RD_RADIUS, n)
        rd(RD_ALPHA, n) = ang(1); rd(RD_BETA, n) = ang(2)
        ri(R_NORMAL_DIR, n) = merge(1, 0, nrm(1))
        if (code == 401) then
          read(line, *, iostat=ios) tok(1:10)
          if (ios /= 0) call line_error('a 401 line without its 10th field G')
          read(tok(10), *) rd(RD_GFIXED, n)
        end if
        if (rd(RD_MASSFLUX, n) /= 0 .and. (rd(RD_VELOCITY, n) == 0 .or. .not. rd(RD_RADIUS, n) > 0)) &
          call line_error('a mass flux with a zero velocity or a non-positive radius')
      case (0, 200, 300, 301, 400)
      case default
        call line_error('unsupported ATLAS BC code '//itoa(code))
      end select
    end do
    close(u)
    if (nrec < ncell) call refuse('Setup_BC', trim(path)//': the file ends before the last boundary cell')
    if (mod(nrec, ncell) /= 0) call refuse('Setup_BC', trim(path)//': the file holds '//itoa(nrec)// &
      ' records, which is not a whole number of copies of the '//itoa(ncell)//' boundary cells of this mesh')
    ncopy = nrec / ncell
    if (ncopy /= 1 .and. ncopy /= ngroups) call refuse('Setup_BC', trim(path)//': the file carries '//itoa(ncopy)// &
      ' copies of the boundary table but this run has '//itoa(ngroups)//' particle families')
    allocate(reci(0:ICE_REC_NI-1, ncell * ngroups), recd(0:ICE_REC_ND-1, ncell * ngroups))
    if (ncopy == 1) then
      ! one record for every family, each carrying its own cell: any order
      do n = 1, ncell
        do p = 1, ngroups
          reci(:, (n - 1) * ngroups + p) = ri(:, n)
          reci(R_P, (n - 1) * ngroups + p) = p
          recd(:, (n - 1) * ngroups + p) = rd(:, n)
        end do
      end do
    else
      ! block by block, the copies of a block together, each over copy 1's cells
      pos = 0
      base = 0
      do b = 1, size(blocks)
        nbb = 2 * (blocks(b)%dim(2) * blocks(b)%dim(3) + blocks(b)%dim(1) * blocks(b)%dim(3) + &
                   blocks(b)%dim(1) * blocks(b)%dim(2))
        do c = 1, ncopy
          do r = 1, nbb
            pos = pos + 1
            if (ri(R_B, pos) /= b) call refuse('Setup_BC', trim(path)//': record '//itoa(pos)//' belongs to block '// &
              itoa(ri(R_B, pos))//' where the mesh expects block '//itoa(b)//': a multi-population file lists '// &
              'the blocks in mesh order, every copy of one block together')
            if (c > 1) then
              if (any(ri([R_I, R_J, R_K, R_F], pos) /= ri([R_I, R_J, R_K, R_F], pos - (c - 1) * nbb))) &
                call refuse('Setup_BC', trim(path)//': copy '//itoa(c)//' of block '//itoa(b)// &
                            ' does not repeat the cells of copy 1 in the same order')
            end if
            n = base + r
            reci(:, (n - 1) * ngroups + c) = ri(:, pos)
            reci(R_P, (n - 1) * ngroups + c) = c
            recd(:, (n - 1) * ngroups + c) = rd(:, pos)
          end do
        end do
        base = base + nbb
      end do
    end if
    allocate(donor(4, max(ndon, 1)), weight(max(ndon, 1)))
    donor = 0
    weight = 0
    donor(:, 1:ndon) = dtmp(:, 1:ndon)
    weight(1:ndon) = wtmp(1:ndon)

  contains

    !> the next non-blank line; eof (when present) says that the file ended there,
    !> else the end of the file stops the setup
    subroutine next_line(eof)
      logical, intent(out), optional :: eof
      if (present(eof)) eof = .false.
      do
        read(u, '(A)', iostat=ios) line
        if (ios /= 0) then
          if (present(eof)) then
            eof = .true.
            return
          end if
          call refuse('Setup_BC', trim(path)//': the file ends before the last boundary cell')
        end if
        nline = nline + 1
        if (len_trim(line) > 0) exit
      end do
    end subroutine

    subroutine line_error(msg)
      character(len=*), intent(in) :: msg
      call refuse('Setup_BC', trim(path)//', line '//itoa(nline)//', record '//itoa(n)//': '//msg)
    end subroutine

    subroutine grow()
      integer, allocatable :: d2(:,:)
      real(R8), allocatable :: w2(:)
      allocate(d2(4, 2 * size(wtmp)), w2(2 * size(wtmp)))
      d2(:, 1:size(wtmp)) = dtmp
      w2(1:size(wtmp)) = wtmp
      call move_alloc(d2, dtmp)
      call move_alloc(w2, wtmp)
    end subroutine
  end subroutine

  ! ------------------------------------------------------------------ metrics
  pure function cross(a, c) result(d)
    real(R8), intent(in) :: a(3), c(3)
    real(R8) :: d(3)
    d(1) = a(2) * c(3) - a(3) * c(2)
    d(2) = a(3) * c(1) - a(1) * c(3)
    d(3) = a(1) * c(2) - a(2) * c(1)
  end function

  pure real(R8) function dot(a, c)
    real(R8), intent(in) :: a(3), c(3)
    dot = (a(1) * c(1) + a(2) * c(2)) + a(3) * c(3)
  end function

  !> Face normals and areas, volumes, metric tensor and lengths of a block from
  !> its nodes (setup_metrics, Mod_Metrics.f90:19-59: extrapolate_nodes :460-495,
  !> normal_signs :500-545, compute_metrics :548-690, compute_metric_tensor
  !> :700-779), in the operation order the reference rounds in.
  subroutine block_metrics(blk)
    type(host_block_t), intent(inout) :: blk
    real(R8), allocatable :: x(:,:,:,:)
    real(R8) :: d1(3), d2(3), d3(3), c1(3), c2(3), sgn(3), a, r(3, 8), ea1(3), ea2(3), ea3(3), vol, dete
    real(R8) :: am(3, 3), det, cof(3, 3)
    integer :: im, jm, km, i, j, k, m, row
    im = blk%dim(1); jm = blk%dim(2); km = blk%dim(3)
    allocate(x(3, -1:im+1, -1:jm+1, -1:km+1))
    x(:, 0:im, 0:jm, 0:km) = blk%node
    x(:, -1, 0:jm, 0:km) = 2.0_R8 * x(:, 0, 0:jm, 0:km) - x(:, 1, 0:jm, 0:km)
    x(:, im+1, 0:jm, 0:km) = 2.0_R8 * x(:, im, 0:jm, 0:km) - x(:, im-1, 0:jm, 0:km)
    x(:, :, -1, 0:km) = 2.0_R8 * x(:, :, 0, 0:km) - x(:, :, 1, 0:km)
    x(:, :, jm+1, 0:km) = 2.0_R8 * x(:, :, jm, 0:km) - x(:, :, jm-1, 0:km)
    x(:, :, :, -1) = 2.0_R8 * x(:, :, :, 0) - x(:, :, :, 1)
    x(:, :, :, km+1) = 2.0_R8 * x(:, :, :, km) - x(:, :, :, km-1)
    ! orientation of the three face families from the first cell
    d3 = cross(x(:, 1, 1, 1) - x(:, 1, 0, 0), x(:, 1, 0, 1) - x(:, 1, 1, 0))
    c1 = 0.25_R8 * (((x(:, 1, 1, 1) + x(:, 1, 0, 1)) + x(:, 1, 0, 0)) + x(:, 1, 1, 0))
    c2 = 0.25_R8 * (((x(:, 0, 1, 1) + x(:, 0, 0, 1)) + x(:, 0, 0, 0)) + x(:, 0, 1, 0))
    sgn(1) = sign(1.0_R8, ((c1(1) - c2(1)) * d3(1) + (c1(2) - c2(2)) * d3(2)) + (c1(3) - c2(3)) * d3(3))
    d3 = cross(x(:, 0, 1, 1) - x(:, 1, 1, 0), x(:, 1, 1, 1) - x(:, 0, 1, 0))
    c1 = 0.25_R8 * (((x(:, 1, 1, 1) + x(:, 0, 1, 1)) + x(:, 0, 1, 0)) + x(:, 1, 1, 0))
    c2 = 0.25_R8 * (((x(:, 1, 0, 1) + x(:, 0, 0, 1)) + x(:, 0, 0, 0)) + x(:, 1, 0, 0))
    sgn(2) = sign(1.0_R8, ((c1(1) - c2(1)) * d3(1) + (c1(2) - c2(2)) * d3(2)) + (c1(3) - c2(3)) * d3(3))
    d3 = cross(x(:, 0, 1, 1) - x(:, 1, 0, 1), x(:, 0, 0, 1) - x(:, 1, 1, 1))
    c1 = 0.25_R8 * (((x(:, 1, 1, 1) + x(:, 0, 1, 1)) + x(:, 0, 0, 1)) + x(:, 1, 0, 1))
    c2 = 0.25_R8 * (((x(:, 1, 1, 0) + x(:, 0, 1, 0)) + x(:, 0, 0, 0)) + x(:, 1, 0, 0))
    sgn(3) = sign(1.0_R8, ((c1(1) - c2(1)) * d3(1) + (c1(2) - c2(2)) * d3(2)) + (c1(3) - c2(3)) * d3(3))

    allocate(blk%fi(4, 0:im, 1:jm, 1:km), blk%fj(4, 1:im, 0:jm, 1:km), blk%fk(4, 1:im, 1:jm, 0:km))
    do k = 1, km
      do j = 1, jm
        do i = 0, im
          call face(x(:, i, j, k) - x(:, i, j-1, k-1), x(:, i, j-1, k) - x(:, i, j, k-1), sgn(1), blk%fi(:, i, j, k))
        end do
      end do
    end do
    do k = 1, km
      do j = 0, jm
        do i = 1, im
          call face(x(:, i-1, j, k) - x(:, i, j, k-1), x(:, i, j, k) - x(:, i-1, j, k-1), sgn(2), blk%fj(:, i, j, k))
        end do
      end do
    end do
    do k = 0, km
      do j = 1, jm
        do i = 1, im
          call face(x(:, i-1, j, k) - x(:, i, j-1, k), x(:, i-1, j-1, k) - x(:, i, j, k), sgn(3), blk%fk(:, i, j, k))
        end do
      end do
    end do

    allocate(blk%vol(im, jm, km))
    do k = 1, km
      do j = 1, jm
        do i = 1, im
          r(:, 1) = 0
          do m = 1, 7
            r(:, m + 1) = x(:, i-1+iand(m, 1), j-1+iand(ishft(m, -1), 1), k-1+iand(ishft(m, -2), 1)) - x(:, i-1, j-1, k-1)
          end do
          ea1 = 0.25_R8 * (((r(:, 2) + (r(:, 6) - r(:, 5))) + (r(:, 4) - r(:, 3))) + (r(:, 8) - r(:, 7)))
          ea2 = 0.25_R8 * (((r(:, 3) + (r(:, 7) - r(:, 5))) + (r(:, 4) - r(:, 2))) + (r(:, 8) - r(:, 6)))
          ea3 = 0.25_R8 * (((r(:, 5) + (r(:, 7) - r(:, 3))) + (r(:, 6) - r(:, 2))) + (r(:, 8) - r(:, 4)))
          vol = dot(ea1, cross(r(:, 7), r(:, 5) - r(:, 3)) + cross(r(:, 8) - r(:, 2), r(:, 6) - r(:, 4)))
          vol = vol + dot(ea2, cross(r(:, 6), r(:, 2) - r(:, 5)) + cross(r(:, 8) - r(:, 3), r(:, 4) - r(:, 7)))
          vol = vol + dot(ea3, cross(r(:, 4), r(:, 3) - r(:, 2)) + cross(r(:, 8) - r(:, 5), r(:, 7) - r(:, 6)))
          vol = vol / 12.0_R8
          dete = dot(ea1, cross(ea2, ea3))
          if (.not. vol * dete > 0) call refuse('compute_metrics', 'a cell whose volume is zero or has not the '// &
                                                'sign of det at its centre')
          blk%vol(i, j, k) = abs(vol)
        end do
      end do
    end do

    allocate(blk%dl(3, 0:im+1, 0:jm+1, 0:km+1), blk%m(3, 3, im, jm, km))
    do k = 0, km + 1
      do j = 0, jm + 1
        do i = 0, im + 1
          am(1, :) = 0.25_R8 * (((((((P(1,1,1) - P(0,1,1)) + P(1,0,1)) - P(0,0,1)) + P(1,1,0)) - P(0,1,0)) &
                                + P(1,0,0)) - P(0,0,0))
          am(2, :) = 0.25_R8 * (((((((P(1,1,1) - P(1,0,1)) + P(0,1,1)) - P(0,0,1)) + P(1,1,0)) - P(1,0,0)) &
                                + P(0,1,0)) - P(0,0,0))
          am(3, :) = 0.25_R8 * (((((((P(1,1,1) - P(1,1,0)) + P(0,1,1)) - P(0,1,0)) + P(1,0,1)) - P(1,0,0)) &
                                + P(0,0,1)) - P(0,0,0))
          do row = 1, 3
            blk%dl(row, i, j, k) = sqrt((am(row, 1) * am(row, 1) + am(row, 2) * am(row, 2)) + am(row, 3) * am(row, 3))
          end do
          if (i < 1 .or. j < 1 .or. k < 1 .or. i > im .or. j > jm .or. k > km) cycle
          det = (((((am(1,1) * am(2,2)) * am(3,3) - (am(1,1) * am(2,3)) * am(3,2)) - (am(1,2) * am(2,1)) * am(3,3)) &
                 + (am(1,2) * am(2,3)) * am(3,1)) + (am(1,3) * am(2,1)) * am(3,2)) - (am(1,3) * am(2,2)) * am(3,1)
          cof(1,1) = am(2,2) * am(3,3) - am(2,3) * am(3,2)
          cof(1,2) = -(am(2,1) * am(3,3) - am(2,3) * am(3,1))
          cof(1,3) = am(2,1) * am(3,2) - am(2,2) * am(3,1)
          cof(2,1) = -(am(1,2) * am(3,3) - am(1,3) * am(3,2))
          cof(2,2) = am(1,1) * am(3,3) - am(1,3) * am(3,1)
          cof(2,3) = -(am(1,1) * am(3,2) - am(1,2) * am(3,1))
          cof(3,1) = am(1,2) * am(2,3) - am(1,3) * am(2,2)
          cof(3,2) = -(am(1,1) * am(2,3) - am(1,3) * am(2,1))
          cof(3,3) = am(1,1) * am(2,2) - am(1,2) * am(2,1)
          if (abs(det) == 0) then
            blk%m(:, :, i, j, k) = 0
            do row = 1, 3
              blk%m(row, row, i, j, k) = 1
            end do
          else
            blk%m(:, :, i, j, k) = cof / det
          end if
        end do
      end do
    end do

  contains

    function P(a1, b1, c1_) result(v)
      integer, intent(in) :: a1, b1, c1_
      real(R8) :: v(3)
      v = x(:, i-1+a1, j-1+b1, k-1+c1_)
    end function

    subroutine face(e1, e2, s, out)
      real(R8), intent(in) :: e1(3), e2(3), s
      real(R8), intent(out) :: out(4)
      real(R8) :: h(3), ar
      h = cross(e1, e2) * 0.5_R8
      ar = sqrt((h(1) * h(1) + h(2) * h(2)) + h(3) * h(3))
      if (ar == 0) then
        out(1:3) = 0
      else
        out(1:3) = h / ar * s
      end if
      out(4) = ar
    end subroutine
  end subroutine

  ! ------------------------------------------------------------------ solution
  !> The solution in the layout of the reference's part-field.tec: one zone per
  !> block, nodal x y z and the cell-centred primitive variables, 17 digits.
  subroutine write_solution(path, names, blocks, iterations)
    character(len=*), intent(in) :: path, names
    type(host_block_t), intent(in) :: blocks(:)
    integer, intent(in) :: iterations
    integer :: u, b, v, i, j, k, nc, ios
    open(newunit=u, file=path, status='replace', action='write', iostat=ios)
    if (ios /= 0) call refuse('write_solution', 'cannot write '//trim(path))
    write(u, '(A)') ' VARIABLES = x y z '//trim(names)
    do b = 1, size(blocks)
      nc = size(blocks(b)%prim, 1)
      write(u, '(A,I0,A,I0,A,I0,A,I0,A,I0,A)') ' ZONE  T = Block', b, ', I=', blocks(b)%dim(1) + 1, ', J=', &
        blocks(b)%dim(2) + 1, ', K=', blocks(b)%dim(3) + 1, ', DATAPACKING=BLOCK, VARLOCATION=([1-3]=NODAL,[4-', &
        3 + nc, ']=CELLCENTERED)'
      do v = 1, 3
        write(u, '(ES25.17E3)') (((blocks(b)%node(v, i, j, k), i = 0, blocks(b)%dim(1)), j = 0, blocks(b)%dim(2)), &
                                 k = 0, blocks(b)%dim(3))
      end do
      do v = 1, nc
        write(u, '(ES25.17E3)') (((blocks(b)%prim(v, i, j, k), i = 1, blocks(b)%dim(1)), j = 1, blocks(b)%dim(2)), &
                                 k = 1, blocks(b)%dim(3))
      end do
    end do
    write(u, '(A,I0,A)') 'DATASETAUXDATA ICE_ITERATION = "', iterations, '"'
    close(u)
  end subroutine

  function itoa(n) result(s)
    integer, intent(in) :: n
    character(len=:), allocatable :: s
    character(len=16) :: b
    write(b, '(I0)') n
    s = trim(b)
  end function

end module ice_host_m
