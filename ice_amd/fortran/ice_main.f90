!> ICE's driver loop over the device path: set-up from the case directory's
!> input.ini and INPUT/ files (Wrap_Setup.f90:48-121 for the parts the hot path
!> reads), then Explicit_Step through the C-ABI until iter-threshold
!> (main.f90:50-55, Wrap_Solve.f90:10-42), printing the residual norms every
!> shell-diter iterations, and the solution in OUTPUT/part-field.tec.
!>
!> usage: ice_driver <case directory> [device]
program ice_driver
  use, intrinsic :: iso_c_binding
  use, intrinsic :: iso_fortran_env, only: R8 => real64, I8 => int64, output_unit
  use ice_gpu_m
  use ice_host_m
  implicit none
  !> one grid level: its blocks and its context (Setup_Multigrid: a domain per level)
  type :: level_t
    type(host_block_t), allocatable :: b(:)
    type(c_ptr) :: ctx = c_null_ptr
  end type
  character(len=1024) :: dir, arg, bcpath, gaspath
  type(host_block_t), allocatable :: gas(:)
  logical :: coupled
  type(registry_t) :: reg
  type(level_t) :: lev(MAXLEV)
  type(host_block_t), allocatable :: blocks(:)
  type(ice_config_t) :: cfg
  type(ice_family_t), allocatable :: fams(:)
  type(ice_status_t) :: st
  integer(c_int32_t), allocatable :: reci(:,:), donor(:,:), dims(:)
  real(R8), allocatable :: recd(:,:), weight(:), fam(:), prop_cols(:,:)
  integer :: t_min, prop_flags
  real(R8) :: rho_mat, cs, total(5), step_ms, kernel_ms, tile_ms(24), res_ref
  integer(c_int32_t) :: tile(4), stopped
  integer(c_int64_t) :: launches
  type(c_ptr) :: ctx
  integer :: device, nb, b, nc, ndir, iters, it, n, shell, rc, ndonor, mgl, m, total_iters
  integer(I8) :: cells, t0, t1, rate
  character(len=2048) :: names
  character(len=16) :: key
  integer :: p, nvar
  integer :: off(14), ncf(14)

  call get_command_argument(1, dir)
  if (len_trim(dir) == 0) call refuse('ice_driver', 'usage: ice_driver <case directory> [device]')
  device = 0
  call get_command_argument(2, arg)
  if (len_trim(arg) > 0) read(arg, *) device

  call read_registry(trim(dir)//'/input.ini', reg)
  ! the initial condition holds every family's primitive variables, family after
  ! family ([ICE-Family<p>] closure), as ICE's part-ic.tec does
  nvar = 0
  names = ''
  do p = 1, reg%nfam
    select case (trim(reg%closure(p)))
    case ('MK'); ncf(p) = 6;  call add_names('rho u v w T n', p)
    case ('IG'); ncf(p) = 7;  call add_names('rho u v w P T n', p)
    case default; ncf(p) = 12; call add_names('rho u v w P11 P12 P13 P22 P23 P33 T n', p)
    end select
    off(p) = nvar
    nvar = nvar + ncf(p)
  end do
  nc = nvar
  mgl = reg_int(reg, 'levels')
  if (mgl < 1 .or. mgl > MAXLEV) call refuse('Register_Multigrid_Levels', 'levels must be 1 to 4 here')
  call read_ic(trim(dir)//'/INPUT/part-ic.tec', nc, blocks)
  call read_material(trim(dir)//'/INPUT/part-properties.dat', rho_mat, cs, t_min, prop_cols, prop_flags)
  ! one-way coupling when the gas field is there (Assign_Setup.f90:50-54): gas.tec of
  ! gas-path, ten cell variables rho u v w T R gam k mu dt (setup_gas, Mod_Phase.f90:46-58)
  gaspath = trim(dir)//'/'//trim(reg_str(reg, 'gas-path'))//'gas.tec'
  inquire(file=trim(gaspath), exist=coupled)
  if (coupled) then
    call read_ic(trim(gaspath), 10, gas)
    if (size(gas) /= size(blocks)) call refuse('setup_gas', 'the gas field has another number of blocks')
    do b = 1, size(blocks)
      if (any(gas(b)%dim /= blocks(b)%dim)) call refuse('setup_gas', 'the gas field of block '//itoa_m(b)// &
                                                        ' has other dimensions')
    end do
  end if
  nb = size(blocks)
  ndir = 0
  do n = 1, 3
    if (any([(blocks(b)%dim(n) > 1, b = 1, nb)])) ndir = ndir + 1
  end do
  call config_from_registry(reg, ndir, rho_mat, cs, coupled, cfg, fams)
  do b = 1, nb
    call block_metrics(blocks(b))
  end do
  call move_alloc(blocks, lev(1)%b)
  do m = 2, mgl
    call coarse_blocks(lev(m - 1)%b, lev(m)%b)
  end do

  ! a context per level: geometry, boundary table (part-bc.txt, part-bc<m>.txt), material
  allocate(dims(3 * nb))
  do m = 1, mgl
    do b = 1, nb
      dims(3 * b - 2:3 * b) = lev(m)%b(b)%dim
    end do
    bcpath = trim(dir)//'/INPUT/part-bc.txt'
    if (m > 1) write(bcpath, '(A,I0,A)') trim(dir)//'/INPUT/part-bc', m, '.txt'
    call read_bc(trim(bcpath), lev(m)%b, cfg%ngroups, reci, recd, donor, weight)
    ctx = c_null_ptr
    call check(ice_gpu_create(device, cfg, fams, nb, dims, ctx), 'create')
    lev(m)%ctx = ctx
    call check(ice_gpu_set_properties(ctx, 1, t_min, size(prop_cols, 1), prop_cols, int(prop_flags, c_int32_t)), &
               'set_properties')
    do b = 1, nb
      call check(ice_gpu_set_geometry(ctx, b, lev(m)%b(b)%vol, lev(m)%b(b)%dl, lev(m)%b(b)%m, lev(m)%b(b)%fi, &
                                      lev(m)%b(b)%fj, lev(m)%b(b)%fk), 'set_geometry')
    end do
    ndonor = 0
    if (any(reci(R_TYPE, :) == 102)) ndonor = size(weight)
    call check(ice_gpu_set_boundaries(ctx, size(reci, 2), reci, recd, ndonor, donor, weight), 'set_boundaries')
    if (coupled) then
      do b = 1, nb
        call put_gas(b, lev(m)%b(b)%dim)
      end do
    end if
    if (m == 1) then
      do b = 1, nb
        do p = 1, cfg%ngroups
          call check(ice_gpu_put_state(ctx, b, p, lev(1)%b(b)%prim(off(p) + 1:off(p) + ncf(p), :, :, :)), 'put_state')
        end do
      end do
      ! res-threshold ends steady runs, on the finest level (Mod_Explicit.f90:196-200)
      call check(ice_gpu_set_res_threshold(ctx, reg_real(reg, 'res-threshold')), 'set_res_threshold')
    else
      ! the initial solution of each coarser level (Mod_Multigrid.f90:76)
      ctx = lev(m - 1)%ctx
      call check(ice_gpu_restrict(lev(m - 1)%ctx, lev(m)%ctx), 'restrict')
    end if
    if (reg_bool(reg, 'autotune')) then
      ctx = lev(m)%ctx
      call check(ice_gpu_autotune(ctx, 2, tile, tile_ms), 'autotune')
      write(output_unit, '(A,I0,A,3(1X,I0),A,I0)') ' level ', m, ': autotuned stage tile:', tile(1:3), &
        ', dispatch ', tile(4)
    end if
  end do

  shell = max(1, reg_int(reg, 'shell-diter'))
  allocate(fam(5 * cfg%ngroups))
  cells = 0
  do b = 1, nb
    cells = cells + product(int(lev(1)%b(b)%dim, I8))
  end do
  write(output_unit, '(A,I0,A,I0,A,I0,A)') ' ICE device path: ', nb, ' blocks, ', cells, ' cells, ', mgl, ' levels'
  call system_clock(t0, rate)
  total_iters = 0
  ! the levels, coarsest first, each for its level<m>-iter iterations, then prolongation
  ! onto the next finer one (Wrap_Solve.f90:28-39)
  do m = mgl, 1, -1
    ctx = lev(m)%ctx
    write(key, '(A,I0,A)') 'level', m, '-iter'
    iters = reg_int(reg, trim(key))
    if (m == 1) iters = min(iters, reg_int(reg, 'iter-threshold'))
    it = 0
    do while (it < iters)
      n = min(shell, iters - it)
      rc = ice_gpu_step(ctx, n, st)
      if (st%bad_recon /= 0) call refuse('state_reconstruction', 'unphysical state at first order')
      if (st%bad_state /= 0) then
        write(output_unit, '(A,5(1X,I0))') ' [ERROR] [ICE::state_update] non-physical state at block, i, j, k, family', &
          st%bad_block, st%bad_i, st%bad_j, st%bad_k, st%bad_p
        error stop 1
      end if
      call check(rc, 'step')
      call print_inflow400_lines(m)
      call check(ice_gpu_res_stop(ctx, stopped, res_ref), 'res_stop')
      if (stopped /= 0) then
        n = st%iter - (total_iters + it)
        it = it + n
        call check(ice_gpu_residual_norm(ctx, total, fam), 'residual_norm')
        write(output_unit, '(A,I8,A,5ES13.5)') ' iter', st%iter, '  res', total
        write(output_unit, '(A,ES12.4,A)') ' converged: density residual below res-threshold x ', res_ref, &
          ' (the first non-zero one)'
        exit
      end if
      it = it + n
      call check(ice_gpu_residual_norm(ctx, total, fam), 'residual_norm')
      if (mgl > 1) then
        write(output_unit, '(A,I0,A,I8,A,5ES13.5)') ' level ', m, ' iter', st%iter, '  res', total
      else
        write(output_unit, '(A,I8,A,5ES13.5)') ' iter', st%iter, '  res', total
      end if
    end do
    total_iters = total_iters + it
    if (m > 1) then
      ctx = lev(m - 1)%ctx
      call check(ice_gpu_prolong(lev(m)%ctx, lev(m - 1)%ctx), 'prolong')
    end if
  end do
  ctx = lev(1)%ctx
  call system_clock(t1)
  call check(ice_gpu_timing(ctx, step_ms, kernel_ms, launches), 'timing')
  write(output_unit, '(A,F10.3,A)') ' wall ', real(t1 - t0, R8) / rate, ' s'
  if (st%n_floored > 0) write(output_unit, '(A,I0)') ' [WARNING] floored densities: ', st%n_floored

  do m = mgl, 1, -1
    ctx = lev(m)%ctx
    call print_inflow400_totals(m)
  end do

  ctx = lev(1)%ctx
  do b = 1, nb
    do p = 1, cfg%ngroups
      call check(ice_gpu_get_state(ctx, b, p, lev(1)%b(b)%prim(off(p) + 1:off(p) + ncf(p), :, :, :)), 'get_state')
    end do
  end do
  call execute_command_line('mkdir -p '//trim(dir)//'/OUTPUT')
  call write_solution(trim(dir)//'/OUTPUT/part-field.tec', names, lev(1)%b, st%iter)
  do m = 1, mgl
    call ice_gpu_destroy(lev(m)%ctx)
  end do

contains

  !> The gas of block b (set_gas, Mod_Exchange.f90:63-81): prim (5, ni, nj, nk) and
  !> R, gam, k, mu from the gas field's cells; a coarse level takes none (it runs
  !> uncoupled levels only)
  subroutine put_gas(b, d)
    integer, intent(in) :: b, d(3)
    real(R8), allocatable :: g5(:,:,:,:), gR(:,:,:), gG(:,:,:), gK(:,:,:), gM(:,:,:)
    if (any(d /= gas(b)%dim)) call refuse('setup_gas', 'a coupled run on several grid levels is not supported')
    allocate(g5(5, d(1), d(2), d(3)))
    g5 = gas(b)%prim(1:5, 1:d(1), 1:d(2), 1:d(3))
    gR = gas(b)%prim(6, 1:d(1), 1:d(2), 1:d(3))
    gG = gas(b)%prim(7, 1:d(1), 1:d(2), 1:d(3))
    gK = gas(b)%prim(8, 1:d(1), 1:d(2), 1:d(3))
    gM = gas(b)%prim(9, 1:d(1), 1:d(2), 1:d(3))
    call check(ice_gpu_put_gas(ctx, b, g5, gR, gG, gK, gM), 'put_gas')
  end subroutine

  !> the names of family p's variables, suffixed _p<p>
  subroutine add_names(words, p)
    character(len=*), intent(in) :: words
    integer, intent(in) :: p
    character(len=8) :: w(12)
    integer :: q, ios
    w = ''
    read(words, *, iostat=ios) w
    do q = 1, 12
      if (len_trim(w(q)) > 0) names = trim(names)//' '//trim(w(q))//'_p'//trim(itoa_m(p))
    end do
  end subroutine

  function itoa_m(i) result(s)
    integer, intent(in) :: i
    character(len=12) :: s
    write(s, '(I0)') i
  end function

  !> The 400-inflow lines of the steps just run, in ICE's text (Mod_Inflow400.f90:118-123)
  subroutine print_inflow400_lines(level)
    integer, intent(in) :: level
    type(ice_inflow400_line_t), save :: lines(ICE_INFLOW400_LOG)
    integer(c_int32_t) :: nl
    integer(c_int64_t) :: dropped
    integer :: m
    character(len=48) :: which
    call check(ice_gpu_inflow400_lines(ctx, lines, ICE_INFLOW400_LOG, nl, dropped), 'inflow400_lines')
    do m = 1, nl
      which = ', faces: '
      if (lines(m)%relative /= 0) which = ', faces above 10 times its last line: '
      write(output_unit, '(A,I0,A,I0,A)') ' 400 boundary of '//trim(boundary_name(lines(m)%block, lines(m)%face, &
        lines(m)%family, level))//': particles entering at iteration ', lines(m)%iter, trim(which)//' ', lines(m)%faces, &
        ', largest '//trim(measure_text(lines(m)%family, lines(m)%largest))
    end do
    if (dropped > 0) write(output_unit, '(A,I0,A)') ' [WARNING] ', dropped, ' lines of the 400 report were dropped'
    flush(output_unit)
  end subroutine

  !> The boundaries through which particles entered over the run (Mod_Inflow400.f90:156-169)
  subroutine print_inflow400_totals(level)
    integer, intent(in) :: level
    type(ice_inflow400_total_t), allocatable :: t(:)
    integer(c_int32_t) :: nt
    integer :: m
    allocate(t(6 * nb * cfg%ngroups))
    call check(ice_gpu_inflow400_totals(ctx, t, size(t), nt), 'inflow400_totals')
    do m = 1, nt
      write(output_unit, '(A,I0,A,I0,A,I0,A)') ' 400 boundary of '//trim(boundary_name(t(m)%block, t(m)%face, &
        t(m)%family, level))//': particles entered in ', t(m)%stages, ' stages, first at iteration ', t(m)%first_iter, &
        ' (faces: ', t(m)%first_faces, ', largest '//trim(measure_text(t(m)%family, t(m)%first_largest))// &
        '), largest over the run: '//trim(number_text(t(m)%family, t(m)%largest))
    end do
  end subroutine

  function boundary_name(b, f, p, level) result(s)
    integer(c_int32_t), intent(in) :: b, f, p
    integer, intent(in) :: level
    character(len=80) :: s
    write(s, '(A,I0,A,I0)') 'block ', b, ', face ', f
    if (cfg%ngroups > 1) write(s, '(A,I0)') trim(s)//', family ', p
    if (mgl > 1) write(s, '(A,I0)') trim(s)//', grid level ', level
  end function

  function measure_text(p, x) result(s)
    integer(c_int32_t), intent(in) :: p
    real(R8), intent(in) :: x
    character(len=60) :: s
    select case (fams(p)%closure)
    case (ICE_IG); s = '|w|/sqrt(P/rho): '//number_text(p, x)
    case (ICE_AG); s = '|w|/sqrt(Pnn/rho): '//number_text(p, x)
    case default;  s = '|w|: '//number_text(p, x)
    end select
  end function

  function number_text(p, x) result(s)
    integer(c_int32_t), intent(in) :: p
    real(R8), intent(in) :: x
    character(len=30) :: s
    if (x >= 1e100_R8 .or. (x > 0._R8 .and. x < 1e-99_R8)) then
      write(s, '(ES12.4E3)') x
    else
      write(s, '(ES11.4)') x
    end if
    s = adjustl(s)
    if (fams(p)%closure == ICE_MK) s = trim(s)//' m/s'
  end function

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) call refuse('ice_gpu_'//what, ice_gpu_error_text(ctx))
  end subroutine

end program ice_driver
