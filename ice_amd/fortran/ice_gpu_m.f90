!> ISO_C_BINDING interface of the MI355X explicit residual-and-update path
!> (include/ice_gpu.h).  A Fortran host that keeps ICE's block and boundary types
!> calls these in place of the body of Explicit_Step
!> (src/lib/numerics/time/explicit/Mod_Explicit.f90:11-245), passing its arrays
!> with c_loc and their Fortran layouts unchanged.
module ice_gpu_m
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: ICE_MK = 0, ICE_IG = 1, ICE_AG = 2
  integer(c_int), parameter, public :: ICE_SAUREL = 0, ICE_RUSANOV = 1, ICE_HLLE = 2
  integer(c_int), parameter, public :: ICE_MODE_PLAIN = 0, ICE_MODE_BATCHED = 1
  integer(c_int), parameter, public :: ICE_REC_NI = 16, ICE_REC_ND = 7
  !> record fields, 0-based as in the header (ICE_R_*, ICE_RD_*)
  integer(c_int), parameter, public :: R_B = 0, R_I = 1, R_J = 2, R_K = 3, R_F = 4, R_TYPE = 5, R_P = 6, &
    R_BS = 7, R_IS = 8, R_JS = 9, R_KS = 10, R_FS = 11, R_NORMAL_DIR = 12, R_NCHI1 = 13, R_NCHI2 = 14, R_DONOR0 = 15
  integer(c_int), parameter, public :: RD_MASSFLUX = 0, RD_VELOCITY = 1, RD_TEMPERATURE = 2, RD_RADIUS = 3, &
    RD_ALPHA = 4, RD_BETA = 5, RD_GFIXED = 6

  !> ice_family_t
  type, bind(C), public :: ice_family_t
    integer(c_int32_t) :: closure = ICE_IG, riemann = ICE_RUSANOV, nfrozen = 0, frozen(6) = 0
    integer(c_int32_t) :: evap_select = 0, intf_select = 0, surrogate = 0
    real(c_double)     :: rho_mat = 0, cs = 0, lv = 0, emiss = 0, ep(8) = 0
  end type

  !> ice_config_t
  type, bind(C), public :: ice_config_t
    integer(c_int32_t) :: ngroups = 1, nrk = 2, time_accurate = 1, cfl_ramp_iter = 0, ndir = 3, limiter = 0
    integer(c_int32_t) :: shock_detector = 0, irs = 0, coupled = 0, drag_select = 13, heat_select = 8
    integer(c_int32_t) :: blow_select = 0, mode = ICE_MODE_PLAIN, var_yinf = 0, reserved(2) = 0
    real(c_double)     :: cfl = 0.5d0, dt_max = 1.d-4, tau_factor = 1.d0, irs_beta = 0.5d0
  end type

  !> ice_status_t
  type, bind(C), public :: ice_status_t
    integer(c_int64_t) :: n_floored = 0
    integer(c_int32_t) :: bad_state = 0, bad_block = 0, bad_i = 0, bad_j = 0, bad_k = 0, bad_p = 0
    integer(c_int32_t) :: bad_recon = 0, iter = 0
    real(c_double)     :: dtglobal = 0, time = 0
  end type

  !> ice_inflow400_line_t: a line of count_400_inflow
  type, bind(C), public :: ice_inflow400_line_t
    integer(c_int32_t) :: iter = 0, block = 0, face = 0, family = 0, faces = 0, relative = 0
    real(c_double)     :: largest = 0
  end type

  !> ice_inflow400_total_t: a boundary's totals (report_400_inflow)
  type, bind(C), public :: ice_inflow400_total_t
    integer(c_int32_t) :: block = 0, face = 0, family = 0, stages = 0, first_iter = 0, first_faces = 0
    real(c_double)     :: first_largest = 0, largest = 0
  end type
  integer, parameter, public :: ICE_INFLOW400_LOG = 4096
  !> columns and flags of ice_gpu_set_properties
  integer, parameter, public :: ICE_PROP_RHO = 0, ICE_PROP_CP = 1, ICE_PROP_E = 2, ICE_PROP_PSAT = 3, &
                                ICE_PROP_LV = 4, ICE_PROP_CPV = 5, ICE_PROP_EMISS = 6, ICE_PROP_NCOL = 7
  integer, parameter, public :: ICE_PROP_RHO_VARIES = 1, ICE_PROP_CS_VARIES = 2, ICE_PROP_USE_PSAT = 4, &
                                ICE_PROP_USE_LV = 8, ICE_PROP_USE_CPV = 16, ICE_PROP_USE_EMISS = 32

  public :: ice_gpu_create, ice_gpu_set_geometry, ice_gpu_set_boundaries, ice_gpu_put_state, ice_gpu_get_state, &
            ice_gpu_put_gas, ice_gpu_put_tau, ice_gpu_step, ice_gpu_residual_norm, ice_gpu_timing, &
            ice_gpu_last_error, ice_gpu_destroy, ice_gpu_error_text, ice_gpu_autotune, ice_gpu_inflow400_lines, &
            ice_gpu_inflow400_totals, ice_gpu_set_properties, ice_gpu_get_bc_maps, ice_gpu_get_source, &
            ice_gpu_put_gas_yinf, ice_gpu_set_keep_source, ice_gpu_set_res_threshold, ice_gpu_res_stop, &
            ice_gpu_restrict, ice_gpu_prolong, ice_gpu_set_tiles, ice_gpu_set_dispatch, ice_gpu_set_kernel_timing, &
            ice_gpu_mlp_grad, ice_gpu_mlp_forward, ice_gpu_surrogate_train, ice_gpu_surrogate_retrain_every, &
            ice_gpu_surrogate_params, ice_gpu_surrogate_set_params, ice_gpu_set_lk_stats, ice_gpu_lk_stats, &
            ice_gpu_set_mode, ice_gpu_set_surrogate, ice_gpu_set_counters, ice_gpu_get_counters, &
            ice_gpu_fill_ghosts, ice_gpu_put_imports, ice_gpu_add_births, ice_gpu_get_tau, ice_gpu_surrogate_retrain, &
            ice_gpu_residual_norm_noalloc, ice_gpu_lk_timing
  integer, parameter, public :: ICE_MLP_PARAMS = 6288

  interface
    integer(c_int) function ice_gpu_create(device, cfg, fams, nb, dims, ctx) bind(C, name='ice_gpu_create')
      import :: c_int, c_int32_t, c_ptr, ice_config_t, ice_family_t
      integer(c_int), value      :: device, nb
      type(ice_config_t)         :: cfg
      type(ice_family_t)         :: fams(*)
      integer(c_int32_t)         :: dims(*)
      type(c_ptr)                :: ctx
    end function
    integer(c_int) function ice_gpu_set_geometry(ctx, b, vol, dl, m, face_i, face_j, face_k) &
        bind(C, name='ice_gpu_set_geometry')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: b
      real(c_double) :: vol(*), dl(*), m(*), face_i(*), face_j(*), face_k(*)
    end function
    integer(c_int) function ice_gpu_set_boundaries(ctx, nrec, reci, recd, ndonor, donor, weight) &
        bind(C, name='ice_gpu_set_boundaries')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nrec, ndonor
      integer(c_int32_t) :: reci(*), donor(*)
      real(c_double) :: recd(*), weight(*)
    end function
    integer(c_int) function ice_gpu_put_state(ctx, b, p, prim) bind(C, name='ice_gpu_put_state')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: b, p
      real(c_double) :: prim(*)
    end function
    integer(c_int) function ice_gpu_get_state(ctx, b, p, prim) bind(C, name='ice_gpu_get_state')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: b, p
      real(c_double) :: prim(*)
    end function
    integer(c_int) function ice_gpu_put_gas(ctx, b, gprim, r, gam, k, mu) bind(C, name='ice_gpu_put_gas')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: b
      real(c_double) :: gprim(*), r(*), gam(*), k(*), mu(*)
    end function
    integer(c_int) function ice_gpu_put_tau(ctx, b, p, tau) bind(C, name='ice_gpu_put_tau')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: b, p
      real(c_double) :: tau(*)
    end function
    integer(c_int) function ice_gpu_step(ctx, nsteps, st) bind(C, name='ice_gpu_step')
      import :: c_int, c_ptr, ice_status_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      type(ice_status_t) :: st
    end function
    integer(c_int) function ice_gpu_residual_norm(ctx, total, fam) bind(C, name='ice_gpu_residual_norm')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: total(5), fam(*)
    end function
    integer(c_int) function ice_gpu_timing(ctx, step_ms, kernel_ms, launches) bind(C, name='ice_gpu_timing')
      import :: c_int, c_double, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: step_ms, kernel_ms
      integer(c_int64_t) :: launches
    end function
    integer(c_int) function ice_gpu_lk_timing(ctx, ms, launches) bind(C, name='ice_gpu_lk_timing')
      import :: c_int, c_double, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: ms
      integer(c_int64_t) :: launches
    end function
    integer(c_int) function ice_gpu_autotune(ctx, steps, best, ms) bind(C, name='ice_gpu_autotune')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: steps
      integer(c_int32_t) :: best(4)
      real(c_double) :: ms(24)
    end function
    integer(c_int) function ice_gpu_set_properties(ctx, p, t_min, n, cols, flags) bind(C, name='ice_gpu_set_properties')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: p, t_min, n
      real(c_double) :: cols(*)
      integer(c_int32_t), value :: flags
    end function
    integer(c_int) function ice_gpu_inflow400_lines(ctx, lines, max, n, dropped) bind(C, name='ice_gpu_inflow400_lines')
      import :: c_int, c_int32_t, c_int64_t, c_ptr, ice_inflow400_line_t
      type(c_ptr), value :: ctx
      type(ice_inflow400_line_t) :: lines(*)
      integer(c_int), value :: max
      integer(c_int32_t) :: n
      integer(c_int64_t) :: dropped
    end function
    integer(c_int) function ice_gpu_inflow400_totals(ctx, totals, max, n) bind(C, name='ice_gpu_inflow400_totals')
      import :: c_int, c_int32_t, c_ptr, ice_inflow400_total_t
      type(c_ptr), value :: ctx
      type(ice_inflow400_total_t) :: totals(*)
      integer(c_int), value :: max
      integer(c_int32_t) :: n
    end function
    integer(c_int) function ice_gpu_get_bc_maps(ctx, partner, owner, cf, p, ncell, first, rec, cap) &
        bind(C, name='ice_gpu_get_bc_maps')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: partner(*), owner(*), cf(*), ncell, first(*), rec(*)
      integer(c_int), value :: p, cap
    end function
    integer(c_int) function ice_gpu_get_source(ctx, b, p, source, tau) bind(C, name='ice_gpu_get_source')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: b, p
      real(c_double) :: source(*), tau(*)
    end function
    integer(c_int) function ice_gpu_put_gas_yinf(ctx, b, yinf) bind(C, name='ice_gpu_put_gas_yinf')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: b
      real(c_double) :: yinf(*)
    end function
    integer(c_int) function ice_gpu_set_keep_source(ctx, on) bind(C, name='ice_gpu_set_keep_source')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function ice_gpu_set_res_threshold(ctx, threshold) bind(C, name='ice_gpu_set_res_threshold')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: threshold
    end function
    integer(c_int) function ice_gpu_res_stop(ctx, stopped, reference) bind(C, name='ice_gpu_res_stop')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: stopped
      real(c_double) :: reference
    end function
    integer(c_int) function ice_gpu_restrict(fine, coarse) bind(C, name='ice_gpu_restrict')
      import :: c_int, c_ptr
      type(c_ptr), value :: fine, coarse
    end function
    integer(c_int) function ice_gpu_prolong(coarse, fine) bind(C, name='ice_gpu_prolong')
      import :: c_int, c_ptr
      type(c_ptr), value :: coarse, fine
    end function
    integer(c_int) function ice_gpu_set_tiles(ctx, tx, ty, tz) bind(C, name='ice_gpu_set_tiles')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: tx, ty, tz
    end function
    integer(c_int) function ice_gpu_set_dispatch(ctx, dispatch) bind(C, name='ice_gpu_set_dispatch')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: dispatch
    end function
    integer(c_int) function ice_gpu_set_kernel_timing(ctx, on) bind(C, name='ice_gpu_set_kernel_timing')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function ice_gpu_mlp_grad(ctx, n, x, y, params, grad, loss) bind(C, name='ice_gpu_mlp_grad')
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      real(c_float) :: x(*), y(*), params(*), grad(*), loss
    end function
    integer(c_int) function ice_gpu_mlp_forward(ctx, n, x, params, out) bind(C, name='ice_gpu_mlp_forward')
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      real(c_float) :: x(*), params(*), out(*)
    end function
    integer(c_int) function ice_gpu_surrogate_train(ctx, p, epochs, batch, lr, seed, report) &
        bind(C, name='ice_gpu_surrogate_train')
      import :: c_int, c_double, c_float, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: p, epochs, batch
      real(c_double), value :: lr
      integer(c_int64_t), value :: seed
      real(c_float) :: report(6)
    end function
    integer(c_int) function ice_gpu_surrogate_retrain_every(ctx, every, epochs) &
        bind(C, name='ice_gpu_surrogate_retrain_every')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: every, epochs
    end function
    integer(c_int) function ice_gpu_surrogate_params(ctx, p, params) bind(C, name='ice_gpu_surrogate_params')
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: p
      real(c_float) :: params(*)
    end function
    integer(c_int) function ice_gpu_surrogate_set_params(ctx, p, params) bind(C, name='ice_gpu_surrogate_set_params')
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: p
      real(c_float) :: params(*)
    end function
    integer(c_int) function ice_gpu_set_lk_stats(ctx, on) bind(C, name='ice_gpu_set_lk_stats')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function ice_gpu_lk_stats(ctx, p, counts) bind(C, name='ice_gpu_lk_stats')
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: p
      integer(c_int64_t) :: counts(4)
    end function
    integer(c_int) function ice_gpu_residual_norm_noalloc(ctx, total, fam) bind(C, name='ice_gpu_residual_norm_noalloc')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: total(5), fam(*)
    end function
    integer(c_int) function ice_gpu_surrogate_retrain(ctx, p, epochs) bind(C, name='ice_gpu_surrogate_retrain')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: p, epochs
    end function
    integer(c_int) function ice_gpu_get_tau(ctx, b, p, tau) bind(C, name='ice_gpu_get_tau')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: b, p
      real(c_double) :: tau(*)
    end function
    integer(c_int) function ice_gpu_fill_ghosts(ctx) bind(C, name='ice_gpu_fill_ghosts')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ice_gpu_put_imports(ctx, b, p, imports) bind(C, name='ice_gpu_put_imports')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, imports
      integer(c_int), value :: b, p
    end function
    integer(c_int) function ice_gpu_add_births(ctx, p, n, cells, pk) bind(C, name='ice_gpu_add_births')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: p, n
      integer(c_int32_t) :: cells(*)
      real(c_double) :: pk(*)
    end function
    integer(c_int) function ice_gpu_set_mode(ctx, mode) bind(C, name='ice_gpu_set_mode')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function
    integer(c_int) function ice_gpu_set_surrogate(ctx, p, mode) bind(C, name='ice_gpu_set_surrogate')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: p, mode
    end function
    integer(c_int) function ice_gpu_set_counters(ctx, iter, time, res_reference) bind(C, name='ice_gpu_set_counters')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: iter
      real(c_double), value :: time, res_reference
    end function
    integer(c_int) function ice_gpu_get_counters(ctx, iter, time, res_reference) bind(C, name='ice_gpu_get_counters')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: iter
      real(c_double) :: time, res_reference
    end function
    integer(c_int) function ice_gpu_last_error(ctx, buf, len) bind(C, name='ice_gpu_last_error')
      import :: c_int, c_char, c_ptr
      type(c_ptr), value :: ctx
      character(kind=c_char) :: buf(*)
      integer(c_int), value :: len
    end function
    subroutine ice_gpu_destroy(ctx) bind(C, name='ice_gpu_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
  end interface

contains

  !> The context's last message as a Fortran string
  function ice_gpu_error_text(ctx) result(s)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: s
    character(kind=c_char) :: buf(1024)
    integer :: n, rc
    buf = c_null_char
    rc = ice_gpu_last_error(ctx, buf, 1024)
    n = 0
    do while (n < 1024)
      if (buf(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: s)
    s = transfer(buf(1:n), s)
  end function

end module ice_gpu_m
