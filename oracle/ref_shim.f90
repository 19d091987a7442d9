! ORACLE -- test infrastructure only.  A bind(C) shim over the reference's own
! Fortran leaf modules (compiled from /root/reference by build_ref.sh into
! oracle/_ref/libiceref.so), so tests can evaluate the reference's closures,
! Riemann solvers, limiters, reconstruction, sensor, drag/heat laws, RK stage
! and evaporation on the inputs the CPU restatement sees.  This file only
! calls the reference; it contains none of its arithmetic.
module ice_ref_shim
  use, intrinsic :: iso_c_binding
  use ICE_Config_Types_m,     only: condensed_phase_t, obj_time_scheme
  use ICE_Global_m,           only: solid_of
  use ICE_Lib_Model
  use ICE_Lib_Riemann,        only: assign_riemann, riemann
  use ICE_Lib_Limiters,       only: assign_limiter, limiter
  use ICE_Lib_Reconstruction, only: state_reconstruction, bad_recon
  use ICE_Lib_Shock_Detector, only: SD_rho
  use ICE_Lib_Drag,           only: drag
  use ICE_Lib_Heat,           only: heat
  use ICE_Lib_RK,             only: RK_stage
  use ICE_Lib_Evaporation,    only: evaporation, blowingFactor, nep
  use ICE_Lib_Properties,     only: mat_rho, mat_e, mat_T_from_e, mat_T_from_rhoe, mat_lv, mat_cpv, mat_emiss, &
                                    mat_psat
  implicit none
  type(condensed_phase_t), save :: mat
  character(len=16), parameter :: lim_words(0:10) = [character(len=16) :: 'IORD', 'minmod', 'vanalbada', &
      'vanleer', 'ospre', 'umist', 'osher', 'sweby', 'mc', 'koren', 'superbee']

contains

  ! closure model (0 MK, 1 IG, 2 AG) and constant material properties
  subroutine ref_setup(model, rho, cs, lv, emiss) bind(C, name='ref_setup')
    integer(c_int), value :: model
    real(c_double), value :: rho, cs, lv, emiss
    if (.not. allocated(obj_time_scheme%model)) allocate(obj_time_scheme%model(1))
    if (.not. allocated(solid_of)) allocate(solid_of(1))
    solid_of(1) = .false.
    select case (model)
    case (0); obj_time_scheme%model(1) = 'MK'
    case (1); obj_time_scheme%model(1) = 'IG'
    case default; obj_time_scheme%model(1) = 'AG'
    end select
    call assign_all(1)
    mat%rho_al = rho; mat%cs_al = cs; mat%lv_al = lv; mat%emiss = emiss
    mat%use_table = .false.; mat%cs_varies = .false.; mat%rho_varies = .false.
  end subroutine ref_setup

  ! a property table on T = t_min .. t_min+n-1 in ice_gpu_set_properties' layout: columns
  ! density, cp, e, Psat, Lv, Cpv, emissivity; flags 1 rho varies, 2 cs varies, 4 Psat,
  ! 8 Lv, 16 Cpv, 32 emissivity
  subroutine ref_setup_table(t_min, n, cols, flags) bind(C, name='ref_setup_table')
    integer(c_int), value :: t_min, n, flags
    real(c_double), intent(in) :: cols(n, 7)
    integer :: hi
    hi = t_min + n - 1
    if (allocated(mat%rho_tab)) deallocate(mat%rho_tab, mat%cs_tab, mat%e_tab, mat%psat_tab, mat%lv_tab, &
                                           mat%cpv_tab, mat%emiss_tab)
    allocate(mat%rho_tab(t_min:hi), source=cols(:, 1))
    allocate(mat%cs_tab(t_min:hi), source=cols(:, 2))
    allocate(mat%e_tab(t_min:hi), source=cols(:, 3))
    allocate(mat%psat_tab(t_min:hi), source=cols(:, 4))
    allocate(mat%lv_tab(t_min:hi), source=cols(:, 5))
    allocate(mat%cpv_tab(t_min:hi), source=cols(:, 6))
    allocate(mat%emiss_tab(t_min:hi), source=cols(:, 7))
    mat%T_min = t_min; mat%T_max = hi
    mat%use_table  = .true.
    mat%rho_varies = iand(flags, 1) /= 0
    mat%cs_varies  = iand(flags, 2) /= 0
    mat%use_psat   = iand(flags, 4) /= 0
    mat%use_lv     = iand(flags, 8) /= 0
    mat%use_cpv    = iand(flags, 16) /= 0
    mat%use_emiss  = iand(flags, 32) /= 0
  end subroutine ref_setup_table

  ! one property of the table at T: 1 density, 4 Psat, 5 Lv, 6 Cpv, 7 emissivity, as the
  ! reference's mat_* functions return them; 3 the energy mat_e
  real(c_double) function ref_mat(col, T) bind(C, name='ref_mat')
    integer(c_int), value :: col
    real(c_double), value :: T
    select case (col)
    case (1); ref_mat = mat_rho(mat, T)
    case (3); ref_mat = mat_e(mat, T)
    case (4); ref_mat = mat_psat(mat, T)
    case (5); ref_mat = mat_lv(mat, T)
    case (6); ref_mat = mat_cpv(mat, T)
    case default; ref_mat = mat_emiss(mat, T)
    end select
  end function ref_mat

  real(c_double) function ref_mat_T_from_e(e) bind(C, name='ref_mat_T_from_e')
    real(c_double), value :: e
    ref_mat_T_from_e = mat_T_from_e(mat, e)
  end function ref_mat_T_from_e

  real(c_double) function ref_mat_T_from_rhoe(rhoe, rho, eps) bind(C, name='ref_mat_T_from_rhoe')
    real(c_double), value :: rhoe, rho, eps
    ref_mat_T_from_rhoe = mat_T_from_rhoe(mat, rhoe, rho, eps)
  end function ref_mat_T_from_rhoe

  subroutine ref_p2c(nc, prim, cons) bind(C, name='ref_p2c')
    integer(c_int), value :: nc
    real(c_double), intent(in) :: prim(nc)
    real(c_double), intent(out) :: cons(nc)
    cons = prim_2_cons(prim, mat)
  end subroutine ref_p2c

  subroutine ref_c2p(nc, cons, prim) bind(C, name='ref_c2p')
    integer(c_int), value :: nc
    real(c_double), intent(in) :: cons(nc)
    real(c_double), intent(out) :: prim(nc)
    prim = cons_2_prim(cons, mat)
  end subroutine ref_c2p

  integer(c_int) function ref_check(nc, prim) bind(C, name='ref_check')
    integer(c_int), value :: nc
    real(c_double), intent(in) :: prim(nc)
    ref_check = merge(1, 0, check_prim(prim))
  end function ref_check

  real(c_double) function ref_wavespeed(nc, prim, nrm) bind(C, name='ref_wavespeed')
    integer(c_int), value :: nc
    real(c_double), intent(in) :: prim(nc), nrm(3)
    ref_wavespeed = wavespeed_make(prim, nrm)
  end function ref_wavespeed

  subroutine ref_flux(nc, prim, nrm, f) bind(C, name='ref_flux')
    integer(c_int), value :: nc
    real(c_double), intent(in) :: prim(nc), nrm(3)
    real(c_double), intent(out) :: f(nc)
    f = flux_make(prim, nrm, mat)
  end subroutine ref_flux

  subroutine ref_source_make(nc, prim, force, s) bind(C, name='ref_source_make')
    integer(c_int), value :: nc
    real(c_double), intent(in) :: prim(nc), force(6)
    real(c_double), intent(out) :: s(nc)
    s = source_make(prim, force, mat)
  end subroutine ref_source_make

  subroutine ref_riemann(solver, nc, p1, p4, nrm, f) bind(C, name='ref_riemann')
    integer(c_int), value :: solver, nc
    real(c_double), intent(in) :: p1(nc), p4(nc), nrm(3)
    real(c_double), intent(out) :: f(nc)
    select case (solver)
    case (0); call assign_riemann('Saurel', .false.)
    case (1); call assign_riemann('Rusanov')
    case default; call assign_riemann('HLLE')
    end select
    f = riemann(p1, p4, nrm, mat)
  end subroutine ref_riemann

  real(c_double) function ref_limiter(id, x, y, cx, cy) bind(C, name='ref_limiter')
    integer(c_int), value :: id
    real(c_double), value :: x, y, cx, cy
    call assign_limiter(trim(lim_words(id)))
    ref_limiter = limiter(x, y, cx, cy)
  end function ref_limiter

  integer(c_int) function ref_reconstruct(lim, nc, prev, local, next, next2, dl0, dl1, dl2, dll, dlr, beta, &
                                          ql, qr) bind(C, name='ref_reconstruct')
    integer(c_int), value :: lim, nc
    real(c_double), intent(in) :: prev(nc), local(nc), next(nc), next2(nc)
    real(c_double), value :: dl0, dl1, dl2, dll, dlr, beta
    real(c_double), intent(out) :: ql(nc), qr(nc)
    call assign_limiter(trim(lim_words(lim)))
    bad_recon = .false.
    call state_reconstruction(prev, local, next, next2, dl0, dl1, dl2, dll, dlr, ql, qr, beta, mat, .false.)
    ref_reconstruct = merge(0, 1, bad_recon)
  end function ref_reconstruct

  ! rho(0:2,0:2,0:2) in Fortran order
  real(c_double) function ref_sd(rho) bind(C, name='ref_sd')
    real(c_double), intent(in) :: rho(0:2, 0:2, 0:2)
    ref_sd = SD_rho(rho)
  end function ref_sd

  real(c_double) function ref_drag(Re, Ma, G, Tr, sel) bind(C, name='ref_drag')
    real(c_double), value :: Re, Ma, G, Tr
    integer(c_int), value :: sel
    ref_drag = drag(Re, Ma, G, Tr, sel)
  end function ref_drag

  real(c_double) function ref_heat(Re, Pr, Ma, sel) bind(C, name='ref_heat')
    real(c_double), value :: Re, Pr, Ma
    integer(c_int), value :: sel
    ref_heat = heat(Re, Pr, Ma, sel)
  end function ref_heat

  subroutine ref_rk_stage(irk, nrk, nc, cons, cons_old, res, out) bind(C, name='ref_rk_stage')
    integer(c_int), value :: irk, nrk, nc
    real(c_double), intent(in) :: cons(nc), cons_old(nc), res(nc)
    real(c_double), intent(out) :: out(nc)
    out = RK_stage(irk, nrk, cons, cons_old, res)
  end subroutine ref_rk_stage

  subroutine ref_evaporation(rhog, Tg, gam, Rg, mug, kg, Tp, dp, Re, evap, intf, ep, mdot, qdot, ovr) &
      bind(C, name='ref_evaporation')
    real(c_double), value :: rhog, Tg, gam, Rg, mug, kg, Tp, dp, Re
    integer(c_int), value :: evap, intf
    real(c_double), intent(in) :: ep(nep)
    real(c_double), intent(out) :: mdot, qdot
    integer(c_int), intent(out) :: ovr
    logical :: o
    call evaporation(rhog, Tg, gam, Rg, mug, kg, Tp, dp, Re, evap, intf, ep, mdot, qdot, o)
    ovr = merge(1, 0, o)
  end subroutine ref_evaporation

  subroutine ref_evaporation_psat(rhog, Tg, gam, Rg, mug, kg, Tp, dp, Re, evap, intf, ep, psat, mdot, qdot, ovr) &
      bind(C, name='ref_evaporation_psat')
    real(c_double), value :: rhog, Tg, gam, Rg, mug, kg, Tp, dp, Re, psat
    integer(c_int), value :: evap, intf
    real(c_double), intent(in) :: ep(nep)
    real(c_double), intent(out) :: mdot, qdot
    integer(c_int), intent(out) :: ovr
    logical :: o
    call evaporation(rhog, Tg, gam, Rg, mug, kg, Tp, dp, Re, evap, intf, ep, mdot, qdot, o, psatExt=psat)
    ovr = merge(1, 0, o)
  end subroutine ref_evaporation_psat

  real(c_double) function ref_blowing(gam, Rg, mug, kg, rhop, dp, mp, mdot) bind(C, name='ref_blowing')
    real(c_double), value :: gam, Rg, mug, kg, rhop, dp, mp, mdot
    ref_blowing = blowingFactor(gam, Rg, mug, kg, rhop, dp, mp, mdot)
  end function ref_blowing

end module ice_ref_shim
