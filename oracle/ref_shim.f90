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

  real(c_double) function ref_blowing(gam, Rg, mug, kg, rhop, dp, mp, mdot) bind(C, name='ref_blowing')
    real(c_double), value :: gam, Rg, mug, kg, rhop, dp, mp, mdot
    ref_blowing = blowingFactor(gam, Rg, mug, kg, rhop, dp, mp, mdot)
  end function ref_blowing

end module ice_ref_shim
