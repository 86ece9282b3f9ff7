! ref_driver.f90 -- C-callable driver around the flang-built reference.
!
! TEST INFRASTRUCTURE ONLY (the oracle's checker).  oracle/Makefile compiles
! the reference's own Fortran sources where they lie under /root/reference
! and links them with this driver into oracle/_ref/libref.so.  The driver
! sets the reference's module globals the way main.f90 does (main.f90:39-60,
! 193-195) and exposes its public routines through bind(C) entry points so
! that Python tests can call the real reference on seeded inputs.
!
! STEPPER_NOGmres and Rheology are private to STEPPER_MOD
! (stepper.f90:26-29), so ref_nogmres replays stepper.f90:369-407 through
! the reference's public/external routines.
module ref_driver
  use iso_c_binding
  implicit none
contains

  subroutine ref_init() bind(C, name="ref_init")
    call INIT_KDirac()
    call INIT_EPS()
    call INIT_Y2()
    call INIT_II()
  end subroutine ref_init

  ! Global state for the periodic FTS path (control flags of the hot path).
  subroutine ref_setup(n, nb, conf, radii_in, lb_in, gamma_in, linerr_in, mu_in, corr, grav, dtdem0) &
      bind(C, name="ref_setup")
    use SIZE, only: NN, Nb_ => Nb, Np, Nswimer, Nfilament
    use CONFIG, only: CONF_ => CONF, RADII, Floc_index, u_bg, omega_bg, EI_bg
    use LATTICE_BASE, only: LB, LINERR
    use SYS_property, only: GAMMA, mu_f, gravity
    use method
    use control, only: dt_dem0
    use filament, only: F_rb
    use rb_conglomerate, only: K_rb
    use BROWN_proerty, only: BROWN
    integer(c_int), value :: n, nb, corr
    real(c_double), intent(in) :: conf(3, n), radii_in(n), lb_in(3)
    real(c_double), value :: gamma_in, linerr_in, mu_in, grav, dtdem0
    NN = n
    Nb_ = nb
    Np = n - nb
    Nswimer = 0
    Nfilament = 0
    if (allocated(CONF_)) deallocate(CONF_)
    if (allocated(RADII)) deallocate(RADII)
    if (allocated(Floc_index)) deallocate(Floc_index)
    allocate(CONF_(3, n), RADII(n), Floc_index(n))
    CONF_ = conf
    RADII = radii_in
    Floc_index = 1
    if (nb > 0) Floc_index(n - nb + 1:n) = 0
    LB = lb_in
    GAMMA = gamma_in
    LINERR = linerr_in
    mu_f = mu_in
    gravity = grav
    dt_dem0 = dtdem0
    correction_method = corr
    IsPeriod = .true.
    simplePeriod = .true.
    FTS_method = .true.
    uselub = .true.
    useDEM = .false.
    useCollision = .false.
    useDEMstepper = .false.
    useGMRESmethod = .false.
    Oscillation_shear = .false.
    fix = .false.
    Boundary = .false.
    Wall_method = .false.
    BROWN = .false.
    F_rb = 0
    K_rb = 0
    u_bg = 0.0_8
    omega_bg = 0.0_8
    EI_bg = 0.0_8
  end subroutine ref_setup

  subroutine ref_set_positions(n, conf) bind(C, name="ref_set_positions")
    use CONFIG, only: CONF_ => CONF
    integer(c_int), value :: n
    real(c_double), intent(in) :: conf(3, n)
    CONF_ = conf
  end subroutine ref_set_positions

  subroutine ref_set_lattice_skew(t, lattice) bind(C, name="ref_set_lattice_skew")
    use period_bdy_tools, only: SET_LATTICE_SKEW
    real(c_double), value :: t
    real(c_double), intent(out) :: lattice(3, 3)
    call SET_LATTICE_SKEW(t, lattice)
  end subroutine ref_set_lattice_skew

  subroutine ref_calc_lattice_inverse(ews_err) bind(C, name="ref_calc_lattice_inverse")
    use period_bdy_tools, only: CALC_LATTICE_INVERSE
    use LATTICE_BASE, only: LR
    real(c_double), value :: ews_err
    call CALC_LATTICE_INVERSE(LR, ews_err)
  end subroutine ref_calc_lattice_inverse

  subroutine ref_get_lattice(lr_o, li_o, lv_o, sig_o, lam_o, logerr_o, phi_o, nr_o, ni_o) &
      bind(C, name="ref_get_lattice")
    use LATTICE_BASE
    real(c_double), intent(out) :: lr_o(3, 3), li_o(3, 3), lv_o, sig_o, lam_o, logerr_o, phi_o
    integer(c_int), intent(out) :: nr_o, ni_o
    lr_o = LR
    li_o = LI
    lv_o = LV
    sig_o = SIG
    lam_o = Lammda
    logerr_o = LOGERR
    phi_o = phi
    nr_o = NR
    ni_o = NI
  end subroutine ref_get_lattice

  subroutine ref_set_ewald(lam_i, phi_i, nr_i, ni_i) bind(C, name="ref_set_ewald")
    use LATTICE_BASE
    real(c_double), value :: lam_i, phi_i
    integer(c_int), value :: nr_i, ni_i
    Lammda = lam_i
    phi = phi_i
    NR = nr_i
    NI = ni_i
  end subroutine ref_set_ewald

  ! Flushes the Fortran runtime's buffered standard output (unit 6).
  subroutine ref_flush() bind(C, name="ref_flush")
    flush(6)
  end subroutine ref_flush

  ! CONFIG::radii_test (control_file.dat, initial_conf.f90:137): force scale of averageU.
  subroutine ref_set_radii_test(x) bind(C, name="ref_set_radii_test")
    use CONFIG, only: radii_test
    real(c_double), value :: x
    radii_test = x
  end subroutine ref_set_radii_test

  subroutine ref_per_skew_correction(r, dmin) bind(C, name="ref_per_skew_correction")
    use period_bdy_tools, only: PER_SKEW_CORRECTION
    real(c_double), intent(inout) :: r(3)
    real(c_double), intent(out) :: dmin
    call PER_SKEW_CORRECTION(r, dmin)
  end subroutine ref_per_skew_correction

  subroutine ref_period_correction(r, t) bind(C, name="ref_period_correction")
    use period_bdy_tools, only: PERIOD_CORRECTION
    use LATTICE_BASE, only: LB
    real(c_double), intent(inout) :: r(3)
    real(c_double), value :: t
    call PERIOD_CORRECTION(r, LB, t)
  end subroutine ref_period_correction

  subroutine ref_rp_self(a, tts, rrs, qqs) bind(C, name="ref_rp_self")
    use Ewald_summation, only: PER_ROTNE_PRAGER_SELF
    use LATTICE_BASE, only: LOGERR
    real(c_double), value :: a
    real(c_double), intent(out) :: tts(3, 3), rrs(3, 3), qqs(5, 5)
    call PER_ROTNE_PRAGER_SELF(tts, rrs, qqs, a, LOGERR)
  end subroutine ref_rp_self

  subroutine ref_rp_ij(r, ai, aj, tt, rr, tr, qq, gq, hq) bind(C, name="ref_rp_ij")
    use Ewald_summation, only: PER_ROTNE_PRAGER_IJ
    use LATTICE_BASE, only: LOGERR
    real(c_double), intent(in) :: r(3)
    real(c_double), value :: ai, aj
    real(c_double), intent(out) :: tt(3, 3), rr(3, 3), tr(3, 3), qq(5, 5), gq(3, 5), hq(3, 5)
    call PER_ROTNE_PRAGER_IJ(tt, rr, tr, qq, gq, hq, r, ai, aj, LOGERR)
  end subroutine ref_rp_ij

  subroutine ref_rp_regular(r, ai, aj, tt, rr, tr, qq, gq, hq) bind(C, name="ref_rp_regular")
    use Regularization, only: Regular_ROTNE_PRAGER_IJ
    real(c_double), intent(in) :: r(3)
    real(c_double), value :: ai, aj
    real(c_double), intent(out) :: tt(3, 3), rr(3, 3), tr(3, 3), qq(5, 5), gq(3, 5), hq(3, 5)
    call Regular_ROTNE_PRAGER_IJ(tt, rr, tr, qq, gq, hq, r, ai, aj)
  end subroutine ref_rp_regular

  subroutine ref_rp_free_fts(r, ai, aj, tt, rr, tr, qq, gq, hq) bind(C, name="ref_rp_free_fts")
    use stokesian_Green, only: ROTNE_PRAGER_IJ_FTS
    real(c_double), intent(in) :: r(3)
    real(c_double), value :: ai, aj
    real(c_double), intent(out) :: tt(3, 3), rr(3, 3), tr(3, 3), qq(5, 5), gq(3, 5), hq(3, 5)
    call ROTNE_PRAGER_IJ_FTS(tt, rr, tr, qq, gq, hq, r, ai, aj)
  end subroutine ref_rp_free_fts

  subroutine ref_per_grpery_inv_fri(n, pos, radii_in, ews_err, app, apq, aqq) &
      bind(C, name="ref_per_grpery_inv_fri")
    use LATTICE_BASE, only: LR
    integer(c_int), value :: n
    real(c_double), intent(in) :: pos(3, n), radii_in(n)
    real(c_double), value :: ews_err
    real(c_double), intent(out) :: app(6*n, 6*n), apq(6*n, 5*n), aqq(5*n, 5*n)
    real(c_double) :: lattice(3, 3)
    lattice = LR
    call PER_GRPERY_INV_FRI(app, apq, aqq, pos, radii_in, n, lattice, ews_err)
  end subroutine ref_per_grpery_inv_fri

  subroutine ref_lubmxcalc(n, pos, rfu, rfe, rse) bind(C, name="ref_lubmxcalc")
    use hydro_lub_mod, only: lubmxcalc
    integer(c_int), value :: n
    real(c_double), intent(in) :: pos(3, n)
    real(c_double), intent(out) :: rfu(6*n, 6*n), rfe(6*n, 5*n), rse(5*n, 5*n)
    call lubmxcalc(n, pos, rfu, rfe, rse)
  end subroutine ref_lubmxcalc

  subroutine ref_matrev(n, a) bind(C, name="ref_matrev")
    integer(c_int), value :: n
    real(c_double), intent(inout) :: a(n, n)
    call MATREV(a, n)
  end subroutine ref_matrev

  subroutine ref_mob_to_resist_fts(n, app, apq, aqq, rpp, rpq, rqq) bind(C, name="ref_mob_to_resist_fts")
    integer(c_int), value :: n
    real(c_double), intent(in) :: app(6*n, 6*n), apq(6*n, 5*n), aqq(5*n, 5*n)
    real(c_double), intent(out) :: rpp(6*n, 6*n), rpq(6*n, 5*n), rqq(5*n, 5*n)
    call MobToResist_FTS(app, apq, aqq, n, rpp, rpq, rqq)
  end subroutine ref_mob_to_resist_fts

  subroutine ref_arrangement(n, rpp, rpq, rqq, mu) bind(C, name="ref_arrangement")
    integer(c_int), value :: n
    real(c_double), intent(in) :: rpp(6*n, 6*n), rpq(6*n, 5*n), rqq(5*n, 5*n)
    real(c_double), intent(out) :: mu(5*n, 5*n)
    call arrangement(rpp, rpq, rqq, mu, n)
  end subroutine ref_arrangement

  subroutine ref_ein(n, ein) bind(C, name="ref_ein")
    use CONFIG, only: Eij
    use hydro_tools, only: INIT_u_bg
    integer(c_int), value :: n
    real(c_double), intent(out) :: ein(5*n)
    call INIT_u_bg()
    call EIN_transY2(ein, Eij, n)
  end subroutine ref_ein

  ! stepper.f90:369-407 (STEPPER_NOGmres) with Fe given, u_brown = 0 and
  ! Fe_inter = 0 as STEPPER_Stokesian passes them (stepper.f90:95-99).
  subroutine ref_nogmres(n, app, apq, aqq, rfu, rfe, rse, fe, ein, u_par, sijn_hyd, rpp_out) &
      bind(C, name="ref_nogmres")
    integer(c_int), value :: n
    real(c_double), intent(in) :: app(6*n, 6*n), apq(6*n, 5*n), aqq(5*n, 5*n)
    real(c_double), intent(in) :: rfu(6*n, 6*n), rfe(6*n, 5*n), rse(5*n, 5*n), fe(6*n), ein(5*n)
    real(c_double), intent(out) :: u_par(6*n), sijn_hyd(5*n), rpp_out(6*n, 6*n)
    real(c_double), allocatable :: rpp0(:, :), rpq0(:, :), rqq0(:, :), rpq(:, :), rqq(:, :), appr(:, :), mu(:, :)
    allocate(rpp0(6*n, 6*n), rpq0(6*n, 5*n), rqq0(5*n, 5*n), rpq(6*n, 5*n), rqq(5*n, 5*n))
    allocate(appr(6*n, 6*n), mu(5*n, 5*n))
    call MobToResist_FTS(app, apq, aqq, n, rpp0, rpq0, rqq0)
    call arrangResist(n, rfu, rfe, rse, rpp0, rpq0, rqq0, rpp_out, rpq, rqq)
    appr = rpp_out
    call MATREV(appr, 6*n)
    u_par = matmul(appr, fe + matmul(rpq, ein))
    call arrangement(rpp_out, rpq, rqq, mu, n)
    sijn_hyd = matmul(mu, ein)
  end subroutine ref_nogmres

  ! One full reference step: main.f90:193-195 (INIT_u_bg, STEPPER_Stokesian).
  subroutine ref_stepper_stokesian(n, conf, radii_in, dt, t, yeta_mu, u_pos) &
      bind(C, name="ref_stepper_stokesian")
    use STEPPER_MOD, only: STEPPER_Stokesian
    use hydro_tools, only: INIT_u_bg
    integer(c_int), value :: n
    real(c_double), intent(inout) :: conf(3, n), u_pos(6*n)
    real(c_double), intent(in) :: radii_in(n)
    real(c_double), value :: dt, t
    real(c_double), intent(out) :: yeta_mu(5)
    real(c_double) :: sijn(5*n)
    call INIT_u_bg()
    call STEPPER_Stokesian(conf, radii_in, dt, t, yeta_mu, u_pos, sijn)
  end subroutine ref_stepper_stokesian

  subroutine ref_get_ntdem(nt, dtdem) bind(C, name="ref_get_ntdem")
    use SYS_property, only: Nt_DEM, DT_DEM
    integer(c_int), intent(out) :: nt
    real(c_double), intent(out) :: dtdem
    nt = Nt_DEM
    dtdem = DT_DEM
  end subroutine ref_get_ntdem

end module ref_driver
