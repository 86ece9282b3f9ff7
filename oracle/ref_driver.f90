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

  ! IsPeriod = simplePeriod and Wall_method (the geometry flags of
  ! STEPPER_Stokesian, stepper.f90:54-70 and STEPPER_conf:203-219).
  subroutine ref_set_geometry(is_period_in, wall_in) bind(C, name="ref_set_geometry")
    use method, only: IsPeriod, simplePeriod, Wall_method
    use Mobilitymatrix, only: Hydrosave, Brownsave
    integer(c_int), value :: is_period_in, wall_in
    IsPeriod = is_period_in /= 0
    simplePeriod = is_period_in /= 0
    Wall_method = wall_in /= 0
    Hydrosave = .false.
    Brownsave = .false.
  end subroutine ref_set_geometry

  ! BROWN and kBT (BROWN_proerty), and the intrinsic generator's seed: every
  ! element of the seed array set to `seed` (as INIT_READ_CONTROL does with
  ! INP, initial_conf.f90:176-179).
  subroutine ref_set_brownian(on, kbt_in, seed) bind(C, name="ref_set_brownian")
    use BROWN_proerty, only: BROWN, kBT
    integer(c_int), value :: on, seed
    real(c_double), value :: kbt_in
    integer :: k
    integer, allocatable :: sd(:)
    BROWN = on /= 0
    kBT = kbt_in
    call random_seed(size=k)
    allocate(sd(k))
    sd = seed
    call random_seed(put=sd)
  end subroutine ref_set_brownian

  ! CHOLESKY (matrices.f90:126-157): lower factor in place, strict upper zeroed.
  subroutine ref_cholesky(n, a) bind(C, name="ref_cholesky")
    integer(c_int), value :: n
    real(c_double), intent(inout) :: a(n, n)
    call CHOLESKY(a, n)
  end subroutine ref_cholesky

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

  ! YAMAKAWA_IJ (GRPY_YZD.f90:768-786): the Rotne-Prager-Yamakawa blocks of an
  ! overlapping pair (the free-space block with correction_method = 1).
  subroutine ref_rp_yamakawa(r, ai, aj, tt, rr, tr, qq, gq, hq) bind(C, name="ref_rp_yamakawa")
    real(c_double), intent(in) :: r(3)
    real(c_double), value :: ai, aj
    real(c_double), intent(out) :: tt(3, 3), rr(3, 3), tr(3, 3), qq(5, 5), gq(3, 5), hq(3, 5)
    external :: YAMAKAWA_IJ
    call YAMAKAWA_IJ(tt, rr, tr, qq, gq, hq, r, ai, aj)
  end subroutine ref_rp_yamakawa

  ! overlap_Yamakawa_CORRECTION_IJ (GRPY_YZD.f90:789-811): Yamakawa minus
  ! free-space RPY, the periodic sum's overlap correction with correction_method = 1.
  subroutine ref_overlap_yamakawa(r, ai, aj, tt, rr, tr, qq, gq, hq) bind(C, name="ref_overlap_yamakawa")
    real(c_double), intent(in) :: r(3)
    real(c_double), value :: ai, aj
    real(c_double), intent(out) :: tt(3, 3), rr(3, 3), tr(3, 3), qq(5, 5), gq(3, 5), hq(3, 5)
    external :: overlap_Yamakawa_CORRECTION_IJ
    call overlap_Yamakawa_CORRECTION_IJ(tt, rr, tr, qq, gq, hq, r, ai, aj)
  end subroutine ref_overlap_yamakawa

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

  ! GRPERY_INV_FRI (GRPY_YZD.f90:14-264), the free-space grand mobility, on
  ! the FTS path (Hydrosave and Brownsave false, as main.f90:118-119 leaves them).
  subroutine ref_grpery_inv_fri(n, pos, radii_in, app, apq, aqq) bind(C, name="ref_grpery_inv_fri")
    use Mobilitymatrix, only: Hydrosave, Brownsave
    integer(c_int), value :: n
    real(c_double), intent(in) :: pos(3, n), radii_in(n)
    real(c_double), intent(out) :: app(6*n, 6*n), apq(6*n, 5*n), aqq(5*n, 5*n)
    Hydrosave = .false.
    Brownsave = .false.
    call GRPERY_INV_FRI(app, apq, aqq, pos, radii_in, n)
  end subroutine ref_grpery_inv_fri

  ! Wall_GRPERY_INV_FRI (Wall_interaction.f90:29-149): TPP(3N,3N).
  subroutine ref_wall_grpery_inv_fri(n, pos, radii_in, tpp) bind(C, name="ref_wall_grpery_inv_fri")
    use Wall_interaction, only: Wall_GRPERY_INV_FRI
    integer(c_int), value :: n
    real(c_double), intent(in) :: pos(3, n), radii_in(n)
    real(c_double), intent(out) :: tpp(3*n, 3*n)
    call Wall_GRPERY_INV_FRI(tpp, pos, radii_in, n)
  end subroutine ref_wall_grpery_inv_fri

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

  ! STEPPER_Stokesian (stepper.f90:33-156) on the periodic FTS path, replayed
  ! call for call through the reference's public and external routines so that
  ! the stresslet SijN_hyd and U_par of the last sub-step, which the real
  ! routine keeps to itself, come out too.  INIT_u_bg first, as main.f90:193.
  ! Private STEPPER_NOGmres is replayed as in ref_nogmres (stepper.f90:369-407)
  ! and private STEPPER_conf as stepper.f90:203-221; Rheology's SijN_FP and
  ! SijN_B are zero on this path (Fe_inter = 0, u_brown = 0).
  subroutine ref_step_replay(n, conf, radii_in, dt, t, yeta_mu, u_pos, u_par, sijn_hyd, z_out, sijn_b_out) &
      bind(C, name="ref_step_replay")
    use hydro_tools, only: INIT_u_bg, INIT_uo_bg, pos_collision_judge, ppiclf_collision_timestep, yetamu_solve
    use period_bdy_tools, only: SET_LATTICE_SKEW, PERIOD_CORRECTION
    use hydro_lub_mod, only: lubmxcalc
    use Wall_interaction, only: Wall_GRPERY_INV_FRI
    use method, only: IsPeriod, simplePeriod, Wall_method
    use SIZE, only: Nb, Np
    use CONFIG, only: Eij
    use LATTICE_BASE, only: LB, LINERR
    use SYS_property, only: NT_DEM, DT_DEM
    use BROWN_proerty, only: BROWN, kBT
    integer(c_int), value :: n
    real(c_double), intent(inout) :: conf(3, n), u_pos(6*n)
    real(c_double), intent(in) :: radii_in(n)
    real(c_double), value :: dt, t
    real(c_double), intent(out) :: yeta_mu(5), u_par(6*n), sijn_hyd(5*n), z_out(3*n), sijn_b_out(5*n)
    ! explicit-shape automatic arrays, as the reference declares them
    ! (stepper.f90:39-46, 652-657): flang evaluates MATMUL on them the same way
    real(c_double) :: a3(3*n, 3*n), app1(6*n, 6*n), app2(6*n, 6*n), apq1(6*n, 5*n), aqq1(5*n, 5*n)
    real(c_double) :: a1(3*n, 3*n), a2(3*n, 3*n)
    real(c_double) :: z(3*n), vb(3*n), dmdx(3*n), v_dm(3*n), ub(6*n), conf1(3, n), conf2(3, n), ga, gb, lat2(3, 3)
    real(c_double), parameter :: pi_ = 3.141592653589793d0, deta = 0.001d0
    real(c_double), allocatable :: app(:, :), apq(:, :), aqq(:, :), rfu(:, :), rfe(:, :), rse(:, :)
    real(c_double), allocatable :: rpp0(:, :), rpq0(:, :), rqq0(:, :), rpp(:, :), rpq(:, :), rqq(:, :)
    real(c_double), allocatable :: appr(:, :), mu(:, :), tpp(:, :)
    real(c_double) :: ein(5*n), fe(6*n), uo_bg(6*n), dconf(3*n), sijn_fp(5*n), sijn_b(5*n), lattice(3, 3)
    logical :: coll
    integer :: k, i, j
    allocate(app(6*n, 6*n), apq(6*n, 5*n), aqq(5*n, 5*n), rfu(6*n, 6*n), rfe(6*n, 5*n), rse(5*n, 5*n))
    allocate(rpp0(6*n, 6*n), rpq0(6*n, 5*n), rqq0(5*n, 5*n), rpp(6*n, 6*n), rpq(6*n, 5*n), rqq(5*n, 5*n))
    allocate(appr(6*n, 6*n), mu(5*n, 5*n))
    call INIT_u_bg()
    call EIN_transY2(ein, Eij, n)
    call pos_collision_judge(conf, radii_in, coll)
    call ppiclf_collision_timestep(radii_in, dt, coll)
    if (simplePeriod) call SET_LATTICE_SKEW(t, lattice)
    if (.not. IsPeriod) then
      call GRPERY_INV_FRI(app, apq, aqq, conf, radii_in, n)
    else
      call PER_GRPERY_INV_FRI(app, apq, aqq, conf, radii_in, n, lattice, LINERR)
    end if
    if (Wall_method) then
      allocate(tpp(3*n, 3*n))
      call Wall_GRPERY_INV_FRI(tpp, conf, radii_in, n)
      app(1:3*n, 1:3*n) = app(1:3*n, 1:3*n) + tpp
    end if
    ! stepper.f90:72-86 (GAUSS and DMDX_CAL, private there, replayed: 638-688)
    z = 0.0_8
    ub = 0.0_8
    if (BROWN) then
      a3 = app(1:3*n, 1:3*n)
      call CHOLESKY(a3, 3*n)
      do i = 1, 3*n
        call random_number(ga)
        call random_number(gb)
        z(i) = sqrt(-2*log(ga)) * cos(2*pi_*gb)
      end do
      vb = sqrt(2.0_8*kBT) * matmul(a3, z) / dt
      forall (i = 1:3, j = 1:n)
        conf1(i, j) = conf(i, j) + 0.5_8*deta*z(3*(j-1)+i)
        conf2(i, j) = conf(i, j) - 0.5_8*deta*z(3*(j-1)+i)
      end forall
      if (.not. IsPeriod) then
        call GRPERY_INV_FRI(app1, apq1, aqq1, conf1, radii_in, n)
        call GRPERY_INV_FRI(app2, apq1, aqq1, conf2, radii_in, n)
      else
        call SET_LATTICE_SKEW(t, lat2)
        call PER_GRPERY_INV_FRI(app1, apq1, aqq1, conf1, radii_in, n, lat2, LINERR)
        call PER_GRPERY_INV_FRI(app2, apq1, aqq1, conf2, radii_in, n, lat2, LINERR)
      end if
      a1 = app1(1:3*n, 1:3*n)
      a2 = app2(1:3*n, 1:3*n)
      dmdx = matmul((a1 - a2), z)
      dmdx = dmdx / deta
      v_dm = kBT * dmdx
      ub(1:3*n) = vb(1:3*n) + v_dm(1:3*n)
    end if
    z_out = z
    sijn_hyd = 0.0_8
    sijn_b = 0.0_8
    do k = 1, NT_DEM
      call source_body_F(conf, radii_in, fe)
      rfu = 0.0_8
      rfe = 0.0_8
      rse = 0.0_8
      call lubmxcalc(n, conf, rfu, rfe, rse)
      call MobToResist_FTS(app, apq, aqq, n, rpp0, rpq0, rqq0)
      call arrangResist(n, rfu, rfe, rse, rpp0, rpq0, rqq0, rpp, rpq, rqq)
      appr = rpp
      call MATREV(appr, 6*n)
      u_par = matmul(appr, fe + matmul(rpq, ein))
      if (IsPeriod) then
        call arrangement(rpp, rpq, rqq, mu, n)
        sijn_hyd = matmul(mu, ein)
        sijn_b = matmul(transpose(rpq), ub)
      end if
      call INIT_uo_bg(conf, uo_bg)
      u_pos = u_par + uo_bg
      call INIT_uo_bg(conf, uo_bg)
      u_pos = u_par + uo_bg
      dconf = DT_DEM * u_pos(1:3*n)
      forall (i = 1:n-Nb, j = 1:3)
        conf(j, i) = conf(j, i) + dconf(3*(i-1)+j)
      end forall
      if (simplePeriod) then
        do i = 1, Np
          call PERIOD_CORRECTION(conf(:, i), LB, t)
        end do
        call INIT_uo_bg(conf, uo_bg)
        u_pos = u_par + uo_bg
      end if
    end do
    sijn_fp = 0.0_8
    sijn_b_out = sijn_b
    call yetamu_solve(app, rpp, fe, u_par, radii_in, sijn_hyd, sijn_fp, sijn_b, ein, yeta_mu)
  end subroutine ref_step_replay

  ! INIT_READ_CONTROL (initial_conf.f90:109-189) on control_file.dat in the
  ! current directory, then its globals packed in the order of the library's
  ! rsd_control: ints io_i(24), reals io_r(48).
  subroutine ref_read_control(io_i, io_r) bind(C, name="ref_read_control")
    use Init_IO_MOD, only: INIT_READ_CONTROL
    use method
    use control, only: K_WRITE, timestep, beginstep, interval, Xshear, END_TIME, BASE_DT, T0
    use CONFIG, only: u_bg, omega_bg, EI_bg, radii_test, radii_true
    use LATTICE_BASE, only: LINERR, muCalculate
    use FORCE_PAR, only: LJ_EPS, LJ_SIGMA, LJ_CUT
    use SYS_property
    use DLVO_property
    use BROWN_proerty, only: BROWN, kBT
    use filament, only: filament_solve_implicit
    integer(c_int), intent(out) :: io_i(24)
    real(c_double), intent(out) :: io_r(48)
    io_i = 0
    io_r = 0.0_8
    kBT = 0.0_8
    call INIT_READ_CONTROL()
    io_i(1:18) = [correction_method, l2i(FTS_method), l2i(filament_solve_implicit), l2i(uselub), l2i(useDEM), &
                  l2i(usecollision), l2i(useDEMstepper), l2i(useGMRESmethod), l2i(Xshear), l2i(IsPeriod), &
                  l2i(simplePeriod), l2i(Oscillation_shear), l2i(fix), IsCaseLattice, beginstep, interval, &
                  NUM_THREADS, timestep]
    io_i(19:24) = [collision_condition, l2i(muCalculate), K_WRITE, l2i(Boundary), l2i(wall_method), l2i(useDLVO)]
    io_r(1:3) = [radii_test, radii_true, g_gravity]
    io_r(4:6) = u_bg
    io_r(7:9) = omega_bg
    io_r(10:15) = EI_bg
    io_r(16:25) = [mu_f, pho_f, Re_number, Gamma_alpha, Youngs_module, Poisson_ratio, ksp, erest, pho_par, LINERR]
    io_r(26:31) = [T0, END_TIME, BASE_DT, LJ_EPS, LJ_SIGMA, LJ_CUT]
    io_r(32:38) = [DebyeLength, z0, Lambda_l, Permittivity, Potential, AH, Surfaceenergy]
    io_r(39) = kBT
    io_r(40) = merge(1.0_8, 0.0_8, BROWN)
  contains
    integer function l2i(x)
      logical, intent(in) :: x
      l2i = merge(1, 0, x)
    end function l2i
  end subroutine ref_read_control

  ! dimentionless_calc / dimentionless_translate (mymodules.f90:808-881) on the
  ! globals ref_read_control left, and frequency / GAMMA as main.f90:148-155:
  ! lambda_time, mu_f, gravity, END_TIME, BASE_DT, T0, dt_dem0, kBT, frequency, GAMMA.
  subroutine ref_nondimensionalise(out) bind(C, name="ref_nondimensionalise")
    use dimentionless, only: dimentionless_calc, dimentionless_translate
    use lambda_Dim, only: lambda_time
    use control, only: END_TIME, BASE_DT, T0, dt_dem0
    use SYS_property, only: mu_f, gravity, GAMMA, GAMMA_alpha, frequency
    use BROWN_proerty, only: kBT
    real(c_double), intent(out) :: out(10)
    call dimentionless_calc()
    call dimentionless_translate()
    frequency = 1.d0 / lambda_time
    GAMMA = GAMMA_alpha * frequency
    out = [lambda_time, mu_f, gravity, END_TIME, BASE_DT, T0, dt_dem0, kBT, frequency, GAMMA]
  end subroutine ref_nondimensionalise

  ! INITIAL_SIZE + INITIAL_CONF (initial_conf.f90:356-476) on initial_config.dat
  ! in the current directory, IsCaseLattice = 0, no boundary particles.
  subroutine ref_read_config(nmax, n_out, lb_out, t_out, conf_out, radii_out, upos_out) &
      bind(C, name="ref_read_config")
    use Init_IO_MOD, only: INITIAL_SIZE, INITIAL_CONF
    use method, only: IsCaseLattice, Boundary
    use SIZE, only: NN, Np
    use CONFIG, only: CONF_ => CONF, RADII, Floc_index
    use LATTICE_BASE, only: LB
    integer(c_int), value :: nmax
    integer(c_int), intent(out) :: n_out
    real(c_double), intent(out) :: lb_out(3), t_out, conf_out(3, nmax), radii_out(nmax), upos_out(6*nmax)
    real(c_double), allocatable :: upos(:)
    IsCaseLattice = 0
    Boundary = .false.
    call INITIAL_SIZE()
    n_out = NN
    lb_out = LB
    if (NN > nmax) return
    if (allocated(CONF_)) deallocate(CONF_)
    if (allocated(RADII)) deallocate(RADII)
    if (allocated(Floc_index)) deallocate(Floc_index)
    allocate(upos(6*NN))
    upos = 0.0_8
    call INITIAL_CONF(t_out, upos)
    conf_out(:, 1:NN) = CONF_(:, 1:NN)
    radii_out(1:NN) = RADII(1:NN)
    upos_out(1:6*NN) = upos
  end subroutine ref_read_config

  ! VMD_WRITE and yetamu_WRITE (initial_conf.f90:250-292) into output.VTF and
  ! yeta_mu.txt in the current directory (opened by INIT_VMD_WRITE, closed by
  ! END_VMD_WRITE; DIM_property.txt's unit closed too).
  subroutine ref_io_open() bind(C, name="ref_io_open")
    use Init_IO_MOD, only: INIT_VMD_WRITE
    call INIT_VMD_WRITE()
  end subroutine ref_io_open

  subroutine ref_vmd_write(n, conf, t, radii_in, u_pos) bind(C, name="ref_vmd_write")
    use Init_IO_MOD, only: VMD_WRITE
    integer(c_int), value :: n
    real(c_double), intent(in) :: conf(3, n), radii_in(n), u_pos(6*n)
    real(c_double), value :: t
    call VMD_WRITE(conf, t, radii_in, u_pos)
  end subroutine ref_vmd_write

  subroutine ref_yetamu_write(t, yeta, lb1, freq, gam, gangle) bind(C, name="ref_yetamu_write")
    use Init_IO_MOD, only: yetamu_WRITE
    real(c_double), value :: t, lb1, freq, gam, gangle
    real(c_double), intent(in) :: yeta(5)
    call yetamu_WRITE(t, yeta, lb1, freq, gam, gangle)
  end subroutine ref_yetamu_write

  subroutine ref_io_close() bind(C, name="ref_io_close")
    use Init_IO_MOD, only: END_VMD_WRITE
    call END_VMD_WRITE()
    close(51)
  end subroutine ref_io_close

  subroutine ref_get_ntdem(nt, dtdem) bind(C, name="ref_get_ntdem")
    use SYS_property, only: Nt_DEM, DT_DEM
    integer(c_int), intent(out) :: nt
    real(c_double), intent(out) :: dtdem
    nt = Nt_DEM
    dtdem = DT_DEM
  end subroutine ref_get_ntdem

end module ref_driver
