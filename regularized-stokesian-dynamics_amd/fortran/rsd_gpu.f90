! rsd_gpu.f90 -- Fortran drop-in for the periodic FTS hot path.
!
! Link this file (and hydro_lub_gpu.f90) in place of the reference's
! per_GRPY_YZD.f90 / hydro_lub.f90 and ahead of matrices.f90: the external
! subroutines below keep the reference's names and argument lists, so
! stepper.f90 and main.f90 call the MI355X library unchanged.
!
!   PER_GRPERY_INV_FRI  per_GRPY_YZD.f90:13-200  -> rsd_per_grpery_inv_fri
!   MobToResist_FTS     matrices.f90:393-416     -> rsd_mob_to_resist_fts
!   MATREV              matrices.f90:4-34        -> rsd_matrev
!   arrangement         matrices.f90:432-445     -> rsd_arrangement
!
! Hidden inputs are read from, and hidden outputs written back to, the
! reference's modules exactly as the replaced routines do: LATTICE_BASE
! (LR, LI, LV in; SIG, Lammda, LOGERR, NR, NI, phi out), method
! (correction_method), SYS_property (mu_f).  A positive return code is a
! Cholesky INFO: reported and non-fatal, as MATREV does.  A negative one is a
! library error: reported with its text, then the program stops.

module rsd_capi
  use iso_c_binding
  implicit none

  type, bind(C) :: rsd_lattice
    real(c_double) :: lr(9), li(9), lv, sig, lammda, logerr, phi
    integer(c_int) :: nr, ni
  end type rsd_lattice

  interface
    integer(c_int) function rsd_device_init(device) bind(C, name="rsd_device_init")
      import :: c_int
      integer(c_int), value :: device
    end function
    integer(c_int) function rsd_lub_tables_load(dir) bind(C, name="rsd_lub_tables_load")
      import :: c_char, c_int
      character(kind=c_char), dimension(*) :: dir
    end function
    type(c_ptr) function rsd_last_error() bind(C, name="rsd_last_error")
      import :: c_ptr
    end function
    integer(c_int) function rsd_per_grpery_inv_fri(n, pos, radii, lat, ews_err, corr, mu_f, app, apq, aqq, n_overlap) &
        bind(C, name="rsd_per_grpery_inv_fri")
      import :: c_int, c_double, rsd_lattice
      integer(c_int), value :: n, corr
      real(c_double), intent(in) :: pos(*), radii(*)
      type(rsd_lattice), intent(inout) :: lat
      real(c_double), value :: ews_err, mu_f
      real(c_double), intent(out) :: app(*), apq(*), aqq(*)
      integer(c_int), intent(out) :: n_overlap
    end function
    integer(c_int) function rsd_lubmxcalc(n, nb, is_period, pos, radii, lat, mu_f, rfu, rfe, rse, n_contact) &
        bind(C, name="rsd_lubmxcalc")
      import :: c_int, c_double, rsd_lattice
      integer(c_int), value :: n, nb, is_period
      real(c_double), intent(in) :: pos(*), radii(*)
      type(rsd_lattice), intent(in) :: lat
      real(c_double), value :: mu_f
      real(c_double), intent(out) :: rfu(*), rfe(*), rse(*)
      integer(c_int), intent(out) :: n_contact
    end function
    integer(c_int) function rsd_matrev(n, a) bind(C, name="rsd_matrev")
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double), intent(inout) :: a(*)
    end function
    integer(c_int) function rsd_mob_to_resist_fts(n, app, apq, aqq, rpp, rpq, rqq) bind(C, name="rsd_mob_to_resist_fts")
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double), intent(in) :: app(*), apq(*), aqq(*)
      real(c_double), intent(out) :: rpp(*), rpq(*), rqq(*)
    end function
    integer(c_int) function rsd_arrangement(n, rpp, rpq, rqq, mu) bind(C, name="rsd_arrangement")
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double), intent(in) :: rpp(*), rpq(*), rqq(*)
      real(c_double), intent(out) :: mu(*)
    end function
  end interface

  logical, save :: rsd_ready = .false., rsd_tables_ready = .false.

contains

  ! First use of any routine: select device 0.
  subroutine rsd_gpu_init()
    integer(c_int) :: rc
    if (rsd_ready) return
    rc = rsd_device_init(0_c_int)
    call rsd_check(rc, 'rsd_device_init')
    rsd_ready = .true.
  end subroutine rsd_gpu_init

  ! First lubrication call: the tables from the working directory's lubdat/,
  ! when and where the reference reads them (hydro_lub.f90:455-466).
  subroutine rsd_gpu_tables()
    integer(c_int) :: rc
    if (rsd_tables_ready) return
    rc = rsd_lub_tables_load('lubdat' // c_null_char)
    call rsd_check(rc, 'rsd_lub_tables_load')
    rsd_tables_ready = .true.
  end subroutine rsd_gpu_tables

  subroutine rsd_check(rc, what)
    use iso_fortran_env, only: error_unit
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: k
    if (rc >= 0) return
    call c_f_pointer(rsd_last_error(), msg, [1024])
    k = 0
    do while (k < 1024)
      if (msg(k + 1) == c_null_char) exit
      k = k + 1
    end do
    write(error_unit, *) trim(what), ' failed: ', rc, ' ', transfer(msg(1:k), repeat(' ', k))
    error stop 'rsd_gpu: library error'
  end subroutine rsd_check

  ! LATTICE_BASE -> rsd_lattice (LR as given, LI and LV from the module, as
  ! SET_LATTICE_SKEW left them)
  subroutine rsd_lattice_from_module(lattice, lat)
    use LATTICE_BASE, only: LI, LV
    real(c_double), intent(in) :: lattice(3, 3)
    type(rsd_lattice), intent(out) :: lat
    lat%lr = reshape(lattice, [9])
    lat%li = reshape(LI, [9])
    lat%lv = LV
    lat%sig = 0.0_c_double
    lat%lammda = 0.0_c_double
    lat%logerr = 0.0_c_double
    lat%phi = 0.0_c_double
    lat%nr = 0
    lat%ni = 0
  end subroutine rsd_lattice_from_module

end module rsd_capi


SUBROUTINE PER_GRPERY_INV_FRI(APP, APQ, AQQ, p_pos, RADII, NN, LATTICE, EWS_ERR)
  use rsd_capi
  use LATTICE_BASE, only: SIG, Lammda, LOGERR, NR, NI, phi
  use SYS_property, only: mu_f
  use method, only: correction_method
  implicit none
  INTEGER, intent(in) :: NN
  REAL*8, intent(out) :: APP(6*NN, 6*NN), APQ(6*NN, 5*NN), AQQ(5*NN, 5*NN)
  REAL*8, intent(in) :: p_pos(3, NN), RADII(NN), LATTICE(3, 3), EWS_ERR
  type(rsd_lattice) :: lat
  integer(c_int) :: rc, nov
  call rsd_gpu_init()
  call rsd_lattice_from_module(LATTICE, lat)
  rc = rsd_per_grpery_inv_fri(int(NN, c_int), p_pos, RADII, lat, EWS_ERR, int(correction_method, c_int), mu_f, &
                              APP, APQ, AQQ, nov)
  call rsd_check(rc, 'PER_GRPERY_INV_FRI')
  ! CALC_LATTICE_INVERSE's outputs (csd_post.f90:166-181)
  SIG = lat%sig
  Lammda = lat%lammda
  LOGERR = lat%logerr
  NR = lat%nr
  NI = lat%ni
  phi = lat%phi
END SUBROUTINE PER_GRPERY_INV_FRI


SUBROUTINE MATREV(A, NN)
  use rsd_capi
  implicit none
  INTEGER, intent(in) :: NN
  REAL*8, intent(inout) :: A(NN, NN)
  integer(c_int) :: rc
  call rsd_gpu_init()
  rc = rsd_matrev(int(NN, c_int), A)
  call rsd_check(rc, 'MATREV')
  if (rc > 0) then
    write(*, *) 'k=', rc
    write(*, *) 'The leading minor of order k is not'
    write(*, *) 'positive definite, and the factorization could not be'
    write(*, *) 'completed.'
  end if
END SUBROUTINE MATREV


subroutine MobToResist_FTS(APP, APQ, AQQ, NN, RPP, RPQ, RQQ)
  use rsd_capi
  implicit none
  INTEGER, intent(in) :: NN
  REAL*8, intent(in) :: APP(6*NN, 6*NN), APQ(6*NN, 5*NN), AQQ(5*NN, 5*NN)
  real(8), intent(out) :: RPP(6*NN, 6*NN), RPQ(6*NN, 5*NN), RQQ(5*NN, 5*NN)
  integer(c_int) :: rc
  call rsd_gpu_init()
  rc = rsd_mob_to_resist_fts(int(NN, c_int), APP, APQ, AQQ, RPP, RPQ, RQQ)
  call rsd_check(rc, 'MobToResist_FTS')
  if (rc > 0) write(*, *) 'k=', rc, ' (MobToResist_FTS: leading minor not positive definite)'
end subroutine MobToResist_FTS


subroutine arrangement(RPP, RPQ, RQQ, mu, NN)
  use rsd_capi
  implicit none
  INTEGER, intent(in) :: NN
  REAL*8, intent(in) :: RPP(6*NN, 6*NN), RPQ(6*NN, 5*NN), RQQ(5*NN, 5*NN)
  real(8), intent(out) :: mu(5*NN, 5*NN)
  integer(c_int) :: rc
  call rsd_gpu_init()
  rc = rsd_arrangement(int(NN, c_int), RPP, RPQ, RQQ, mu)
  call rsd_check(rc, 'arrangement')
end subroutine arrangement
