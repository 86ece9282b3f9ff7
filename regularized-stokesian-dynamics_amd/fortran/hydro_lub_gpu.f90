! hydro_lub_gpu.f90 -- drop-in hydro_lub_mod for the MI355X library.
!
! Replaces the reference's hydro_lub.f90 (module hydro_lub_mod, whose only
! public procedure is lubmxcalc, hydro_lub.f90:22-308).  Same module name,
! same public interface, so stepper.f90 (`use hydro_lub_mod`,
! stepper.f90:13,109) links against it unchanged.  The routine reads the
! same module state the reference reads -- CONFIG::RADII, size::NB,
! method::IsPeriod, SYS_property::mu_f and LATTICE_BASE::LR/LI/LV -- and
! the tables come from the working directory's lubdat/ on first use.
module hydro_lub_mod
  use iso_c_binding
  use rsd_capi
  implicit none
  private

  public :: lubmxcalc

contains

  subroutine lubmxcalc(ntotal, P_pos, rfu, rfe, rse)
    use size, only: NB
    use SYS_property, only: mu_f
    use method, only: IsPeriod
    use CONFIG, only: RADII
    use LATTICE_BASE, only: LR
    integer, intent(in) :: ntotal
    real(8), dimension(3, ntotal), intent(in) :: P_pos
    real(8), dimension(6*ntotal, 6*ntotal), intent(out) :: rfu
    real(8), dimension(6*ntotal, 5*ntotal), intent(out) :: rfe
    real(8), dimension(5*ntotal, 5*ntotal), intent(out) :: rse
    type(rsd_lattice) :: lat
    integer(c_int) :: rc, ncontact, iper
    call rsd_gpu_init()
    call rsd_gpu_tables()
    call rsd_lattice_from_module(LR, lat)
    iper = 0
    if (IsPeriod) iper = 1
    rc = rsd_lubmxcalc(int(ntotal, c_int), int(NB, c_int), iper, P_pos, RADII, lat, mu_f, rfu, rfe, rse, ncontact)
    call rsd_check(rc, 'lubmxcalc')
  end subroutine lubmxcalc

end module hydro_lub_mod
