! Link-time replacement for the reference's external pot_ttm
! (pot_ttm.f90:3-15): same symbol, same implicit interface, one bead per call
! on the GPU.  Build it in place of pot_ttm.f90 to keep the reference's own
! force_calc bead loop; build force_calc_gpu.f90 instead to evaluate every
! bead in one launch.
!
! One predictor history serves every call, as the reference's tx_dip
! (pot_mod.f90:53) does in a serial run.  Side effects on system_mod match
! the reference: Umon, Uvdw, Uelec, Uind (system_mod.f90:30).  The
! CONTRACTION branch (pot_ttm.f90:92-132) follows system_mod's flag.
subroutine pot_ttm(RR, RRc, En, dRR, virt, virialc, dip_momI, Edip_mom, chg, t)
  use system_mod
  use pimd_gpu
  implicit none
  double precision, dimension(3, Natoms), intent(in) :: RR, RRc
  double precision, intent(out) :: En, virialc
  double precision, dimension(3, Natoms), intent(out) :: dRR
  double precision, dimension(3, 3), intent(out) :: virt
  double precision, dimension(3, Nwaters), intent(out) :: dip_momI, Edip_mom
  double precision, dimension(Natoms), intent(out) :: chg
  integer, intent(in) :: t
  double precision :: e5(5)
  ! the reference's pot_ttm reads CONTRACTION (system_mod.f90:40) on every call
  ! and then leaves out the monomer PES terms (pot_ttm.f90:92-132)
  if (.not. gpu_ready() .or. (gpu_contraction() .neqv. CONTRACTION)) then
    call gpu_init_pot(pot_model, Natoms, Nbeads, box, Rc, rc1, eps_ewald, polar_maxiter, &
                      polar_sor, polar_eps, guess_initdip, massO, massH, contraction=CONTRACTION)
  endif
  call gpu_potential(RR, RRc, En, dRR, virt, virialc, dip_momI, Edip_mom, chg, t, e5)
  Umon = e5(1)
  Uvdw = e5(2)
  Uelec = e5(3)
  Uind = e5(4)
end subroutine pot_ttm

! pot_ttm.f90 also holds get_dd3 and calc_dpcf (pot_ttm.f90:633-675); outside
! pot_ttm only dip_ttm (the SIESTA model, pot_model 6) calls them, which this
! build does not support.
subroutine get_dd3(R, ts1, ts2, dd3)
  implicit none
  double precision :: R(3), ts1, ts2, dd3(3, 3)
  error stop 'pot_ttm (GPU): get_dd3 is only used by the SIESTA model, which is not supported'
end subroutine get_dd3

subroutine calc_dpcf(md_step, dpcf)
  implicit none
  integer :: md_step
  double precision :: dpcf(4)
  error stop 'pot_ttm (GPU): calc_dpcf is only used by the SIESTA model, which is not supported'
end subroutine calc_dpcf
