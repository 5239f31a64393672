! Drop-in replacement for the reference's force_calc module (force_calc.f90).
!
! Same module name and the same public procedures, so the reference's MD.f90
! compiles against it unchanged; build it in place of force_calc.f90.  The
! MPI master/worker loop over beads (force_calc.f90:31-92) becomes one call
! that evaluates every bead on the GPU.
module force_calc
  use dans_timer   ! re-exported, as the reference module does
  use pimd_gpu
  implicit none
  double precision, allocatable, private, save :: chg_all(:,:), virt_all(:,:,:), terms(:,:)

contains

  ! force_calc.f90:26-106.  Reads and writes the main_stuff globals exactly as
  ! the reference's serial run does: chg, virt and the energy terms
  ! Umon..Uind (system_mod.f90:30) keep the last bead's values.
  subroutine full_bead_forces
    use main_stuff
    implicit none
    if (.not. gpu_ready() .or. (gpu_contraction() .neqv. CONTRACTION)) then
      call gpu_init_pot(pot_model, Natoms, Nbeads, box, Rc, rc1, eps_ewald, polar_maxiter, &
                        polar_sor, polar_eps, guess_initdip, massO, massH, contraction=CONTRACTION)
    endif
    if (.not. allocated(chg_all)) allocate(chg_all(Natoms, Nbeads), virt_all(3, 3, Nbeads), terms(5, Nbeads))
    if (BAROSTAT) error stop 'force_calc (GPU): the barostat path is not supported'
    call gpu_full_bead_forces(RRt, RRc, t, Upott, dRRt, virt_all, virialt, virialct, &
                              dip_momIt, dip_momEt, chg_all)
    virt = virt_all(:, :, Nbeads)
    chg = chg_all(:, Nbeads)
    call gpu_energy_terms(terms)
    Umon = terms(1, Nbeads)
    Uvdw = terms(2, Nbeads)
    Uelec = terms(3, Nbeads)
    Uind = terms(4, Nbeads)
    Upot = sum(Upott)
    virial = sum(virialt)
    virialc = sum(virialct) / Nbeads
  end subroutine full_bead_forces

  ! force_calc.f90:112-170 for the TTM models (pot_model 2 and 3) and the
  ! fixed-charge models of pot_spc (4 and 5; virialc and Edip_mom, which
  ! pot_spc leaves unset, come back as zeros).
  subroutine potential(RR, RRc, Upot, dRR, virt, virialc, dip_momI, Edip_mom, chg, t, BAROSTAT, sys_lab)
    use system_mod
    implicit none
    double precision, dimension(3, Natoms), intent(in) :: RR, RRc
    double precision, intent(out) :: Upot, virialc
    double precision, dimension(3, Natoms), intent(out) :: dRR
    double precision, dimension(3, 3), intent(out) :: virt
    double precision, dimension(Natoms), intent(out) :: chg
    double precision, dimension(3, Nwaters), intent(out) :: dip_momI, Edip_mom
    integer, intent(in) :: t
    logical, intent(in) :: BAROSTAT
    character(len=*), intent(in) :: sys_lab
    double precision :: e5(5)
    if (pot_model < 2 .or. pot_model > 5) &
      error stop 'force_calc (GPU): only TTM2.1-F, TTM3-F, q-SPC/Fw and SPC/F'
    if (BAROSTAT) error stop 'force_calc (GPU): the barostat path is not supported'
    ! the context follows the run's CONTRACTION flag: contracted_MD calls
    ! potential at the centroid with the monomer PES left out (MD.f90:398)
    if (.not. gpu_ready() .or. (gpu_contraction() .neqv. CONTRACTION)) then
      call gpu_init_pot(pot_model, Natoms, Nbeads, box, Rc, rc1, eps_ewald, polar_maxiter, &
                        polar_sor, polar_eps, guess_initdip, massO, massH, contraction=CONTRACTION)
    endif
    call gpu_potential(RR, RRc, Upot, dRR, virt, virialc, dip_momI, Edip_mom, chg, t, e5)
    Umon = e5(1)
    Uvdw = e5(2)
    Uelec = e5(3)
    Uind = e5(4)
  end subroutine potential

  ! The SIESTA monomer helper of the reference module (force_calc.f90:176-198)
  ! is not on the TTM path.
  subroutine siesta_monomer(r1, dr1, e1)
    double precision, dimension(3, 3) :: r1, dr1
    double precision :: e1
    error stop 'force_calc (GPU): SIESTA monomer calculations are not supported'
  end subroutine siesta_monomer

  ! force_calc.f90:205-264, every bead in one launch (contracted_MD and the
  ! SIESTA branch call it after each step, MD.f90:229,407).
  subroutine calc_monomer_dip_moments(dip_momIt, RRt)
    use main_stuff, only: pot_model, Natoms, Nbeads, box, Rc, rc1, eps_ewald, polar_maxiter, &
                          polar_sor, polar_eps, guess_initdip, massO, massH, CONTRACTION
    double precision, dimension(:, :, :), intent(out) :: dip_momIt
    double precision, dimension(:, :, :), intent(in) :: RRt
    if (.not. gpu_ready() .or. (gpu_contraction() .neqv. CONTRACTION)) then
      call gpu_init_pot(pot_model, Natoms, Nbeads, box, Rc, rc1, eps_ewald, polar_maxiter, &
                        polar_sor, polar_eps, guess_initdip, massO, massH, contraction=CONTRACTION)
    endif
    call gpu_monomer_dipoles(RRt, dip_momIt)
  end subroutine calc_monomer_dip_moments

end module force_calc
