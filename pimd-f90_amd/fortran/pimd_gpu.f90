! Fortran host side of the GPU force path and ring-polymer propagation.
!
! Mirrors the reference's interfaces on this path so a Fortran PIMD driver can
! switch to the GPU by calling these instead of (or through the drop-in
! force_calc module built from force_calc_gpu.f90):
!
!   gpu_init_pot          init_pot + ewald_set   (potential.f90:43-199)
!   gpu_full_bead_forces  full_bead_forces       (force_calc.f90:26-106), all beads at once
!   gpu_potential         potential / pot_ttm    (force_calc.f90:112-170, pot_ttm.f90:3)
!   gpu_init_normal_modes InitNormalModes        (NormalModes.f90:68-218)
!   gpu_evolve_rings      EvolveRing for every atom (NormalModes.f90:34-63, MD.f90:52-58)
!   gpu_init_langevin_nm  Init_Langevin_NM       (Langevin.f90:68-98)
!   gpu_langevin_nm       Langevin_NM            (Langevin.f90:30-64)
!
! Errors follow the reference's convention: the run stops with a message and a non-zero exit status (error stop).
module pimd_gpu
  use iso_c_binding
  use pimd_ttm_c
  implicit none
  private
  public :: gpu_init_pot, gpu_full_bead_forces, gpu_potential, gpu_energy_terms, gpu_monomer_dipoles, &
            gpu_init_normal_modes, gpu_evolve_rings, gpu_init_langevin_nm, gpu_langevin_nm, &
            gpu_finalize, gpu_ready, gpu_nm_data, gpu_contraction

  type(c_ptr), save :: ctx_all = c_null_ptr   ! every bead in one launch
  type(c_ptr), save :: ctx_one = c_null_ptr   ! one bead (potential)
  integer, save :: g_natoms = 0, g_nwaters = 0, g_nbeads = 0
  type(pimd_params), save :: g_params

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= PIMD_OK) then
      write(*, '(a)') 'pimd_gpu: ' // what // ' failed: ' // trim(pimd_error_message())
      error stop 'pimd_gpu: GPU force path error'
    endif
  end subroutine check

  logical function gpu_ready()
    gpu_ready = c_associated(ctx_all)
  end function gpu_ready

  ! Whether the contexts were made with pot_ttm's CONTRACTION branch.
  logical function gpu_contraction()
    gpu_contraction = g_params%contraction /= 0
  end function gpu_contraction

  ! precision (optional): PIMD_FP64 (default) or PIMD_MIXED.
  ! contraction (optional): the reference's CONTRACTION flag (system_mod.f90:40);
  ! when true pot_ttm leaves out the monomer PES energy, gradient and virial
  ! (pot_ttm.f90:92-132), as contracted_MD's centroid call needs (MD.f90:398).
  subroutine gpu_init_pot(pot_model, Natoms, Nbeads, box, Rc, rc1, eps_ewald, polar_maxiter, &
                          polar_sor, polar_eps, guess_initdip, massO, massH, precision, &
                          contraction)
    integer, intent(in) :: pot_model, Natoms, Nbeads, polar_maxiter
    double precision, intent(in) :: box(3), Rc, rc1, eps_ewald, polar_sor, polar_eps, massO, massH
    logical, intent(in) :: guess_initdip
    integer, intent(in), optional :: precision
    logical, intent(in), optional :: contraction
    call gpu_finalize()
    g_params%pot_model = pot_model
    g_params%nwaters = Natoms / 3
    g_params%nbeads = Nbeads
    g_params%polar_maxiter = polar_maxiter
    g_params%guess_initdip = merge(1, 0, guess_initdip)
    g_params%precision = PIMD_FP64
    if (present(precision)) g_params%precision = precision
    g_params%device = 0
    g_params%contraction = 0
    if (present(contraction)) g_params%contraction = merge(1, 0, contraction)
    g_params%box = box
    g_params%rc = Rc
    g_params%rc1 = rc1
    g_params%eps_ewald = eps_ewald
    g_params%polar_sor = polar_sor
    g_params%polar_eps = polar_eps
    g_params%mass_o = massO
    g_params%mass_h = massH
    call check(pimd_ctx_create(g_params, ctx_all), 'pimd_ctx_create')
    g_natoms = Natoms
    g_nwaters = Natoms / 3
    g_nbeads = Nbeads
  end subroutine gpu_init_pot

  subroutine gpu_finalize()
    integer(c_int) :: rc
    if (c_associated(ctx_all)) rc = pimd_ctx_destroy(ctx_all)
    if (c_associated(ctx_one)) rc = pimd_ctx_destroy(ctx_one)
    ctx_all = c_null_ptr
    ctx_one = c_null_ptr
  end subroutine gpu_finalize

  ! All beads: Upott(P), dRRt(3,Natoms,P), virt(3,3,P), virialt(P) (trace),
  ! virialct(P), dip_momIt/EmomEt(3,Nwaters,P), chg(Natoms,P).
  subroutine gpu_full_bead_forces(RRt, RRc, t, Upott, dRRt, virt, virialt, virialct, &
                                  dip_momIt, dip_momEt, chg, iters)
    double precision, intent(in) :: RRt(3, g_natoms, g_nbeads), RRc(3, g_natoms)
    integer, intent(in) :: t
    double precision, intent(out) :: Upott(g_nbeads), dRRt(3, g_natoms, g_nbeads)
    double precision, intent(out) :: virt(3, 3, g_nbeads), virialt(g_nbeads), virialct(g_nbeads)
    double precision, intent(out) :: dip_momIt(3, g_nwaters, g_nbeads), dip_momEt(3, g_nwaters, g_nbeads)
    double precision, intent(out) :: chg(g_natoms, g_nbeads)
    integer, intent(out), optional :: iters(g_nbeads)
    integer(c_int) :: it(g_nbeads)
    integer :: k
    if (.not. c_associated(ctx_all)) error stop 'pimd_gpu: gpu_init_pot has not been called'
    call check(pimd_forces_all_beads(ctx_all, RRt, RRc, int(t, c_int), Upott, dRRt, virt, &
               virialct, dip_momIt, dip_momEt, chg, it), 'pimd_forces_all_beads')
    do k = 1, g_nbeads
      virialt(k) = virt(1, 1, k) + virt(2, 2, k) + virt(3, 3, k)
    enddo
    if (present(iters)) iters = it
  end subroutine gpu_full_bead_forces

  ! Energy terms Umon, Uvdw, Uelec, Uind, Uvdw_lrc per bead (system_mod.f90:30).
  subroutine gpu_energy_terms(terms)
    double precision, intent(out) :: terms(5, g_nbeads)
    call check(pimd_energy_terms(ctx_all, terms), 'pimd_energy_terms')
  end subroutine gpu_energy_terms

  ! Monomer dipole moments of every bead (force_calc.f90:205-264).
  subroutine gpu_monomer_dipoles(RRt, dip_momIt)
    double precision, intent(in) :: RRt(3, g_natoms, g_nbeads)
    double precision, intent(out) :: dip_momIt(3, g_nwaters, g_nbeads)
    if (.not. c_associated(ctx_all)) error stop 'pimd_gpu: gpu_init_pot has not been called'
    call check(pimd_monomer_dipoles(ctx_all, RRt, dip_momIt), 'pimd_monomer_dipoles')
  end subroutine gpu_monomer_dipoles

  ! One bead with the argument list of pot_ttm (pot_ttm.f90:3-15).
  ! terms (optional): Umon, Uvdw, Uelec, Uind, Uvdw_lrc of this call.
  subroutine gpu_potential(RR, RRc, En, dRR, virt, virialc, dip_momI, Edip_mom, chg, t, terms)
    double precision, intent(in) :: RR(3, g_natoms), RRc(3, g_natoms)
    double precision, intent(out) :: En, virialc, dRR(3, g_natoms), virt(3, 3)
    double precision, intent(out) :: dip_momI(3, g_nwaters), Edip_mom(3, g_nwaters), chg(g_natoms)
    integer, intent(in) :: t
    double precision, intent(out), optional :: terms(5)
    type(pimd_params) :: p1
    double precision :: e(1), vc(1)
    integer(c_int) :: it(1)
    if (.not. c_associated(ctx_all)) error stop 'pimd_gpu: gpu_init_pot has not been called'
    if (.not. c_associated(ctx_one)) then
      p1 = g_params
      p1%nbeads = 1
      call check(pimd_ctx_create(p1, ctx_one), 'pimd_ctx_create (one bead)')
    endif
    call check(pimd_forces_all_beads(ctx_one, RR, RRc, int(t, c_int), e, dRR, virt, vc, &
               dip_momI, Edip_mom, chg, it), 'pimd_forces_all_beads (one bead)')
    En = e(1)
    virialc = vc(1)
    if (present(terms)) call check(pimd_energy_terms(ctx_one, terms), 'pimd_energy_terms (one bead)')
  end subroutine gpu_potential

  subroutine gpu_init_normal_modes(omegan, delta, setNMfreq)
    double precision, intent(in) :: omegan, delta, setNMfreq
    call check(pimd_nm_init(ctx_all, omegan, delta, setNMfreq), 'pimd_nm_init')
  end subroutine gpu_init_normal_modes

  ! The NormalModes module data C, A11..A22, omegalist, MassScaleFactor.
  subroutine gpu_nm_data(C, A11, A12, A21, A22, omegalist, MassScaleFactor)
    double precision, intent(out) :: C(g_nbeads, g_nbeads), A11(g_nbeads), A12(g_nbeads), &
                                     A21(g_nbeads), A22(g_nbeads), omegalist(g_nbeads), &
                                     MassScaleFactor(g_nbeads)
    double precision :: cm(g_nbeads, g_nbeads)
    call check(pimd_nm_get(ctx_all, cm, A11, A12, A21, A22, omegalist, MassScaleFactor), &
               'pimd_nm_get')
    C = transpose(cm)   ! the C ABI stores C[mode][bead]; the reference's C(mode, bead)
  end subroutine gpu_nm_data

  ! EvolveRing(RRt(:,i,:), PPt(:,i,:), Nbeads, mass(i)) for every atom i.
  subroutine gpu_evolve_rings(RRt, PPt, masses)
    double precision, intent(inout) :: RRt(3, g_natoms, g_nbeads), PPt(3, g_natoms, g_nbeads)
    double precision, intent(in) :: masses(g_natoms)
    call check(pimd_nm_evolve(ctx_all, RRt, PPt, masses), 'pimd_nm_evolve')
  end subroutine gpu_evolve_rings

  subroutine gpu_init_langevin_nm(delt2, CENTROIDTHERMOSTAT, tau_bead, temp)
    double precision, intent(in) :: delt2, tau_bead, temp
    logical, intent(in) :: CENTROIDTHERMOSTAT
    call check(pimd_langevin_init(ctx_all, delt2, merge(1_c_int, 0_c_int, CENTROIDTHERMOSTAT), &
               tau_bead, temp), 'pimd_langevin_init')
  end subroutine gpu_init_langevin_nm

  ! Langevin_NM(PPt) with the standard normals noise(3, Nbeads, Natoms) in the
  ! order the reference draws them (Langevin.f90:41-55); drawing them on the
  ! host with the reference's rand_norm keeps its random stream.
  subroutine gpu_langevin_nm(PPt, noise)
    double precision, intent(inout) :: PPt(3, g_natoms, g_nbeads)
    double precision, intent(in), target :: noise(3, g_nbeads, g_natoms)
    call check(pimd_langevin(ctx_all, PPt, c_loc(noise), 0_c_int64_t, 0_c_int64_t), &
               'pimd_langevin')
  end subroutine gpu_langevin_nm

end module pimd_gpu
