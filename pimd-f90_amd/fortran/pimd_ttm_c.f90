! ISO_C_BINDING interfaces to the C ABI of include/pimd_ttm.h.
!
! Arrays are passed as contiguous assumed-size dummies, so the reference's own
! column-major arrays (RRt(3,Natoms,Nbeads), virt(3,3), ...) go across the
! boundary without copies or transposes.
module pimd_ttm_c
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: PIMD_OK = 0, PIMD_EINVAL = 1, PIMD_EDEVICE = 2, &
                                       PIMD_ENOMEM = 3, PIMD_ESTATE = 4
  integer(c_int), parameter, public :: PIMD_TTM21F = 2, PIMD_TTM3F = 3, PIMD_QSPCFW = 4, &
                                       PIMD_SPCF = 5, PIMD_FP64 = 0, PIMD_MIXED = 1, &
                                       PIMD_NOBS = 9

  type, bind(C), public :: pimd_params
    integer(c_int) :: pot_model = 3
    integer(c_int) :: nwaters = 0
    integer(c_int) :: nbeads = 1
    integer(c_int) :: polar_maxiter = 15
    integer(c_int) :: guess_initdip = 0
    integer(c_int) :: precision = 0
    integer(c_int) :: device = 0
    integer(c_int) :: contraction = 0
    real(c_double) :: box(3) = 0.d0
    real(c_double) :: rc = -1.d0
    real(c_double) :: rc1 = -1.d0
    real(c_double) :: eps_ewald = 1.d-6
    real(c_double) :: polar_sor = real(0.7_c_float, c_double)   ! single-precision literal, as the reference
    real(c_double) :: polar_eps = 1.d-5
    real(c_double) :: mass_o = real(15.994_c_float, c_double)
    real(c_double) :: mass_h = real(1.008_c_float, c_double)
  end type pimd_params

  type, bind(C), public :: pimd_md_params
    real(c_double) :: delt_fs = 0.25d0
    real(c_double) :: temp = 300.d0
    integer(c_int) :: bead_thermostat = 1
    integer(c_int) :: centroid_thermostat = 0
    real(c_double) :: tau_centroid = real(0.1_c_float, c_double)
    integer(c_int) :: global_nh = 0
    integer(c_int) :: intra_timesteps = 0
    real(c_double) :: tau = real(0.1_c_float, c_double)
    real(c_double) :: setNMfreq = 0.d0
  end type pimd_md_params

  public :: pimd_ctx_create, pimd_ctx_destroy, pimd_last_error, pimd_ctx_info, &
            pimd_forces_all_beads, pimd_energy_terms, pimd_monomer_dipoles, pimd_nm_init, pimd_nm_get, &
            pimd_nm_evolve, pimd_langevin_init, pimd_langevin, pimd_nose_hoover, &
            pimd_md_init, pimd_state_set, pimd_state_get, pimd_md_run, pimd_sync, &
            pimd_md_observables, pimd_error_message

  interface
    integer(c_int) function pimd_ctx_create(p, ctx) bind(C, name="pimd_ctx_create")
      import :: c_int, c_ptr, pimd_params
      type(pimd_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function pimd_ctx_destroy(ctx) bind(C, name="pimd_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pimd_last_error(buf, n) bind(C, name="pimd_last_error")
      import :: c_int, c_char, c_size_t
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_size_t), value :: n
    end function
    integer(c_int) function pimd_ctx_info(ctx, out) bind(C, name="pimd_ctx_info")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function pimd_forces_all_beads(ctx, RRt, RRc, t, Upott, dRRt, virt, &
        virialct, dipI, dipE, chg, iters) bind(C, name="pimd_forces_all_beads")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: RRt(*), RRc(*)
      integer(c_int), value :: t
      real(c_double), intent(out) :: Upott(*), dRRt(*), virt(*), virialct(*), dipI(*), dipE(*), chg(*)
      integer(c_int), intent(out) :: iters(*)
    end function
    integer(c_int) function pimd_energy_terms(ctx, terms) bind(C, name="pimd_energy_terms")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: terms(*)
    end function
    integer(c_int) function pimd_monomer_dipoles(ctx, RRt, dip) bind(C, name="pimd_monomer_dipoles")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: RRt(*)
      real(c_double), intent(out) :: dip(*)
    end function
    integer(c_int) function pimd_nm_init(ctx, omegan, delta, setnm) bind(C, name="pimd_nm_init")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: omegan, delta, setnm
    end function
    integer(c_int) function pimd_nm_get(ctx, C, A11, A12, A21, A22, om, msf) bind(C, name="pimd_nm_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: C(*), A11(*), A12(*), A21(*), A22(*), om(*), msf(*)
    end function
    integer(c_int) function pimd_nm_evolve(ctx, RRt, PPt, masses) bind(C, name="pimd_nm_evolve")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: RRt(*), PPt(*)
      real(c_double), intent(in) :: masses(*)
    end function
    integer(c_int) function pimd_langevin_init(ctx, delt2, centroid, tau_bead, temp) &
        bind(C, name="pimd_langevin_init")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: delt2, tau_bead, temp
      integer(c_int), value :: centroid
    end function
    ! noise: c_loc of [Natoms][P][3] normals, or c_null_ptr for device draws
    integer(c_int) function pimd_langevin(ctx, PPt, noise, seed, counter) bind(C, name="pimd_langevin")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: PPt(*)
      type(c_ptr), value :: noise
      integer(c_int64_t), value :: seed, counter
    end function
    integer(c_int) function pimd_nose_hoover(s, uk, vxi, tau, delt2, L, temp) bind(C, name="pimd_nose_hoover")
      import :: c_int, c_double
      real(c_double), intent(out) :: s
      real(c_double), intent(inout) :: uk, vxi(2)
      real(c_double), value :: tau, delt2, temp
      integer(c_int), value :: L
    end function
    integer(c_int) function pimd_md_init(ctx, md) bind(C, name="pimd_md_init")
      import :: c_int, c_ptr, pimd_md_params
      type(c_ptr), value :: ctx
      type(pimd_md_params), intent(in) :: md
    end function
    ! dRRt: c_loc of the gradient, or c_null_ptr to start from zero
    integer(c_int) function pimd_state_set(ctx, RRt, PPt, dRRt) bind(C, name="pimd_state_set")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: RRt(*), PPt(*)
      type(c_ptr), value :: dRRt
    end function
    integer(c_int) function pimd_state_get(ctx, RRt, PPt, dRRt, RRc) bind(C, name="pimd_state_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: RRt(*), PPt(*), dRRt(*), RRc(*)
    end function
    integer(c_int) function pimd_md_run(ctx, t0, nsteps, noise, seed, Upot, virial, virialc, uk) &
        bind(C, name="pimd_md_run")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: t0, nsteps
      type(c_ptr), value :: noise
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: Upot(*), virial(*), virialc(*), uk(*)
    end function
    ! obs(PIMD_NOBS, nsteps): sys_temp, virial energy and pressure, simple
    ! energy and pressure, total dipole (3), sum of dipole magnitudes
    integer(c_int) function pimd_md_observables(ctx, nsteps, obs) bind(C, name="pimd_md_observables")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: obs(*)
    end function
    integer(c_int) function pimd_sync(ctx) bind(C, name="pimd_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface

contains

  ! Message of the last failed call, as a Fortran string.
  function pimd_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char) :: buf(512)
    integer :: i, rc
    rc = pimd_last_error(buf, int(512, c_size_t))
    msg = ' '
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    enddo
  end function pimd_error_message

end module pimd_ttm_c
