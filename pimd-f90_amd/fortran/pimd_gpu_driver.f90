! Fortran program over the GPU force path, reading the same input records as
! the reference-side test driver and writing the same output records, so one
! comparison covers the Fortran binding, the C ABI and the kernels.
!
! Input  (stream): int32 ih(16), real*8 dh(32), then arrays.
!   ih: mode, pot_model, nwaters, nbeads, polar_maxiter, guess_initdip, t
!   dh: box(3), Rc, rc1, eps_ewald, polar_sor, polar_eps, massO, massH,
!       delt_fs, temp, tau_centroid, setNMfreq
!   mode 1  forces   RRt(3,Natoms,P), RRc(3,Natoms) -> per bead:
!           En, Umon, Uvdw, Uelec, Uind, virialc, virt(3,3), dRR, dipI, dipE, chg
!   mode 2  nm       RRt, PPt -> C, A11, A12, A21, A22, omegalist, MassScaleFactor, RRt, PPt
!   mode 3  langevin PPt, noise(3,P,Natoms) -> PPt
program pimd_gpu_driver
  use pimd_gpu
  implicit none
  integer :: ih(16), mode, Nwaters, Natoms, P, ib, ia
  double precision :: dh(32), delt, temp, omegan
  character(len=512) :: fin, fout
  double precision, allocatable :: RRt(:,:,:), RRc(:,:), PPt(:,:,:), dRRt(:,:,:), virt(:,:,:)
  double precision, allocatable :: Upott(:), virialt(:), virialct(:), dipI(:,:,:), dipE(:,:,:)
  double precision, allocatable :: chg(:,:), terms(:,:), masses(:), noise(:,:,:)
  double precision, allocatable :: C(:,:), A11(:), A12(:), A21(:), A22(:), om(:), msf(:)
  double precision, parameter :: KB_amuA2ps2perK = 0.831447148d0, hbar = 6.35077993041d0

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  read(10) ih
  read(10) dh
  mode = ih(1)
  Nwaters = ih(3)
  Natoms = 3 * Nwaters
  P = ih(4)
  call gpu_init_pot(ih(2), Natoms, P, dh(1:3), dh(4), dh(5), dh(6), ih(5), dh(7), dh(8), &
                    ih(6) /= 0, dh(9), dh(10))
  delt = dh(11) / 1000.d0
  temp = dh(12)

  select case (mode)
  case (1)
    allocate(RRt(3, Natoms, P), RRc(3, Natoms), dRRt(3, Natoms, P), virt(3, 3, P))
    allocate(Upott(P), virialt(P), virialct(P), dipI(3, Nwaters, P), dipE(3, Nwaters, P))
    allocate(chg(Natoms, P), terms(5, P))
    read(10) RRt
    read(10) RRc
    call gpu_full_bead_forces(RRt, RRc, ih(7), Upott, dRRt, virt, virialt, virialct, dipI, dipE, chg)
    call gpu_energy_terms(terms)
    do ib = 1, P
      write(11) Upott(ib), terms(1:4, ib), virialct(ib), virt(:, :, ib), dRRt(:, :, ib), &
                dipI(:, :, ib), dipE(:, :, ib), chg(:, ib)
    enddo
  case (2, 3)
    allocate(RRt(3, Natoms, P), PPt(3, Natoms, P), masses(Natoms))
    allocate(C(P, P), A11(P), A12(P), A21(P), A22(P), om(P), msf(P))
    do ia = 1, Natoms
      masses(ia) = dh(10)
      if (mod(ia + 2, 3) == 0) masses(ia) = dh(9)
    enddo
    omegan = KB_amuA2ps2perK * temp * P / hbar
    call gpu_init_normal_modes(omegan, delt, dh(14))
    if (mode == 2) then
      read(10) RRt
      read(10) PPt
      call gpu_nm_data(C, A11, A12, A21, A22, om, msf)
      write(11) C, A11, A12, A21, A22, om, msf
      call gpu_evolve_rings(RRt, PPt, masses)
      write(11) RRt, PPt
    else
      allocate(noise(3, P, Natoms))
      read(10) PPt
      read(10) noise
      call gpu_init_langevin_nm(delt / 2.d0, ih(9) /= 0, dh(13), P * temp)
      call gpu_langevin_nm(PPt, noise)
      write(11) PPt
    endif
  case default
    error stop 'pimd_gpu_driver: unknown mode'
  end select
  close(11)
  call gpu_finalize()
end program pimd_gpu_driver
