! Test-only driver around the reference PIMD-F90 routines (compiled from the
! reference sources by oracle/Makefile into oracle/_ref/).  It is part of the
! oracle: it is never linked into, loaded by, or shipped with the product.
!
! Input  (stream, little endian): int32 ih(16), real*8 dh(32), then arrays.
! Output (stream): real*8 values, layout per mode (see oracle/refrun.py).
!
!   mode 1  forces     pot_ttm per bead            (pot_ttm.f90:3-630)
!                      (pot_spc for pot_model 4, 5: pot_spc.f90:1-277)
!   mode 2  nm         InitNormalModes+EvolveRing  (NormalModes.f90:34-218)
!   mode 3  langevin   Init_Langevin_NM+Langevin_NM(Langevin.f90:30-98)
!   mode 4  nosehoover Nose_Hoover                 (Nose_Hoover.f90:33-112)
!   mode 5  traj       PIMD_VV step order          (MD.f90:35-110)
!   mode 6  funcs      erfc/gammq2_3/PBCsmear*/self* (math.f90, smear.f90)
!   mode 7  setup      init_pot + ewald_set tables (potential.f90:43-199)
!   mode 8  monomer    pot_nasa + dms_nasa         (nasa.f90:2-293)
!   mode 9  contracted contracted_MD step order    (MD.f90:276-449)
!
! Mode 5 also writes, per step, the observables of write_out
! (InputOutput.f90:556-621) from the reference's estimators
! (estimators.f90:30-96): temperature, quantum virial and simple energy and
! pressure, the total molecular dipole and the sum of the molecular dipole
! magnitudes (Debye).
program ref_driver
use consts
use system_mod
use pot_mod
use nasa_mod, only: dms_param1
use NormalModes
use Langevin
use estimators, only: quantum_virial_estimators, simple_quantum_estimators
use math, only: erfc, gammq2_3, rand_norm
use main_stuff, only: RRt, PPt, dRRt, RRc, PPc, delt, delt2, uk, imassO, imassH, &
                      temp, tau, tau_centroid, s, sbead, global_chain_length, &
                      bead_chain_length, vxi_beads, vxi_global, omegan, setNMfreq, &
                      bead_thermostat_type, BEADTHERMOSTAT, CENTROIDTHERMOSTAT, &
                      THERMOSTAT, PBCs, bead_thermostat, calc_uk
implicit none
integer :: ih(16)
double precision :: dh(32)
character(len=512) :: fin, fout
integer :: mode, nb, ib, ia, nf, n, it, nsteps, ist, nrec, ns
double precision :: En, vc, vtmp(3,3), e1, q3(3), dq3(3,3,3), r1(3,3), dr1(3,3)
double precision :: ts0, ts1, ts2, ts3, mass, sNH, ukNH, vxiNH(2)
double precision, allocatable :: dRR(:,:), dipI(:,:), dipE(:,:), chgb(:)
double precision, allocatable :: xs(:), drsq(:), apol(:), hist(:,:,:,:), noise(:)
integer, allocatable :: seedst(:)
integer :: chain2, LNH, per_bead_hist, nseed
logical :: have_hist
double precision, allocatable :: dRRfast(:,:,:), Upostep(:)
double precision :: deltfast, delt2fast, virialmon, virialcmon, roh1(3), roh2(3)
integer :: intra, tintra, iw, iO, iH1, iH2
double precision :: Usum, visum, vcsum, sys_temp, eq, pq, es, ps, dipmag, dipmom(3)
double precision, allocatable :: dipIt(:,:,:)

call get_command_argument(1, fin)
call get_command_argument(2, fout)
open(10, file=trim(fin), access='stream', form='unformatted', status='old')
open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
read(10) ih
read(10) dh

mode = ih(1)
pot_model = ih(2)
Nwaters = ih(3)
Nbeads = ih(4)
polar_maxiter = ih(5)
guess_initdip = (ih(6) /= 0)
per_bead_hist = ih(15)
Natoms = 3*Nwaters
nb = Nbeads
box = dh(1:3)
Rc = dh(4)
rc1 = dh(5)
eps_ewald = dh(6)
polar_sor = dh(7)
polar_eps = dh(8)
massO = dh(9)
massH = dh(10)
Rc2 = Rc*Rc
boxi = 1.d0/box
volume = box(1)*box(2)*box(3)
volume_init = volume
CONTRACTION = .false.
MONOMERPIMD = .false.
print_dipiters = .true.

select case (mode)
case (1)
   call init_pot
   allocate(RRt(3,Natoms,nb), RRc(3,Natoms))
   read(10) RRt
   read(10) RRc
   allocate(dRR(3,Natoms), dipI(3,Nwaters), dipE(3,Nwaters), chgb(Natoms))
   allocate(hist(3, fd:ld, 4, nb))
   hist = 0.d0
   do ib = 1, nb
      if (pot_model >= 4) then
         ! pot_spc (pot_spc.f90:1-277) sets neither virialc nor Edip_mom,
         ! and has no induction energy: written as zeros
         call pot_spc(RRt(:,:,ib), En, dRR, vtmp, dipI, chgb)
         vc = 0.d0; Uind = 0.d0; dipE = 0.d0
         write(11) En, Umon, Uvdw, Uelec, Uind, vc, vtmp, dRR, dipI, dipE, chgb
         cycle
      endif
      if (per_bead_hist /= 0) tx_dip = hist(:,:,:,ib)
      call pot_ttm(RRt(:,:,ib), RRc, En, dRR, vtmp, vc, dipI, dipE, chgb, ih(7))
      if (per_bead_hist /= 0) hist(:,:,:,ib) = tx_dip
      write(11) En, Umon, Uvdw, Uelec, Uind, vc, vtmp, dRR, dipI, dipE, chgb
   enddo

case (2)
   allocate(RRt(3,Natoms,nb), PPt(3,Natoms,nb))
   read(10) RRt
   read(10) PPt
   temp = dh(12); delt = dh(11)/1000.d0; setNMfreq = dh(14)
   omegan = KB_amuA2ps2perK*temp*Nbeads/hbar
   call InitNormalModes(Nbeads, omegan, delt, setNMfreq, 6)
   write(11) C, A11, A12, A21, A22, omegalist, MassScaleFactor
   do ia = 1, Natoms
      mass = massH
      if (mod(ia+2,3) == 0) mass = massO
      call EvolveRing(RRt(:,ia,:), PPt(:,ia,:), Nbeads, mass)
   enddo
   write(11) RRt, PPt

case (3)
   allocate(PPt(3,Natoms,nb))
   read(10) PPt
   temp = dh(12); delt = dh(11)/1000.d0; delt2 = delt/2.d0; setNMfreq = dh(14)
   tau_centroid = dh(13)
   omegan = KB_amuA2ps2perK*temp*Nbeads/hbar
   call InitNormalModes(Nbeads, omegan, delt, setNMfreq, 6)
   call Init_Langevin_NM(delt2, ih(9) /= 0, tau_centroid, Nbeads, Nbeads*temp)
   call seed_rng(ih(10))
   n = 3*Natoms*Nbeads
   allocate(noise(n))
   call record_noise(noise, n)
   call Langevin_NM(PPt)
   write(11) c1, c2, noise, PPt

case (4)
   ! Nose_Hoover(s, uk, chain_length, vxi, tau, delt2, L, temp)
   ukNH = dh(16); vxiNH = dh(17:18); chain2 = 2; LNH = ih(14)
   call Nose_Hoover(sNH, ukNH, chain2, vxiNH, dh(15), dh(11)/2000.d0, LNH, dh(12))
   write(11) sNH, ukNH, vxiNH

case (5)
   call init_pot
   allocate(RRt(3,Natoms,nb), PPt(3,Natoms,nb), dRRt(3,Natoms,nb), RRc(3,Natoms), PPc(3,Natoms))
   read(10) RRt
   read(10) PPt
   read(10) dRRt
   temp = dh(12); delt = dh(11)/1000.d0; delt2 = delt/2.d0; setNMfreq = dh(14)
   tau_centroid = dh(13); tau = dh(15)
   imassO = 1.d0/massO; imassH = 1.d0/massH
   omegan = KB_amuA2ps2perK*temp*Nbeads/hbar
   nsteps = ih(8)
   BEADTHERMOSTAT = (ih(12) /= 0)
   CENTROIDTHERMOSTAT = (ih(9) /= 0)
   THERMOSTAT = (ih(13) /= 0)
   bead_thermostat_type = 'none'
   if (ih(12) == 1) bead_thermostat_type = 'Langevin'
   if (ih(12) == 2) bead_thermostat_type = 'Nose-Hoover'
   global_chain_length = 2; bead_chain_length = 2
   s = 1.d0; sbead = 1.d0
   call InitNormalModes(Nbeads, omegan, delt, setNMfreq, 6)
   if (THERMOSTAT) then
      allocate(vxi_global(2)); vxi_global = 1
   endif
   if (BEADTHERMOSTAT) then
      allocate(vxi_beads(2, Natoms, Nbeads, 3)); vxi_beads = 0
   endif
   if (ih(12) == 1) call Init_Langevin_NM(delt2, CENTROIDTHERMOSTAT, tau_centroid, Nbeads, Nbeads*temp)
   call seed_rng(ih(10))
   n = 3*Natoms*Nbeads
   allocate(noise(n), dRR(3,Natoms), dipI(3,Nwaters), dipE(3,Nwaters), chgb(Natoms))
   allocate(hist(3, fd:ld, 4, nb), dipIt(3, Nwaters, nb))
   hist = 0.d0
   call calc_uk
   do it = 1, nsteps
      ! MD.f90:41-78
      if (BEADTHERMOSTAT) then
         if (ih(12) == 1) call record_noise(noise, n)
         call bead_thermostat
         if (ih(12) == 1) write(11) noise
      endif
      if (THERMOSTAT) then
         call Nose_Hoover(s, uk, global_chain_length, vxi_global, tau, delt2, 3*Natoms*Nbeads, temp*Nbeads)
         PPt = PPt*s
      endif
      PPt = PPt - MASSCON*dRRt*delt2
      if (Nbeads > 1) then
         do ia = 1, Nwaters
            call EvolveRing(RRt(:,3*ia-2,:), PPt(:,3*ia-2,:), Nbeads, massO)
            call EvolveRing(RRt(:,3*ia-1,:), PPt(:,3*ia-1,:), Nbeads, massH)
            call EvolveRing(RRt(:,3*ia-0,:), PPt(:,3*ia-0,:), Nbeads, massH)
         enddo
      else
         do ia = 1, Nwaters
            RRt(:,3*ia-2,1) = RRt(:,3*ia-2,1) + imassO*PPt(:,3*ia-2,1)*delt
            RRt(:,3*ia-1,1) = RRt(:,3*ia-1,1) + imassH*PPt(:,3*ia-1,1)*delt
            RRt(:,3*ia-0,1) = RRt(:,3*ia-0,1) + imassH*PPt(:,3*ia-0,1)*delt
         enddo
      endif
      RRc = sum(RRt,3)/Nbeads
      PPc = sum(PPt,3)/Nbeads
      call PBCs(RRt, RRc)
      ! force_calc.f90:26-106 (serial order, one bead at a time)
      Usum = 0.d0; visum = 0.d0; vcsum = 0.d0
      do ib = 1, nb
         if (pot_model >= 4) then
            ! pot_spc leaves virialc unset (pot_spc.f90:1-13): taken as zero
            call pot_spc(RRt(:,:,ib), En, dRR, vtmp, dipI, chgb)
            vc = 0.d0
         else
            if (per_bead_hist /= 0) tx_dip = hist(:,:,:,ib)
            call pot_ttm(RRt(:,:,ib), RRc, En, dRR, vtmp, vc, dipI, dipE, chgb, it)
            if (per_bead_hist /= 0) hist(:,:,:,ib) = tx_dip
         endif
         dRRt(:,:,ib) = dRR
         dipIt(:,:,ib) = dipI
         write(11) En, vtmp(1,1)+vtmp(2,2)+vtmp(3,3), vc
         ! force_calc.f90:94-97
         Usum = Usum + En
         visum = visum + (vtmp(1,1)+vtmp(2,2)+vtmp(3,3))
         vcsum = vcsum + vc
      enddo
      vcsum = vcsum/Nbeads
      ! MD.f90:85-106
      call calc_uk
      ! write_out (InputOutput.f90:556-621)
      sys_temp = TEMPFACTOR*uk/(Natoms*Nbeads*Nbeads)
      call quantum_virial_estimators(RRt, visum, vcsum, eq, pq, sys_temp, Usum)
      call simple_quantum_estimators(RRt, visum, es, ps, sys_temp, Usum)
      dipIt = dipIt*DEBYE/CHARGECON
      dipmom = 0.d0
      dipmag = 0.d0
      do iw = 1, Nwaters
         do ia = 1, 3
            dipI(ia,iw) = sum(dipIt(ia,iw,:))/Nbeads
         enddo
      enddo
      dipmom(:) = sum(dipI(:,:), dim=2)
      do iw = 1, Nwaters
         dipmag = dipmag + dsqrt(dot_product(dipI(:,iw), dipI(:,iw)))
      enddo
      write(11) sys_temp, eq, pq, es, ps, dipmom, dipmag
      PPt = PPt - MASSCON*dRRt*delt2
      if (BEADTHERMOSTAT) then
         if (ih(12) == 1) call record_noise(noise, n)
         call bead_thermostat
         if (ih(12) == 1) write(11) noise
      endif
      call calc_uk
      if (THERMOSTAT) then
         call Nose_Hoover(s, uk, global_chain_length, vxi_global, tau, delt2, 3*Natoms*Nbeads, temp*Nbeads)
         PPt = PPt*s
      endif
      write(11) uk
   enddo
   write(11) RRt, PPt, dRRt, RRc
   if (THERMOSTAT) write(11) vxi_global
   if (BEADTHERMOSTAT .and. ih(12) == 2) write(11) vxi_beads

case (9)
   ! contracted_MD (MD.f90:276-449): fast intramolecular forces on the beads
   ! in intra inner steps, intermolecular forces of the centroid with
   ! CONTRACTION on (pot_ttm.f90:92-132) every outer step.  dRRfast is kept
   ! from one outer step to the next (the reference's automatic array is set
   ! to zero only at t = 1, MD.f90:293).
   call init_pot
   allocate(RRt(3,Natoms,nb), PPt(3,Natoms,nb), dRRt(3,Natoms,nb), RRc(3,Natoms), PPc(3,Natoms))
   allocate(dRRfast(3,Natoms,nb), Upostep(nb))
   read(10) RRt
   read(10) PPt
   read(10) dRRt
   intra = ih(16)
   temp = dh(12); delt = dh(11)/1000.d0; delt2 = delt/2.d0; setNMfreq = dh(14)
   deltfast = delt/intra; delt2fast = deltfast/2.d0
   tau_centroid = dh(13); tau = dh(15)
   imassO = 1.d0/massO; imassH = 1.d0/massH
   omegan = KB_amuA2ps2perK*temp*Nbeads/hbar
   nsteps = ih(8)
   BEADTHERMOSTAT = (ih(12) /= 0)
   CENTROIDTHERMOSTAT = (ih(9) /= 0)
   THERMOSTAT = .false.
   bead_thermostat_type = 'none'
   if (ih(12) == 1) bead_thermostat_type = 'Langevin'
   s = 1.d0; sbead = 1.d0
   call InitNormalModes(Nbeads, omegan, deltfast, setNMfreq, 6)
   if (ih(12) == 1) call Init_Langevin_NM(delt2, CENTROIDTHERMOSTAT, tau_centroid, Nbeads, Nbeads*temp)
   call seed_rng(ih(10))
   n = 3*Natoms*Nbeads
   allocate(noise(n), dRR(3,Natoms), dipI(3,Nwaters), dipE(3,Nwaters), chgb(Natoms))
   dRRfast = 0.d0
   call calc_uk
   do it = 1, nsteps
      Upostep = 0.d0
      if (BEADTHERMOSTAT) then
         call record_noise(noise, n)
         call bead_thermostat
         write(11) noise
      endif
      PPt = PPt - MASSCON*dRRt*delt2
      RRc = sum(RRt,3)/Nbeads
      PPc = sum(PPt,3)/Nbeads
      call PBCs(RRt, RRc)
      do tintra = 1, intra
         PPt = PPt - MASSCON*dRRfast*delt2fast
         if (nb > 1) then
            do ia = 1, Nwaters
               call EvolveRing(RRt(:,3*ia-2,:), PPt(:,3*ia-2,:), Nbeads, massO)
               call EvolveRing(RRt(:,3*ia-1,:), PPt(:,3*ia-1,:), Nbeads, massH)
               call EvolveRing(RRt(:,3*ia-0,:), PPt(:,3*ia-0,:), Nbeads, massH)
            enddo
         else
            do ia = 1, Nwaters
               RRt(:,3*ia-2,1) = RRt(:,3*ia-2,1) + imassO*PPt(:,3*ia-2,1)*deltfast
               RRt(:,3*ia-1,1) = RRt(:,3*ia-1,1) + imassH*PPt(:,3*ia-1,1)*deltfast
               RRt(:,3*ia-0,1) = RRt(:,3*ia-0,1) + imassH*PPt(:,3*ia-0,1)*deltfast
            enddo
         endif
         virialmon = 0.d0
         virialcmon = 0.d0
         do ib = 1, nb
            do iw = 1, Nwaters
               iO = 3*iw-2; iH1 = 3*iw-1; iH2 = 3*iw
               r1(1:3,1:3) = RRt(1:3, (/iO, iH1, iH2/), ib)
               call pot_nasa(r1, dr1, e1, box, boxi)
               dRRfast(1:3, (/iO, iH1, iH2/), ib) = dr1
               if (tintra == intra) then
                  Upostep(ib) = Upostep(ib) + e1
                  roh1 = RRc(1:3, iH1) - RRc(1:3, iO)
                  roh1 = roh1 - box*anint(roh1*boxi)
                  roh2 = RRc(1:3, iH2) - RRc(1:3, iO)
                  roh2 = roh2 - box*anint(roh2*boxi)
                  virialcmon = virialcmon + dot_product(roh1, dr1(:,2))
                  virialcmon = virialcmon + dot_product(roh2, dr1(:,3))
                  roh1 = RRt(1:3, iH1, ib) - RRt(1:3, iO, ib)
                  roh1 = roh1 - box*anint(roh1*boxi)
                  roh2 = RRt(1:3, iH2, ib) - RRt(1:3, iO, ib)
                  roh2 = roh2 - box*anint(roh2*boxi)
                  virialmon = virialmon + dot_product(roh1, dr1(:,2))
                  virialmon = virialmon + dot_product(roh2, dr1(:,3))
               endif
            enddo
         enddo
         PPt = PPt - MASSCON*dRRfast*delt2fast
      enddo
      RRc = sum(RRt,3)/Nbeads
      PPc = sum(PPt,3)/Nbeads
      call PBCs(RRt, RRc)
      CONTRACTION = .true.
      call pot_ttm(RRc, RRc, En, dRR, vtmp, vc, dipI, dipE, chgb, it)
      CONTRACTION = .false.
      Upostep = Upostep + En
      do ib = 1, nb
         dRRt(:,:,ib) = dRR
      enddo
      write(11) sum(Upostep), virialmon + vtmp(1,1) + vtmp(2,2) + vtmp(3,3), &
                virialcmon/Nbeads + vc
      call calc_uk
      PPt = PPt - MASSCON*dRRt*delt2
      if (BEADTHERMOSTAT) then
         call record_noise(noise, n)
         call bead_thermostat
         write(11) noise
      endif
      call calc_uk
      write(11) uk
   enddo
   write(11) RRt, PPt, dRRt, RRc

case (6)
   call init_pot
   nf = ih(11)
   allocate(xs(nf), drsq(nf), apol(nf))
   read(10) xs
   read(10) drsq
   read(10) apol
   do ia = 1, nf
      write(11) erfc(xs(ia))
      if (xs(ia) >= 0.d0) then
         write(11) gammq2_3(xs(ia))
      else
         write(11) 0.d0
      endif
      call PBCsmear01(drsq(ia), apol(ia), aewald, ts0, ts1)
      write(11) ts0, ts1
      call PBCsmear2(drsq(ia), apol(ia), aewald, ts1, ts2)
      write(11) ts1, ts2
      call PBCsmear3(drsq(ia), apol(ia), aewald, ts1, ts2, ts3)
      write(11) ts1, ts2, ts3
      call self1(drsq(ia), aewald, ts0, ts1)
      write(11) ts0, ts1
      call self2(drsq(ia), aewald, ts1, ts2)
      write(11) ts1, ts2
   enddo

case (7)
   call init_pot
   write(11) aewald, dble(hmax), dble(kmax), dble(lmax), dble(nkvec), Uvdw_lrc, polfac, &
             shiftA, shiftB, const_ts1, const_ts2, const_ts3, volume
   write(11) kvec(:,1:nkvec), Ak(1:nkvec), dble(k_hkl(:,1:nkvec))

case (8)
   call init_pot
   nf = ih(11)
   do ia = 1, nf
      read(10) r1
      call pot_nasa(r1, dr1, e1, box, boxi)
      call dms_nasa(r1, q3, dq3, box, boxi)
      write(11) e1, dr1, q3, dq3
   enddo

case default
   stop 'ref_driver: unknown mode'
end select

close(11)
close(10)

contains

subroutine seed_rng(sd)
   integer, intent(in) :: sd
   integer :: k2
   call random_seed(size=nseed)
   allocate(seedst(nseed))
   seedst = sd + 357*(/ (k2-1, k2=1, nseed) /)
   call random_seed(put=seedst)
end subroutine seed_rng

! Draws the n standard normals the next n rand_norm calls will return, then
! rewinds the generator so the reference routine consumes the same stream.
! Needs rand_norm's one-value cache to be empty on entry, i.e. n even.
subroutine record_noise(buf, nn)
   integer, intent(in) :: nn
   double precision, intent(out) :: buf(nn)
   integer, allocatable :: st(:)
   integer :: k2
   allocate(st(nseed))
   call random_seed(get=st)
   do k2 = 1, nn
      buf(k2) = rand_norm(1.d0)
   enddo
   call random_seed(put=st)
   deallocate(st)
end subroutine record_noise

end program ref_driver
