! Test harness (not part of the product): the reference's own checkpoint
! writer and reader (save_checkpoint / load_checkpoint, InputOutput.f90:987-1074).
!   ref_img write <in.bin> <out.img>   in.bin: Natoms, Nbeads (int32), massO,
!                                      massH, box(3), init_time, delt, t (float64),
!                                      RRt, PPt (3,Natoms,Nbeads) float64
!   ref_img read  <m.bin> <in.img> <out.bin>
!                                      m.bin: as in.bin up to massO, massH;
!                                      out.bin: RRt, PPt, RRc (after PBCs)
! The header line reads use the formats of read_and_initialize_all_nodes
! (InputOutput.f90:86-93).
program ref_img
use system_mod
use consts
use main_stuff
use InputOutput, only: save_checkpoint, load_checkpoint
implicit none
character(len=512) :: op, fin, fout, fimg
integer :: hdr(2), nb_in
double precision :: d(8), tt
call get_command_argument(1, op)
call get_command_argument(2, fin)
call get_command_argument(3, fout)
call get_command_argument(4, fimg)
if (trim(op) == 'write') then
   open(10, file=trim(fin), access='stream', form='unformatted', status='old')
   read(10) hdr
   Natoms = hdr(1); Nbeads = hdr(2); Nwaters = Natoms/3
   read(10) d
   massO = d(1); massH = d(2); box = d(3:5); init_time = d(6); delt = d(7); tt = d(8)
   imassO = 1.d0/massO; imassH = 1.d0/massH
   allocate(RRt(3,Natoms,Nbeads), PPt(3,Natoms,Nbeads))
   read(10) RRt
   read(10) PPt
   close(10)
   open(11, file=trim(fout), status='replace')
   call save_checkpoint(11, RRt, PPt, 0.d0, nint(tt), delt)
   close(11)
else
   open(10, file=trim(fin), access='stream', form='unformatted', status='old')
   read(10) hdr
   read(10) d(1:2)
   close(10)
   massO = d(1); massH = d(2)
   open(12, file=trim(fout), status='old')
   read(12,'(2i10)') Natoms, nb_in
   Nbeads = nb_in; Nwaters = Natoms/3
   read(12,'(f12.6,2x,f12.6,3(1x,f12.6))') init_time, box
   boxi = 1.d0/box
   allocate(RRt(3,Natoms,Nbeads), PPt(3,Natoms,Nbeads), RRc(3,Natoms), PPc(3,Natoms))
   BEADTHERMOSTAT = .false.; THERMOSTAT = .false.
   call load_checkpoint(12)
   close(12)
   open(11, file=trim(fimg), access='stream', form='unformatted', status='replace')
   write(11) RRt, PPt, RRc
   close(11)
endif
end program ref_img
