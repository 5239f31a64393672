! Test harness (not part of the product): runs the reference's own input-file
! reader (input_file_reader.f90:30-431) on the file named by the first
! argument and prints every variable it sets as "key value" lines.
program ref_input
use system_mod
use main_stuff
use input_file_reader
implicit none
character(len=512) :: fname
call get_command_argument(1, fname)
lun = 41
open(lun, file=trim(fname), status='old', action='read')
call read_inputfile()
write(*,'(a)') '#parsed'
call wl('coord_out', coord_out); call wl('momenta_out', momenta_out)
call wl('dip_out', dip_out); call wl('Edip_out', Edip_out); call wl('TD_out', TD_out)
call wl('images_out', images_out); call wl('IMAGEDIPOLESOUT', IMAGEDIPOLESOUT)
call wl('BOXSIZEOUT', BOXSIZEOUT); call wl('TP_out', TP_out); call wl('ENERGYOUT', ENERGYOUT)
call wl('HISTOUT', HISTOUT); call wi('hist_out_freq', hist_out_freq)
call wl('CALCGEOMETRY', CALCGEOMETRY); call wl('CALCDIFFUSION', CALCDIFFUSION)
call wi('read_method', read_method); call wl('CALCIRSPECTRA', CALCIRSPECTRA)
call wl('CALCDOS', CALCDOS); call wl('CALC_RADIUS_GYRATION', CALC_RADIUS_GYRATION)
call wl('DIELECTRICOUT', DIELECTRICOUT); call wl('CHARGESOUT', CHARGESOUT)
call wl('WRITECHECKPOINTS', WRITECHECKPOINTS); call wi('td_freq', td_freq)
call wi('tp_freq', tp_freq); call wi('ti_freq', ti_freq); call wi('t_freq', t_freq)
call wi('checkpoint_freq', checkpoint_freq); call ws('siesta_name', siesta_name)
call wi('num_SIESTA_nodes', num_SIESTA_nodes); call wd('Rc', Rc); call wd('rc1', rc1)
call wd('eps_ewald', eps_ewald); call wi('polar_maxiter', polar_maxiter)
call wd('polar_sor', polar_sor); call wd('polar_eps', polar_eps)
call wl('guess_initdip', guess_initdip); call wl('print_dipiters', print_dipiters)
call wl('GENVEL', GENVEL); call wl('THERMOSTAT', THERMOSTAT)
call wl('BEADTHERMOSTAT', BEADTHERMOSTAT); call wl('CENTROIDTHERMOSTAT', CENTROIDTHERMOSTAT)
call ws('bead_thermostat_type', bead_thermostat_type); call wd('tau', tau)
call wd('tau_centroid', tau_centroid); call wi('global_chain_length', global_chain_length)
call wi('bead_chain_length', bead_chain_length); call wd('temp', temp)
call wl('BAROSTAT', BAROSTAT); call wd('tau_P', tau_P); call wd('press', press)
call wl('PEQUIL', PEQUIL); call wd('setNMfreq', setNMfreq); call ws('PIMD_type', PIMD_type)
call wi('intra_timesteps', intra_timesteps); call wd('massO', massO); call wd('massH', massH)
call wl('RESTART', RESTART); call wl('SIESTA_MON_CALC', SIESTA_MON_CALC)
call wi('Nbeads', Nbeads); call ws('fconfig', fconfig); call ws('fsave', fsave)
call wi('eq_timesteps', eq_timesteps); call wi('run_timesteps', run_timesteps)
call wd('delt', delt); call wi('pot_model', pot_model); call ws('sys_label', sys_label)
call ws('mon_siesta_name', mon_siesta_name)
contains
subroutine wl(k, v)
  character(len=*), intent(in) :: k
  logical, intent(in) :: v
  write(*,'(a,1x,l1)') k, v
end subroutine
subroutine wi(k, v)
  character(len=*), intent(in) :: k
  integer, intent(in) :: v
  write(*,'(a,1x,i0)') k, v
end subroutine
subroutine wd(k, v)
  character(len=*), intent(in) :: k
  double precision, intent(in) :: v
  write(*,'(a,1x,es25.17e3)') k, v
end subroutine
subroutine ws(k, v)
  character(len=*), intent(in) :: k, v
  write(*,'(a,1x,a)') k, trim(v)
end subroutine
end program ref_input
