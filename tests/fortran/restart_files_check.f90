! The restart files written and read by libmpcx against the reference's own
! Fortran statements for them, compiled with flang:
!   timestep.dat         main_pattern.f95:130-136 (read), :243-249 (write)
!   mt_state.dat         mtmod.f90:163-166 (mtsavef 'u'), :214-217 (mtgetf 'u')
!   motor_positions.dat  initial_mod.f95:79,118 (write), restart_mod.f95:410-414 (read)
! Run in a scratch directory; prints "restart files ok" or stops with a code.
program restart_files_check
  use iso_c_binding
  use mpcx_c
  implicit none
  integer(c_int32_t) :: cs2, nthr
  integer(c_int64_t) :: a2, b2, s2
  integer :: mt(624, 2), mti(2), mt2(624, 2), mti2(2), i
  real(c_double) :: cm(3, 3), cm2(3, 3)

  call mpcx_check(mpcx_timestep_write('ts.dat'//c_null_char, 4000_c_int32_t, 1234567_c_int64_t, 89_c_int64_t, &
                                      0_c_int64_t), 'timestep_write')
  open(unit=10, file='ts.dat', status='old', action='read')
  read(10,*) cs2
  read(10,*) a2
  read(10,*) b2
  read(10,*) s2
  close(10)
  if (cs2 /= 4000 .or. a2 /= 1234567 .or. b2 /= 89 .or. s2 /= 0) stop 1
  open(unit=10, file='ts2.dat', status='replace')
  write(10,*) 77_c_int32_t
  write(10,*) 5_c_int64_t
  write(10,*) 6_c_int64_t
  write(10,*) 7_c_int64_t
  close(10)
  call mpcx_check(mpcx_timestep_read('ts2.dat'//c_null_char, cs2, a2, b2, s2), 'timestep_read')
  if (cs2 /= 77 .or. a2 /= 5 .or. b2 /= 6 .or. s2 /= 7) stop 2

  do i = 1, 624
    mt(i, 1) = i * 7919 - 3
    mt(i, 2) = -i
  end do
  mti = (/ 17, 625 /)
  call mpcx_check(mpcx_mt_state_write('mt.dat'//c_null_char, mt, mti, 2_c_int32_t), 'mt_state_write')
  open(unit=10, file='mt.dat', status='old', form='unformatted')
  read(10) mti2
  read(10) mt2
  close(10)
  if (any(mt2 /= mt) .or. any(mti2 /= mti)) stop 3
  open(unit=10, file='mt2.dat', status='replace', form='unformatted', position='append')
  write(10) mti
  write(10) mt
  close(10)
  mt2 = 0
  mti2 = 0
  call mpcx_check(mpcx_mt_state_read('mt2.dat'//c_null_char, mt2, mti2, 2_c_int32_t, nthr), 'mt_state_read')
  if (nthr /= 2 .or. any(mt2 /= mt) .or. any(mti2 /= mti)) stop 4

  cm = reshape((/ 1.0d0, -2.5d-7, 3.0d20, 0.1d0, 5.0d0, -6.25d0, 7.0d0, 8.0d-300, 1.0d0 / 3.0d0 /), (/ 3, 3 /))
  open(unit=10, file='mp.dat', status='replace', position='append')
  do i = 1, 3
    write(10,*) cm(i, 1), cm(i, 2), cm(i, 3)
  end do
  close(10)
  call mpcx_check(mpcx_motor_positions_read('mp.dat'//c_null_char, cm2, 3_c_int32_t), 'motor_positions_read')
  if (any(cm2 /= cm)) stop 5
  call mpcx_check(mpcx_motor_positions_write('mp2.dat'//c_null_char, cm, 3_c_int32_t), 'motor_positions_write')
  cm2 = 0
  open(unit=10, file='mp2.dat', status='old')
  do i = 1, 3
    read(10,*) cm2(i, 1), cm2(i, 2), cm2(i, 3)
  end do
  close(10)
  if (any(cm2 /= cm)) stop 6
  print '(a)', 'restart files ok'
end program
