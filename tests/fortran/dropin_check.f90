! Calls the Fortran drop-in modules of main_amd/fortran in the order of
! collective_code's set-up and first collision (main_pattern.f95:171-188 and
! :219-224) on a system read from `in.bin`, and writes every output to
! `out.bin` (stream files, native byte order).  tests/test_gpu_fortran.py
! writes the input and compares the output with the oracle.
program dropin_check
  use iso_c_binding
  use sort_hilbert_mod
  use neighbour_list_mod
  use distribute_mod
  use mpc_rotation_mod
  use reaction_mod
  implicit none
  integer, parameter :: dp = selected_real_kind(15, 307)
  integer, parameter :: max_nb = 4000
  integer(4) :: NT, numdims
  integer(8) :: vol, mpcnumcell, rxnnumcell, maxs
  real(dp) :: box(3), cutoff, Rc, Rn, skinrad, rotang, a0, MPCt, k1
  real(dp), allocatable :: x(:), y(:), z(:), vx(:), vy(:), vz(:), xd(:, :), yd(:, :), zd(:, :)
  integer(4), allocatable :: flag(:)
  integer, allocatable :: mt(:, :), mti(:)
  integer(8), allocatable :: nbnum(:, :), nb(:, :, :), ol(:), il(:), tA(:), tB(:), tS(:), sA(:, :), sB(:, :), sS(:, :)
  integer(8), allocatable :: ncn(:)
  real(dp), allocatable :: vxcm(:), vycm(:), vzcm(:), tcs(:), uvx(:), uvy(:), uvz(:)
  integer(8) :: onum, inum
  integer :: u, k, i, cstep
  real(dp) :: R(2)

  open(newunit=u, file="in.bin", access="stream", form="unformatted", status="old")
  read(u) NT, numdims, vol, mpcnumcell, rxnnumcell, maxs
  read(u) box, cutoff, Rc, Rn, skinrad, rotang, a0, MPCt, k1
  allocate(x(NT), y(NT), z(NT), vx(NT), vy(NT), vz(NT), flag(NT))
  allocate(xd(numdims, 2), yd(numdims, 2), zd(numdims, 2), mt(624, 1), mti(1))
  read(u) x, y, z, vx, vy, vz, flag, xd, yd, zd, mt, mti
  close(u)
  allocate(nbnum(numdims, 2), nb(numdims, 2, max_nb), ol(NT), il(NT))
  allocate(tA(vol), tB(vol), tS(vol), sA(vol, maxs), sB(vol, maxs), sS(vol, maxs), ncn(NT))
  allocate(vxcm(mpcnumcell), vycm(mpcnumcell), vzcm(mpcnumcell), tcs(mpcnumcell))
  allocate(uvx(mpcnumcell), uvy(mpcnumcell), uvz(mpcnumcell))
  nb = 0; ol = 0; il = 0; sA = 0; sB = 0; sS = 0
  R = [Rc, Rn]

  call sort_hilbert(vol, NT, box, x, y, z, flag, vx, vy, vz)
  do k = 1, numdims
    do i = 1, 2
      nbnum(k, i) = 0
      call neighbour_list(NT, box, cutoff * R(i), skinrad, x, y, z, xd(k, i), yd(k, i), zd(k, i), nbnum(k, i), &
                          nb(k, i, :))
    end do
  end do
  onum = 0; inum = 0
  call inout_particle_list(NT, box, cutoff * Rc, cutoff * Rn, skinrad, x, y, z, xd, yd, zd, onum, ol, inum, il, &
                           numdims)
  cstep = 0
  call distribute(NT, numdims, cutoff, Rc, Rn, flag, x, y, z, xd, yd, zd, box, tA, tB, tS, sA, sB, sS, &
                  (cutoff * Rc)**2, (cutoff * Rn)**2, cstep, onum, ol, inum, il)
  call mpc_rotation(NT, rotang, mpcnumcell, a0, box, x, y, z, vx, vy, vz, vxcm, vycm, vzcm, tcs, ncn, &
                    uvx, uvy, uvz, mt, mti)
  call reaction_simple(MPCt, rxnnumcell, k1, flag, tB, sB, mt, mti)

  open(newunit=u, file="out.bin", access="stream", form="unformatted", status="replace")
  write(u) x, y, z, vx, vy, vz, flag, nbnum, nb, onum, inum, ol, il, tA, tB, tS, sA, sB, sS
  write(u) vxcm, vycm, vzcm, tcs, ncn, uvx, uvy, uvz, mt, mti
  close(u)
end program dropin_check
