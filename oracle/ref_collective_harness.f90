! C-callable harness around the reference collective_code modules (compiled
! from /root/reference by oracle/Makefile into oracle/_ref).  TEST
! INFRASTRUCTURE ONLY: it lets the tests run the reference's own subroutines
! on small inputs to pin the C restatement in oracle/mpc_oracle.c.
!
! Scalars travel in two blocks:
!   ip(int64):  1 NT, 2 numdims, 3 MPCtstep, 4 sort_step, 5 react_dir, 6 cstep,
!               7 time_count, 8 time_penetrate, 9 vol, 10 max_nb, 11 out_num,
!               12 in_num, 13 mpcnumcell, 14 rxnnumcell
!   rp(double): 1-3 box, 4 MDt, 5 Ma, 6 Mb, 7 Ms, 8 Mc, 9 Mn, 10 Dcn, 11 cutoff,
!               12 Rc, 13 Rn, 14 EaC, 15 EbC, 16 EsC, 17 EaN, 18 EbN, 19 EsN,
!               20 Rw, 21 Ew, 22 Rm_delta, 23 Em, 24 RPab, 25 RPba, 26 skinrad,
!               27 rotang, 28 a0, 29 MPCt, 30 k1, 31 TPotential (out)
! The driver loop mirrors main_pattern.f95:170-188 (set-up) and :202-225 (loop).
module ref_collective_harness
use iso_c_binding
use mtmod
use neighbour_list_mod
use reactive_grid_mod
use evolve_mod
use distribute_mod
use mpc_rotation_mod
use reaction_mod
use sort_hilbert_mod
implicit none
integer, parameter :: MAXS = 100
contains

subroutine ref_mtmod_seed(seed, mt, mti) bind(C, name="ref_mtmod_seed")
  integer(c_int32_t), value :: seed
  integer(c_int32_t) :: mt(624), mti
  call sgrnd(seed, mt, mti)
end subroutine

subroutine ref_mtmod_fill(mt, mti, n, out) bind(C, name="ref_mtmod_fill")
  integer(c_int32_t) :: mt(624), mti
  integer(c_int64_t), value :: n
  real(c_double) :: out(n)
  integer(c_int64_t) :: i
  do i = 1, n
    out(i) = grnd(mt, mti)
  end do
end subroutine

subroutine ref_reactive_grid_setup(box, nbox, react_box, rtype) bind(C, name="ref_reactive_grid_setup")
  real(c_double) :: box(3)
  integer(c_int64_t), value :: nbox
  integer(c_int32_t) :: react_box(nbox)
  integer(c_int32_t), value :: rtype
  integer :: t
  t = rtype
  call reactive_grid_setup(box, react_box, t)
end subroutine

subroutine ref_neighbour_list(NT, box, cut, skin, x, y, z, xd, yd, zd, num, nb, maxnb) &
    bind(C, name="ref_neighbour_list")
  integer(c_int32_t), value :: NT
  real(c_double) :: box(3)
  real(c_double), value :: cut, skin, xd, yd, zd
  real(c_double) :: x(NT), y(NT), z(NT)
  integer(c_int64_t) :: num
  integer(c_int64_t), value :: maxnb
  integer(c_int64_t) :: nb(maxnb)
  num = 0
  call neighbour_list(NT, box, cut, skin, x, y, z, xd, yd, zd, num, nb)
end subroutine

subroutine ref_inout(NT, numdims, box, cutC, cutN, skin, x, y, z, xd, yd, zd, onum, ol, inum, il, ilmax) &
    bind(C, name="ref_inout")
  integer(c_int32_t), value :: NT, numdims
  real(c_double) :: box(3)
  real(c_double), value :: cutC, cutN, skin
  real(c_double) :: x(NT), y(NT), z(NT), xd(numdims,2), yd(numdims,2), zd(numdims,2)
  integer(c_int64_t) :: onum, inum
  integer(c_int64_t), value :: ilmax
  integer(c_int64_t) :: ol(NT), il(ilmax)
  integer :: nd
  nd = numdims
  onum = 0
  inum = 0
  call inout_particle_list(NT, box, cutC, cutN, skin, x, y, z, xd, yd, zd, onum, ol, inum, il, nd)
end subroutine

subroutine ref_sort_hilbert(vol, NT, box, x, y, z, flag, vx, vy, vz) bind(C, name="ref_sort_hilbert")
  integer(c_int64_t), value :: vol
  integer(c_int32_t), value :: NT
  real(c_double) :: box(3), x(NT), y(NT), z(NT), vx(NT), vy(NT), vz(NT)
  integer(c_int32_t) :: flag(NT)
  call sort_hilbert(vol, NT, box, x, y, z, flag, vx, vy, vz)
end subroutine

subroutine ref_mpc_rotation(NT, rotang, numcell, a0, box, x, y, z, vx, vy, vz, vxcm, vycm, vzcm, tcs, ncn, &
    uvx, uvy, uvz, mt, mti) bind(C, name="ref_mpc_rotation")
  integer(c_int32_t), value :: NT
  real(c_double), value :: rotang, a0
  integer(c_int64_t), value :: numcell
  real(c_double) :: box(3), x(NT), y(NT), z(NT), vx(NT), vy(NT), vz(NT)
  real(c_double) :: vxcm(numcell), vycm(numcell), vzcm(numcell), tcs(numcell)
  real(c_double) :: uvx(numcell), uvy(numcell), uvz(numcell)
  integer(c_int64_t) :: ncn(NT)
  integer(c_int32_t) :: mt(624,1), mti(1)
  call mpc_rotation(NT, rotang, numcell, a0, box, x, y, z, vx, vy, vz, vxcm, vycm, vzcm, tcs, ncn, &
                    uvx, uvy, uvz, mt, mti)
end subroutine

subroutine ref_reaction_simple(MPCt, numcell, vol, k1, NT, flag, tB, sB, mt, mti) bind(C, name="ref_reaction_simple")
  real(c_double), value :: MPCt, k1
  integer(c_int64_t), value :: numcell, vol
  integer(c_int32_t), value :: NT
  integer(c_int32_t) :: flag(NT)
  integer(c_int64_t) :: tB(vol), sB(vol, MAXS)
  integer(c_int32_t) :: mt(624,1), mti(1)
  call reaction_simple(MPCt, numcell, k1, flag, tB, sB, mt, mti)
end subroutine

! Full state of one system; arrays in the reference's own shapes.
subroutine ref_co_run(ip, rp, nsteps, do_setup, flag, x, y, z, vx, vy, vz, xd, yd, zd, vxd, vyd, vzd, &
    fxd, fyd, fzd, fpx, fpy, fpz, pot, nb, nbnum, msd, ol, il, nrb, react_box, &
    tA, tB, tS, sA, sB, sS, vxcm, vycm, vzcm, tcs, ncn, uvx, uvy, uvz, mt, mti) bind(C, name="ref_co_run")
  integer(c_int64_t) :: ip(14)
  real(c_double) :: rp(31)
  integer(c_int64_t), value :: nsteps
  integer(c_int32_t), value :: do_setup
  integer(c_int32_t) :: flag(ip(1))
  real(c_double) :: x(ip(1)), y(ip(1)), z(ip(1)), vx(ip(1)), vy(ip(1)), vz(ip(1))
  real(c_double), dimension(ip(2),2) :: xd, yd, zd, vxd, vyd, vzd, fxd, fyd, fzd
  real(c_double), dimension(ip(2),2,ip(10)) :: fpx, fpy, fpz, pot
  integer(c_int64_t) :: nb(ip(2),2,ip(10)), nbnum(ip(2),2)
  real(c_double) :: msd(ip(1))
  integer(c_int64_t) :: ol(ip(1)), il(2*ip(2)*ip(10))
  integer(c_int64_t), value :: nrb
  integer(c_int32_t) :: react_box(nrb)
  integer(c_int64_t) :: tA(ip(9)), tB(ip(9)), tS(ip(9))
  integer(c_int64_t) :: sA(ip(9),MAXS), sB(ip(9),MAXS), sS(ip(9),MAXS)
  real(c_double) :: vxcm(ip(13)), vycm(ip(13)), vzcm(ip(13)), tcs(ip(13))
  real(c_double) :: uvx(ip(13)), uvy(ip(13)), uvz(ip(13))
  integer(c_int64_t) :: ncn(ip(1))
  integer(c_int32_t) :: mt(624,1), mti(1)

  integer(c_int32_t) :: NT, cstep, MPCtstep
  integer :: numdims, sort_step, react_dir, time_count, time_penetrate, k
  integer(c_int64_t) :: vol, onum, inum, mpcnumcell, rxnnumcell, step, time
  real(c_double) :: box(3), cutoffC, cutoffN, cutoffC2, cutoffN2, TPotential
  real(c_double), dimension(ip(2),3) :: motorCM
  integer, dimension(2) :: sphere = (/1,2/)

  NT = int(ip(1), c_int32_t); numdims = int(ip(2)); MPCtstep = int(ip(3), c_int32_t)
  sort_step = int(ip(4)); react_dir = int(ip(5)); cstep = int(ip(6), c_int32_t)
  time_count = int(ip(7)); time_penetrate = int(ip(8)); vol = ip(9)
  onum = ip(11); inum = ip(12); mpcnumcell = ip(13); rxnnumcell = ip(14)
  box = rp(1:3)
  cutoffC = rp(12) * rp(11)
  cutoffN = rp(13) * rp(11)
  cutoffC2 = cutoffC**2
  cutoffN2 = cutoffN**2
  motorCM = 0.d0

  if (do_setup /= 0) then
    if (do_setup == 2) call sort_hilbert(vol, NT, box, x, y, z, flag, vx, vy, vz)
    nbnum = 0
    onum = 0
    inum = 0
    do k = 1, numdims
      call neighbour_list(NT, box, cutoffC, rp(26), x, y, z, xd(k,1), yd(k,1), zd(k,1), nbnum(k,1), nb(k,1,:))
      call neighbour_list(NT, box, cutoffN, rp(26), x, y, z, xd(k,2), yd(k,2), zd(k,2), nbnum(k,2), nb(k,2,:))
    end do
    call inout_particle_list(NT, box, cutoffC, cutoffN, rp(26), x, y, z, xd, yd, zd, onum, ol, inum, il, numdims)
    msd = 0
    time_penetrate = floor((0.5*rp(26) / (sqrt(maxval(vx*vx+vy*vy+vz*vz)))) / rp(4))
    time_count = 0
  end if

  do step = 1, nsteps
    time = step
    cstep = cstep + 1
    time_count = time_count + 1
    call evolve(cstep, time, NT, rp(4), MPCtstep, numdims, rp(5), rp(6), rp(7), rp(8), rp(9), rp(10), box, rp(11), &
                rp(12), rp(13), rp(14), rp(15), rp(16), rp(17), rp(18), rp(19), rp(20), rp(21), rp(22), rp(23), &
                rp(24), rp(25), sphere, flag, x, y, z, vx, vy, vz, xd, yd, zd, vxd, vyd, vzd, fxd, fyd, fzd, &
                motorCM, TPotential, pot, fpx, fpy, fpz, rp(26), msd, nbnum, nb, time_count, time_penetrate, &
                onum, ol, inum, il, vol, sort_step, react_box, react_dir)
    if (mod(cstep, MPCtstep) == 0) then
      call distribute(NT, numdims, rp(11), rp(12), rp(13), flag, x, y, z, xd, yd, zd, box, tA, tB, tS, &
                      sA, sB, sS, cutoffC2, cutoffN2, cstep, onum, ol, inum, il)
      call mpc_rotation(NT, rp(27), mpcnumcell, rp(28), box, x, y, z, vx, vy, vz, vxcm, vycm, vzcm, &
                        tcs, ncn, uvx, uvy, uvz, mt, mti)
      call reaction_simple(rp(29), rxnnumcell, rp(30), flag, tB, sB, mt, mti)
    end if
    rp(31) = TPotential
  end do

  ip(6) = cstep; ip(7) = time_count; ip(8) = time_penetrate; ip(11) = onum; ip(12) = inum
end subroutine

end module ref_collective_harness
