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
! The driver loop mirrors main_pattern.f95:170-188 (set-up) and :202-225 (loop);
! do_setup 1/2 is that set-up without/with the sort, 3/4 the same followed by the
! force evaluation of restart_mod.f95:425-495.
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
use forces_mod
implicit none
integer, parameter :: MAXS = 100
! Mersenne-Twister streams the modules draw from: one per OpenMP thread
! (mt(:,thread), mpc_rotation_mod.f95:186-189); 1 unless the library was
! compiled with -fopenmp and ref_set_mt_threads named the thread count
integer :: nthr_mt = 1
contains

subroutine ref_set_mt_threads(n) bind(C, name="ref_set_mt_threads")
  integer(c_int32_t), value :: n
  nthr_mt = n
end subroutine

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
  integer(c_int32_t) :: mt(624,nthr_mt), mti(nthr_mt)

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
    if (do_setup == 2 .or. do_setup == 4) call sort_hilbert(vol, NT, box, x, y, z, flag, vx, vy, vz)
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
    if (do_setup >= 3) call restart_forces(NT, numdims, box, rp, flag, x, y, z, xd, yd, zd, fxd, fyd, fzd, &
                                           fpx, fpy, fpz, pot, nbnum, nb, TPotential)
    if (do_setup >= 3) rp(31) = TPotential
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

! restart_mod.f95:425-495 after the lists: forces, pair forces and potential
! recomputed from the restored positions, with the reference's own force
! routines and the constants of restart_mod.f95:211-251
subroutine restart_forces(NT, numdims, box, rp, flag, x, y, z, xd, yd, zd, fxd, fyd, fzd, &
                          fpx, fpy, fpz, pot, nbnum, nb, TPotential)
  integer(c_int32_t) :: NT
  integer :: numdims
  real(c_double) :: box(3), rp(31), TPotential
  integer(c_int32_t) :: flag(NT)
  real(c_double) :: x(NT), y(NT), z(NT)
  real(c_double), dimension(numdims,2) :: xd, yd, zd, fxd, fyd, fzd
  real(c_double), dimension(:,:,:) :: fpx, fpy, fpz, pot
  integer(c_int64_t) :: nb(:,:,:), nbnum(numdims,2)
  real(c_double) :: Rc, Rn, cutoff, Rm_delta, Rw, Ew, Em
  real(c_double) :: Rn6, Rn12, r6Rn6, r12Rn12, Rc6, Rc12, r6Rc6, r12Rc12, cutoffC2, cutoffN2
  real(c_double) :: Rcc6, Rcc12, Rcn6, Rcn12, Rnn6, Rnn12, r6Rcc6, r12Rcc12, r6Rcn6, r12Rcn12, r6Rnn6, r12Rnn12
  real(c_double) :: mot_cutoff_CC2, mot_cutoff_CN2, mot_cutoff_NN2
  real(c_double) :: Rwc3, Rwc9, r3Rwc3, r9Rwc9, Rwn3, Rwn9, r3Rwn3, r9Rwn9, wall_cutoffC2, wall_cutoffN2
  real(c_double), allocatable :: FmotX(:,:,:,:), FmotY(:,:,:,:), FmotZ(:,:,:,:), Fmot_pot(:,:,:,:)
  real(c_double) :: fwallz(numdims,2), wall_pot(numdims,2)
  integer :: k, i, j, l
  integer(c_int64_t) :: jj
  integer, dimension(2) :: sphere = (/1,2/)

  Rc = rp(12); Rn = rp(13); cutoff = rp(11); Rm_delta = rp(22); Rw = rp(20); Ew = rp(21); Em = rp(23)
  Rn6 = Rn**6; Rn12 = Rn**12; r6Rn6 = 6.d0*Rn6; r12Rn12 = 12.d0*Rn12
  Rc6 = Rc**6; Rc12 = Rc**12; r6Rc6 = 6.d0*Rc6; r12Rc12 = 12.d0*Rc12
  cutoffC2 = (Rc * cutoff)**2
  cutoffN2 = (Rn * cutoff)**2
  Rcc6 = (Rc+Rc+Rm_delta)**6; Rcc12 = (Rc+Rc+Rm_delta)**12
  Rcn6 = (Rc+Rn+Rm_delta)**6; Rcn12 = (Rc+Rn+Rm_delta)**12
  Rnn6 = (Rn+Rn+Rm_delta)**6; Rnn12 = (Rn+Rn+Rm_delta)**12
  r6Rcc6 = 6.d0*Rcc6; r12Rcc12 = 12.d0*Rcc12; r6Rcn6 = 6.d0*Rcn6; r12Rcn12 = 12.d0*Rcn12
  r6Rnn6 = 6.d0*Rnn6; r12Rnn12 = 12.d0*Rnn12
  mot_cutoff_CC2 = ((Rc+Rc+Rm_delta)*cutoff)**2
  mot_cutoff_CN2 = ((Rc+Rn+Rm_delta)*cutoff)**2
  mot_cutoff_NN2 = ((Rn+Rn+Rm_delta)*cutoff)**2
  Rwc3 = (Rw+Rc)**3; Rwc9 = (Rw+Rc)**9; r3Rwc3 = 3.d0*Rwc3; r9Rwc9 = 9.d0*Rwc9
  Rwn3 = (Rw+Rn)**3; Rwn9 = (Rw+Rn)**9; r3Rwn3 = 3.d0*Rwn3; r9Rwn9 = 9.d0*Rwn9
  wall_cutoffC2 = ((Rw+Rc)*3.d0**(1./6.))**2
  wall_cutoffN2 = ((Rw+Rn)*3.d0**(1./6.))**2
  allocate(FmotX(numdims,2,numdims,2), FmotY(numdims,2,numdims,2), FmotZ(numdims,2,numdims,2), &
           Fmot_pot(numdims,2,numdims,2))

  fxd = 0.; fyd = 0.; fzd = 0.
  fwallz = 0.; wall_pot = 0.
  FmotX = 0.; FmotY = 0.; FmotZ = 0.; Fmot_pot = 0.
  fpx = 0.; fpy = 0.; fpz = 0.; pot = 0.
  TPotential = 0.
  do k = 1, numdims
    call LJ_mot_sol(sphere(1), NT, box, flag, x, y, z, xd(k,1), yd(k,1), zd(k,1), fpx(k,1,:), fpy(k,1,:), &
                    fpz(k,1,:), pot(k,1,:), rp(24), rp(25), cutoffC2, rp(14), rp(15), rp(16), Rc6, Rc12, &
                    r12Rc12, r6Rc6, nbnum(k,1), nb(k,1,:))
    call LJ_mot_sol(sphere(2), NT, box, flag, x, y, z, xd(k,2), yd(k,2), zd(k,2), fpx(k,2,:), fpy(k,2,:), &
                    fpz(k,2,:), pot(k,2,:), rp(24), rp(25), cutoffN2, rp(17), rp(18), rp(19), Rn6, Rn12, &
                    r12Rn12, r6Rn6, nbnum(k,2), nb(k,2,:))
    do j = 1, numdims
      if (j /= k) then
        call LJ_mot_mot(box, xd(k,1), yd(k,1), zd(k,1), xd(j,1), yd(j,1), zd(j,1), FmotX(k,1,j,1), &
                        FmotY(k,1,j,1), FmotZ(k,1,j,1), Fmot_pot(k,1,j,1), mot_cutoff_CC2, Em, Rcc6, Rcc12, &
                        r6Rcc6, r12Rcc12)
        call LJ_mot_mot(box, xd(k,1), yd(k,1), zd(k,1), xd(j,2), yd(j,2), zd(j,2), FmotX(k,1,j,2), &
                        FmotY(k,1,j,2), FmotZ(k,1,j,2), Fmot_pot(k,1,j,2), mot_cutoff_CN2, Em, Rcn6, Rcn12, &
                        r6Rcn6, r12Rcn12)
        call LJ_mot_mot(box, xd(k,2), yd(k,2), zd(k,2), xd(j,1), yd(j,1), zd(j,1), FmotX(k,2,j,1), &
                        FmotY(k,2,j,1), FmotZ(k,2,j,1), Fmot_pot(k,2,j,1), mot_cutoff_CN2, Em, Rcn6, Rcn12, &
                        r6Rcn6, r12Rcn12)
        call LJ_mot_mot(box, xd(k,2), yd(k,2), zd(k,2), xd(j,2), yd(j,2), zd(j,2), FmotX(k,2,j,2), &
                        FmotY(k,2,j,2), FmotZ(k,2,j,2), Fmot_pot(k,2,j,2), mot_cutoff_NN2, Em, Rnn6, Rnn12, &
                        r6Rnn6, r12Rnn12)
      end if
    end do
    call LJ_mot_wall(box, zd(k,1), fwallz(k,1), Ew, Rwc3, Rwc9, r3Rwc3, r9Rwc9, wall_cutoffC2, wall_pot(k,1))
    call LJ_mot_wall(box, zd(k,2), fwallz(k,2), Ew, Rwn3, Rwn9, r3Rwn3, r9Rwn9, wall_cutoffN2, wall_pot(k,2))
  end do
  do k = 1, numdims
    do i = 1, 2
      do jj = 1, nbnum(k,i)
        TPotential = TPotential + pot(k,i,jj)
        fxd(k,i) = fxd(k,i) - fpx(k,i,jj)
        fyd(k,i) = fyd(k,i) - fpy(k,i,jj)
        fzd(k,i) = fzd(k,i) - fpz(k,i,jj)
      end do
      do j = 1, numdims
        do l = 1, 2
          fxd(k,i) = fxd(k,i) + FmotX(k,i,j,l)
          fyd(k,i) = fyd(k,i) + FmotY(k,i,j,l)
          fzd(k,i) = fzd(k,i) + FmotZ(k,i,j,l)
          if (j < k) TPotential = TPotential + Fmot_pot(k,i,j,l)
        end do
      end do
      TPotential = TPotential + wall_pot(k,i)
      fzd(k,i) = fzd(k,i) + fwallz(k,i)
    end do
  end do
  deallocate(FmotX, FmotY, FmotZ, Fmot_pot)
end subroutine

! ------------------------------------------------------------ restart files
! Path arguments arrive as (bytes, length) and become a Fortran string.
function c_path(p, n) result(s)
  integer(c_int32_t), value :: n
  character(kind=c_char) :: p(n)
  character(len=n) :: s
  integer :: i
  do i = 1, n
    s(i:i) = p(i)
  end do
end function

! mt_state.dat through the reference's own mtsavef / mtgetf (mtmod.f90:152-229)
subroutine ref_mtsavef(p, n, nthr, mt, mti) bind(C, name="ref_mtsavef")
  integer(c_int32_t), value :: n, nthr
  character(kind=c_char) :: p(n)
  integer(c_int32_t) :: mt(624, nthr), mti(nthr)
  call mtsavef(mt, mti, c_path(p, n), 'u')
end subroutine

subroutine ref_mtgetf(p, n, nthr, mt, mti) bind(C, name="ref_mtgetf")
  integer(c_int32_t), value :: n, nthr
  character(kind=c_char) :: p(n)
  integer(c_int32_t) :: mt(624, nthr), mti(nthr)
  call mtgetf(mt, mti, c_path(p, n), 'u')
end subroutine

! timestep.dat with the statements of main_pattern.f95:243-249 (write) and
! :130-136 (read); cstep is integer(4), the totals integer(8)
subroutine ref_timestep_write(p, n, cstep, tot) bind(C, name="ref_timestep_write")
  integer(c_int32_t), value :: n, cstep
  character(kind=c_char) :: p(n)
  integer(c_int64_t) :: tot(3)
  open(unit=10, file=c_path(p, n), status='replace')
  write(10,*) cstep
  write(10,*) tot(1)
  write(10,*) tot(2)
  write(10,*) tot(3)
  close(10)
end subroutine

subroutine ref_timestep_read(p, n, cstep, tot) bind(C, name="ref_timestep_read")
  integer(c_int32_t), value :: n
  character(kind=c_char) :: p(n)
  integer(c_int32_t) :: cstep
  integer(c_int64_t) :: tot(3)
  open(unit=10, file=c_path(p, n), status='old', action='read')
  read(10,*) cstep
  read(10,*) tot(1)
  read(10,*) tot(2)
  read(10,*) tot(3)
  close(10)
end subroutine

! motor_positions.dat with the statements of initial_mod.f95:79,118 (write)
! and restart_mod.f95:410-414 (read); cm is motorCM(numdims,3)
subroutine ref_motor_positions_write(p, n, nd, cm) bind(C, name="ref_motor_positions_write")
  integer(c_int32_t), value :: n, nd
  character(kind=c_char) :: p(n)
  real(c_double) :: cm(nd, 3)
  integer :: i
  open(unit=10, file=c_path(p, n), status="replace", position="append")
  do i = 1, nd
    write(10,*) cm(i,1), cm(i,2), cm(i,3)
  end do
  close(10)
end subroutine

subroutine ref_motor_positions_read(p, n, nd, cm) bind(C, name="ref_motor_positions_read")
  integer(c_int32_t), value :: n, nd
  character(kind=c_char) :: p(n)
  real(c_double) :: cm(nd, 3)
  integer :: i
  open(unit=10, file=c_path(p, n), status="old")
  do i = 1, nd
    read(10,*) cm(i,1), cm(i,2), cm(i,3)
  end do
  close(10)
end subroutine

end module ref_collective_harness
