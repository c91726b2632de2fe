! C-callable harness around the reference selkov_code subroutines (compiled from
! /root/reference by oracle/Makefile into oracle/_ref).  TEST INFRASTRUCTURE
! ONLY.  The reference keeps its state in module Global_SelkovDimer with a
! compile-time Vol = 64000 (a 40x40x40 box); this harness fills those globals
! from C arrays, runs the Main_SelkovDimer.f90:125-149 loop body for nsteps
! steps with the mt95 stream seeded by "seed", and copies the state back.
!   ip(int64): 1 NT, 2 space
!   rp(double): 1-3 box, 4 Mf, 5 Mg, 6 Mi, 7 Rc, 8 Rn, 9 Dcn, 10 EfC, 11 EgC,
!               12 EiC, 13 EfN, 14 EgN, 15 EiN, 16 RPfg, 17 RPgf, 18 A0,
!               19 MDt, 20 MPCt, 21-26 k1..k6, 27 Mc, 28 Mn, 29 TPotential (out)
!   dm(2,12): columns xd yd zd vxd vyd vzd fxd fyd fzd (spheres C=1, N=2)
module ref_selkov_harness
use iso_c_binding
use Global_SelkovDimer
use mt95
implicit none
contains

subroutine ref_mt95_fill(seed, n, out) bind(C, name="ref_mt95_fill")
  integer(c_int32_t), value :: seed
  integer(c_int64_t), value :: n
  real(c_double) :: out(n)
  real(8) :: r
  integer(c_int64_t) :: q
  call genrand_init(put=seed)
  do q = 1, n
    call genrand_real1(r)
    out(q) = r
  end do
end subroutine

subroutine ref_sk_run(ip, rp, seed, cstep0, nsteps, flg, xs, ys, zs, vxs, vys, vzs, fxs, fys, fzs, &
    pfxs, pfys, pfzs, dm) bind(C, name="ref_sk_run")
  integer(c_int64_t) :: ip(2)
  real(c_double) :: rp(29)
  integer(c_int32_t), value :: seed
  integer(c_int64_t), value :: cstep0, nsteps
  integer(c_int64_t) :: flg(ip(2))
  real(c_double), dimension(ip(2)) :: xs, ys, zs, vxs, vys, vzs, fxs, fys, fzs, pfxs, pfys, pfzs
  real(c_double) :: dm(2,12)
  integer(c_int64_t) :: it, q

  NT = ip(1); space = ip(2)
  Box = rp(1:3); Mf = rp(4); Mg = rp(5); Mi = rp(6); Rc = rp(7); Rn = rp(8); Dcn = rp(9)
  EfC = rp(10); EgC = rp(11); EiC = rp(12); EfN = rp(13); EgN = rp(14); EiN = rp(15)
  RPfg = rp(16); RPgf = rp(17); A0 = rp(18); MDt = rp(19); MPCt = rp(20)
  k1 = rp(21); k2 = rp(22); k3 = rp(23); k4 = rp(24); k5 = rp(25); k6 = rp(26); Mc = rp(27); Mn = rp(28)
  MPCtstep = INT(MPCt/MDt)
  CUTOFFc = Rc*CUTOFF
  CUTOFFn = Rn*CUTOFF
  CUTOFFc2 = CUTOFFc*CUTOFFc
  CUTOFFn2 = CUTOFFn*CUTOFFn
  HDcn = 0.5d0*Dcn
  do q = 1, 3
    HBox(q) = 0.5d0*Box(q)
    iBox(q) = INT(Box(q))
  end do

  ALLOCATE(Xdim(1:2),Ydim(1:2),Zdim(1:2),prXdim(1:2),prYdim(1:2),prZdim(1:2))
  ALLOCATE(Xsol(1:space),Ysol(1:space),Zsol(1:space),FLAG(1:space),R(1:2,1:3),dR(1:3))
  ALLOCATE(VXdim(1:2),VYdim(1:2),VZdim(1:2),prVXdim(1:2),prVYdim(1:2),prVZdim(1:2))
  ALLOCATE(VXsol(1:space),VYsol(1:space),VZsol(1:space),FXsol(1:space),FYsol(1:space),FZsol(1:space))
  ALLOCATE(FXdim(1:2),FYdim(1:2),FZdim(1:2),pFXsol(1:space),pFYsol(1:space),pFZsol(1:space))
  ALLOCATE(pFXdim(1:2),pFYdim(1:2),pFZdim(1:2))
  ALLOCATE(tFCellN(1:Vol),tGCellN(1:Vol),sFCellN(1:Vol,1:MaxF),sGCellN(1:Vol,1:MaxG))
  ALLOCATE(QXdim(1:2),QYdim(1:2),QZdim(1:2),affectedCellN(1:Vol))
  ALLOCATE(VXCM(1:Vol),VYCM(1:Vol),VZCM(1:Vol),tCellshift(1:Vol),NewCellN(1:space))
  ALLOCATE(UVX(1:Vol),UVY(1:Vol),UVZ(1:Vol),Potential(1:space))
  ALLOCATE(FpairX(1:space,1:2),FpairY(1:space,1:2),FpairZ(1:space,1:2))
  ALLOCATE(tICellN(1:Vol),sICellN(1:Vol,1:MaxI))
  ALLOCATE(newXsol(1:space),newYsol(1:space),newZsol(1:space),newFLAG(1:space))
  ALLOCATE(newVXsol(1:space),newVYsol(1:space),newVZsol(1:space))
  ALLOCATE(newFXsol(1:space),newFYsol(1:space),newFZsol(1:space))
  ALLOCATE(newpFXsol(1:space),newpFYsol(1:space),newpFZsol(1:space))

  FLAG = flg; Xsol = xs; Ysol = ys; Zsol = zs; VXsol = vxs; VYsol = vys; VZsol = vzs
  FXsol = fxs; FYsol = fys; FZsol = fzs; pFXsol = pfxs; pFYsol = pfys; pFZsol = pfzs
  Xdim = dm(:,1); Ydim = dm(:,2); Zdim = dm(:,3); VXdim = dm(:,4); VYdim = dm(:,5); VZdim = dm(:,6)
  FXdim = dm(:,7); FYdim = dm(:,8); FZdim = dm(:,9)
  call genrand_init(put=seed)

  cstep = cstep0
  do it = 1, nsteps
    if (MOD(cstep,MPCtstep) /= 0) then
      call Motion_SelkovDimer
      call ForcesMD_SelkovDimer
    else
      if (MOD(cstep,10000) == 0) call Sort
      call Motion_SelkovDimer
      call Distribute_SelkovDimer
      call ForcesMD_SelkovDimer
      call Affected_Cell
      call RMPC_SelkovDimer
    end if
    cstep = cstep + 1
  end do

  ip(1) = NT
  flg = FLAG; xs = Xsol; ys = Ysol; zs = Zsol; vxs = VXsol; vys = VYsol; vzs = VZsol
  fxs = FXsol; fys = FYsol; fzs = FZsol; pfxs = pFXsol; pfys = pFYsol; pfzs = pFZsol
  dm(:,1) = Xdim; dm(:,2) = Ydim; dm(:,3) = Zdim; dm(:,4) = VXdim; dm(:,5) = VYdim; dm(:,6) = VZdim
  dm(:,7) = FXdim; dm(:,8) = FYdim; dm(:,9) = FZdim
  rp(29) = TPotential

  DEALLOCATE(Xdim,Ydim,Zdim,prXdim,prYdim,prZdim,Xsol,Ysol,Zsol,FLAG,R,dR)
  DEALLOCATE(VXdim,VYdim,VZdim,prVXdim,prVYdim,prVZdim,VXsol,VYsol,VZsol,FXsol,FYsol,FZsol)
  DEALLOCATE(FXdim,FYdim,FZdim,pFXsol,pFYsol,pFZsol,pFXdim,pFYdim,pFZdim)
  DEALLOCATE(tFCellN,tGCellN,sFCellN,sGCellN,QXdim,QYdim,QZdim,affectedCellN)
  DEALLOCATE(VXCM,VYCM,VZCM,tCellshift,NewCellN,UVX,UVY,UVZ,Potential,FpairX,FpairY,FpairZ)
  DEALLOCATE(tICellN,sICellN,newXsol,newYsol,newZsol,newFLAG,newVXsol,newVYsol,newVZsol)
  DEALLOCATE(newFXsol,newFYsol,newFZsol,newpFXsol,newpFYsol,newpFZsol)
end subroutine

! The reference's own diagnostics on a given state: fills the globals that
! Output_SelkovDimer (Output_SelkovDimer.f90:1-305) and
! Output_Concentration_Catalytic (Output_Concentration_Catalytic.f90:1-191)
! read, sets cstep and Freq1..Freq6 from freq(7) = (cstep, Freq1..Freq6), and
! calls both in the current working directory, where they write their .dat
! records (DotVZ, Concentration, EnergyMomentum, ShellConcentration_Catalytic
! when cstep is a multiple of Freq3..Freq5; choose Freq6 not dividing cstep so
! the flow-field section stays off).  The census tF/tG/tI and the affected
! mask (Vol = 64000 entries) are taken as given.
!   rp(double): 1-3 box, 4 Mf, 5 Mg, 6 Mi, 7 Mc, 8 Mn, 9 Dcn, 10 TPotential
subroutine ref_sk_output(ip, rp, freq, flg, xs, ys, zs, vxs, vys, vzs, dm, tf, tg, ti, aff) &
    bind(C, name="ref_sk_output")
  integer(c_int64_t) :: ip(2)
  real(c_double) :: rp(10)
  integer(c_int64_t) :: freq(7)
  integer(c_int64_t) :: flg(ip(2))
  real(c_double), dimension(ip(2)) :: xs, ys, zs, vxs, vys, vzs
  real(c_double) :: dm(2,12)
  integer(c_int64_t), dimension(64000) :: tf, tg, ti, aff
  integer(c_int64_t) :: q

  NT = ip(1); space = ip(2)
  Box = rp(1:3); Mf = rp(4); Mg = rp(5); Mi = rp(6); Mc = rp(7); Mn = rp(8); Dcn = rp(9)
  TPotential = rp(10)
  do q = 1, 3
    HBox(q) = 0.5d0*Box(q)
    iBox(q) = INT(Box(q))
  end do
  cstep = freq(1)
  Freq1 = freq(2); Freq2 = freq(3); Freq3 = freq(4); Freq4 = freq(5); Freq5 = freq(6); Freq6 = freq(7)

  ALLOCATE(Xdim(1:2),Ydim(1:2),Zdim(1:2),VXdim(1:2),VYdim(1:2),VZdim(1:2),FXdim(1:2),FYdim(1:2),FZdim(1:2))
  ALLOCATE(Xsol(1:space),Ysol(1:space),Zsol(1:space),FLAG(1:space))
  ALLOCATE(VXsol(1:space),VYsol(1:space),VZsol(1:space),FXsol(1:space),FYsol(1:space),FZsol(1:space))
  ALLOCATE(PXsol(1:space),PYsol(1:space),PZsol(1:space))
  ALLOCATE(tFCellN(1:Vol),tGCellN(1:Vol),tICellN(1:Vol),affectedCellN(1:Vol))
  ALLOCATE(SolCenterX(1:space),SolCenterY(1:space),SolCenterZ(1:space))
  ALLOCATE(Rx(1:space),Ry(1:space),Rz(1:space),R2(1:space))

  FLAG = flg; Xsol = xs; Ysol = ys; Zsol = zs; VXsol = vxs; VYsol = vys; VZsol = vzs
  FXsol = 0.d0; FYsol = 0.d0; FZsol = 0.d0
  PXsol = 0.d0; PYsol = 0.d0; PZsol = 0.d0
  Xdim = dm(:,1); Ydim = dm(:,2); Zdim = dm(:,3); VXdim = dm(:,4); VYdim = dm(:,5); VZdim = dm(:,6)
  FXdim = dm(:,7); FYdim = dm(:,8); FZdim = dm(:,9)
  tFCellN = tf; tGCellN = tg; tICellN = ti; affectedCellN = aff

  call Output_SelkovDimer
  call Output_Concentration_Catalytic

  DEALLOCATE(Xdim,Ydim,Zdim,VXdim,VYdim,VZdim,FXdim,FYdim,FZdim,Xsol,Ysol,Zsol,FLAG)
  DEALLOCATE(VXsol,VYsol,VZsol,FXsol,FYsol,FZsol,PXsol,PYsol,PZsol)
  DEALLOCATE(tFCellN,tGCellN,tICellN,affectedCellN,SolCenterX,SolCenterY,SolCenterZ,Rx,Ry,Rz,R2)
end subroutine

! The fixed-form records of Output_SelkovDimer.f90 (FORMAT 15 :26, 16 :139,
! 18 :186, 19 :238, 17 :305) and Main_SelkovDimer.f90 (23 :158), written with
! the reference's FORMAT statements into the file at p: one record of each.
subroutine ref_sk_records(p, n, cstep, v) bind(C, name="ref_sk_records")
  integer(c_int32_t), value :: n, cstep
  character(kind=c_char) :: p(n)
  real(c_double) :: v(10)
  character(len=n) :: fname
  integer :: i
  do i = 1, n
    fname(i:i) = p(i)
  end do
  OPEN(UNIT=10, FILE=fname, STATUS='replace', ACTION='WRITE')
  WRITE(10,15) cstep, v(1), v(2), v(3)
  WRITE(10,16) v(1), v(2), v(3), cstep
  WRITE(10,18) cstep, v(1), v(2), v(3), v(4)
  WRITE(10,19) cstep, v(1), v(2), v(3), v(4), v(5), v(6), v(7), v(8), v(9), v(10)
  WRITE(10,17) cstep, v(1), v(2), v(3), v(4), v(5), v(6), v(7), v(8)
  WRITE(10,23) cstep
  CLOSE(10)
15 FORMAT(I8,1X,3(E23.16,1X))
16 FORMAT(3(ES23.16,1X),I8)
17 FORMAT(I8,8(1X,ES23.16))
18 FORMAT(I8,4(1X,ES23.16))
19 FORMAT(I8,10(1X,ES23.16))
23 FORMAT(I9)
end subroutine

end module ref_selkov_harness
