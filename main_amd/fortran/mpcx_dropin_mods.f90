! Drop-in replacements of collective_code's module procedures.  Each module
! and procedure has the name, dummy arguments, kinds and intents of the one it
! replaces, so a driver such as main_pattern.f95 compiles unchanged against
! these modules instead of the reference's; the work runs on the GPU through
! libmpcx.  The Mersenne-Twister draws are taken from the caller's mt(:,1),
! mti(1) in the reference's one-thread call order (mpcx_mtmod_fill is the
! sgrnd/grnd stream of mtmod.f90), so mt_state.dat ends as it would.

! collective_code/mpc_rotation_mod.f95:6-36
module mpc_rotation_mod
  use mpcx_c
  implicit none
contains
  subroutine mpc_rotation(NT, rotang, numcell, a0, Box, Xsol, Ysol, Zsol, VXsol, VYsol, VZsol, VXCM, VYCM, VZCM, &
                          tCellShift, NewCellN, UVX, UVY, UVZ, mt, mti)
    integer, parameter :: sp = selected_real_kind(6, 37)
    integer, parameter :: dp = selected_real_kind(15, 307)
    integer(kind=sp), intent(in) :: NT
    real(kind=dp), intent(in) :: rotang
    integer(kind=dp), intent(in) :: numcell
    real(kind=dp), intent(in) :: a0
    real(kind=dp), dimension(:), intent(in) :: Box, Xsol, Ysol, Zsol
    real(kind=dp), dimension(:), intent(inout) :: VXsol, VYsol, VZsol, VXCM, VYCM, VZCM, tCellShift
    integer(kind=dp), dimension(:), intent(inout) :: NewCellN
    real(kind=dp), dimension(:), intent(inout) :: UVX, UVY, UVZ
    integer, dimension(:, :), intent(inout) :: mt
    integer, dimension(:), intent(inout) :: mti
    real(kind=dp), allocatable :: draws(:)
    allocate(draws(3 + 2 * numcell))
    ! grnd() at mpc_rotation_mod.f95:102-106 (shift) and :187-189 (two per cell)
    call mpcx_check(mpcx_mtmod_fill(mt(:, 1), mti(1), draws, int(size(draws), c_int64_t)), 'mpc_rotation')
    call mpcx_check(mpcx_mpc_rotation(NT, rotang, numcell, a0, Box, Xsol, Ysol, Zsol, VXsol, VYsol, VZsol, &
                                      VXCM, VYCM, VZCM, tCellShift, NewCellN, UVX, UVY, UVZ, draws), 'mpc_rotation')
  end subroutine
end module mpc_rotation_mod

! collective_code/reaction_mod.f95:8-27
module reaction_mod
  use mpcx_c
  implicit none
contains
  subroutine reaction_simple(MPCt, numcell, k1, flag, tBCellN, sBCellN, mt, mti)
    integer, parameter :: sp = selected_real_kind(6, 37)
    integer, parameter :: dp = selected_real_kind(15, 307)
    integer(kind=dp), intent(in) :: numcell
    real(kind=dp), intent(in) :: MPCt, k1
    integer(kind=sp), dimension(:), intent(inout) :: flag
    integer(kind=dp), dimension(:), intent(in) :: tBCellN
    integer(kind=dp), dimension(:, :), intent(in) :: sBCellN
    integer, dimension(:, :), intent(inout) :: mt
    integer, dimension(:), intent(inout) :: mti
    real(kind=dp), allocatable :: draws(:)
    allocate(draws(max(numcell, 1_dp)))
    ! one grnd() per cell, reaction_mod.f95:57
    if (numcell > 0) &
      call mpcx_check(mpcx_mtmod_fill(mt(:, 1), mti(1), draws, int(numcell, c_int64_t)), 'reaction_simple')
    call mpcx_check(mpcx_reaction_simple(MPCt, numcell, k1, int(size(flag), c_int32_t), flag, tBCellN, sBCellN, &
                                         int(size(sBCellN, 1), c_int64_t), draws), 'reaction_simple')
  end subroutine
end module reaction_mod

! collective_code/neighbour_list_mod.f95:6-26 and :63-87
module neighbour_list_mod
  use mpcx_c
  implicit none
contains
  subroutine neighbour_list(NT, Box, cutoff, skinrad, Xsol, Ysol, Zsol, Xdim, Ydim, Zdim, neighbour_num, neighbour)
    integer, parameter :: sp = selected_real_kind(6, 37)
    integer, parameter :: dp = selected_real_kind(15, 307)
    integer(kind=sp), intent(in) :: NT
    real(kind=dp), dimension(:), intent(in) :: Box
    real(kind=dp), intent(in) :: cutoff, skinrad
    real(kind=dp), dimension(:), intent(in) :: Xsol, Ysol, Zsol
    real(kind=dp), intent(in) :: Xdim, Ydim, Zdim
    integer(kind=dp), intent(out) :: neighbour_num
    integer(kind=dp), dimension(:), intent(out) :: neighbour
    ! the caller zeroes the count first (evolve_mod.f95:536); the list is
    ! appended from that count on, as the reference's loop does
    call mpcx_check(mpcx_neighbour_list(NT, Box, cutoff, skinrad, Xsol, Ysol, Zsol, Xdim, Ydim, Zdim, &
                                        neighbour_num, neighbour, 1_c_int64_t, int(size(neighbour), c_int64_t)), &
                    'neighbour_list')
  end subroutine

  subroutine inout_particle_list(NT, box, cutoffC, cutoffN, skinrad, xsol, ysol, zsol, xdim, ydim, zdim, &
                                 outside_particle_num, outside_particle, inside_particle_num, inside_particle, numdims)
    integer, parameter :: sp = selected_real_kind(6, 37)
    integer, parameter :: dp = selected_real_kind(15, 307)
    integer(kind=sp), intent(in) :: NT
    real(kind=dp), dimension(:), intent(in) :: box
    real(kind=dp), intent(in) :: cutoffC, cutoffN, skinrad
    real(kind=dp), dimension(:), intent(in) :: xsol, ysol, zsol
    real(kind=dp), dimension(:, :), intent(in) :: xdim, ydim, zdim
    integer(kind=dp), intent(inout) :: outside_particle_num
    integer(kind=dp), dimension(:), intent(inout) :: outside_particle
    integer(kind=dp), intent(inout) :: inside_particle_num
    integer(kind=dp), dimension(:), intent(inout) :: inside_particle
    integer, intent(in) :: numdims
    call mpcx_check(mpcx_inout_particle_list(NT, box, cutoffC, cutoffN, skinrad, xsol, ysol, zsol, xdim, ydim, zdim, &
                                             outside_particle_num, outside_particle, inside_particle_num, &
                                             inside_particle, int(size(inside_particle), c_int64_t), &
                                             int(numdims, c_int32_t)), 'inout_particle_list')
  end subroutine
end module neighbour_list_mod

! collective_code/distribute_mod.f95:6-40
module distribute_mod
  use mpcx_c
  implicit none
contains
  subroutine distribute(NT, numdims, cutoff, Rc, Rn, flag, Xsol, Ysol, Zsol, Xdim, Ydim, Zdim, Box, tACellN, &
                        tBCellN, tSCellN, sACellN, sBCellN, sSCellN, cutoffC2, cutoffN2, cstep, &
                        outside_particle_num, outside_particle, inside_particle_num, inside_particle)
    integer, parameter :: sp = selected_real_kind(6, 37)
    integer, parameter :: dp = selected_real_kind(15, 307)
    integer(kind=sp), intent(in) :: NT
    integer, intent(in) :: numdims
    integer(kind=sp), dimension(:), intent(in) :: flag
    real(kind=dp), dimension(:), intent(in) :: Xsol, Ysol, Zsol
    real(kind=dp), dimension(:, :), intent(in) :: Xdim, Ydim, Zdim
    real(kind=dp), dimension(:), intent(in) :: Box
    integer(kind=dp), dimension(:), intent(out) :: tACellN, tBCellN, tSCellN
    integer(kind=dp), dimension(:, :), intent(out) :: sACellN, sBCellN, sSCellN
    real(kind=dp), intent(in) :: cutoff, Rc, Rn, cutoffC2, cutoffN2
    integer :: cstep
    integer(kind=dp), intent(inout) :: outside_particle_num
    integer(kind=dp), dimension(:), intent(inout) :: outside_particle
    integer(kind=dp), intent(inout) :: inside_particle_num
    integer(kind=dp), dimension(:), intent(inout) :: inside_particle
    call mpcx_check(mpcx_distribute(NT, int(numdims, c_int32_t), flag, Xsol, Ysol, Zsol, Xdim, Ydim, Zdim, Box, &
                                    int(size(tACellN), c_int64_t), tACellN, tBCellN, tSCellN, sACellN, sBCellN, &
                                    sSCellN, int(min(size(sACellN, 2), size(sBCellN, 2), size(sSCellN, 2)), c_int64_t), &
                                    cutoffC2, cutoffN2, outside_particle_num, outside_particle, inside_particle_num, &
                                    inside_particle), 'distribute')
  end subroutine
end module distribute_mod

! collective_code/sort_hilbert_mod.f95:6-24; the tour is read once from
! hilbert.txt in the working directory, in the same "x y z" per-line format
! (sort_hilbert_mod.f95:57-62), and kept for later calls
module sort_hilbert_mod
  use mpcx_c
  implicit none
  integer(c_int64_t), allocatable, private, save :: tour(:)
contains
  subroutine sort_hilbert(vol, NT, box, xsol, ysol, zsol, flag, vxsol, vysol, vzsol)
    integer, parameter :: sp = selected_real_kind(6, 37)
    integer, parameter :: dp = selected_real_kind(15, 307)
    integer(kind=dp), intent(in) :: vol
    integer(kind=sp), intent(in) :: NT
    real(kind=dp), dimension(:), intent(in) :: box
    real(kind=dp), dimension(:), intent(inout) :: xsol, ysol, zsol
    integer, dimension(:), intent(inout) :: flag
    real(kind=dp), dimension(:), intent(inout) :: vxsol, vysol, vzsol
    integer :: i, u, c(3)
    if (.not. allocated(tour)) then
      allocate(tour(vol))
      open(newunit=u, file="hilbert.txt", status="old", action="read")
      do i = 1, int(vol)
        read(u, *) c(1), c(2), c(3)
        tour(i) = int(c(1) + box(1) * c(2) + box(1) * box(2) * c(3) + 1, c_int64_t)
      end do
      close(u)
    end if
    call mpcx_check(mpcx_sort_hilbert(vol, NT, box, tour, xsol, ysol, zsol, flag, vxsol, vysol, vzsol), &
                    'sort_hilbert')
  end subroutine
end module sort_hilbert_mod
