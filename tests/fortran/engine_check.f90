! Runs the device-resident engines from Fortran through the bind(C) mirrors
! and interfaces of main_amd/fortran/mpcx_c.f90, as the engine loops of the
! integration guide do: `engine_check co` replaces main_pattern.f95:202-225
! (collective_code), `engine_check sk` Main_SelkovDimer.f90:125-149
! (selkov_code), each in chunks of collisions with the reference's
! Mersenne-Twister draws taken from the caller's stream.  Inputs come from
! co_in.bin / sk_in.bin (the configuration struct as its C bytes, then the
! state), outputs go to co_out.bin / sk_out.bin (stream files);
! tests/test_gpu_fortran.py writes the one and compares the other with the oracle.
program engine_check
  use iso_c_binding
  use mpcx_c
  implicit none
  character(len=8) :: mode
  call get_command_argument(1, mode)
  if (trim(mode) == "co") then
    call run_co()
  else if (trim(mode) == "sk") then
    call run_sk()
  else
    error stop "usage: engine_check co|sk"
  end if
  call mpcx_check(mpcx_shutdown(), "shutdown")

contains

  subroutine run_co()
    type(mpcx_co_config) :: cfg
    type(c_ptr) :: eng
    integer(c_int64_t) :: NT, nd, nrb, vol, nchunk, ncoll_chunk, ndraw, i
    real(c_double), allocatable :: x(:), y(:), z(:), vx(:), vy(:), vz(:), draws(:), dimer12(:, :)
    real(c_double), allocatable :: xd(:, :), yd(:, :), zd(:, :), vxd(:, :), vyd(:, :), vzd(:, :)
    real(c_double), allocatable :: fxd(:, :), fyd(:, :), fzd(:, :)
    integer(c_int32_t), allocatable :: flag(:), react_box(:)
    integer(c_int64_t), allocatable :: tour(:)
    integer(c_int32_t) :: mt(624), mti, scal(4)
    integer :: u

    open(newunit=u, file="co_in.bin", access="stream", form="unformatted", status="old")
    read(u) cfg
    read(u) nrb, nchunk, ncoll_chunk
    NT = cfg%NT
    nd = cfg%numdims
    vol = cfg%vol
    allocate(x(NT), y(NT), z(NT), vx(NT), vy(NT), vz(NT), flag(NT), react_box(nrb), tour(vol))
    allocate(xd(nd, 2), yd(nd, 2), zd(nd, 2), vxd(nd, 2), vyd(nd, 2), vzd(nd, 2), fxd(nd, 2), fyd(nd, 2), fzd(nd, 2))
    read(u) x, y, z, vx, vy, vz, flag, xd, yd, zd, vxd, vyd, vzd, fxd, fyd, fzd, react_box, tour, mt, mti
    close(u)
    allocate(dimer12(2 * nd, 12))

    call mpcx_check(mpcx_co_create(eng, cfg), "mpcx_co_create")
    call mpcx_check(mpcx_co_set_state(eng, 0_c_int32_t, x, y, z, vx, vy, vz, flag, xd, yd, zd, vxd, vyd, vzd, &
                                      fxd, fyd, fzd, react_box, tour), "mpcx_co_set_state")
    call mpcx_check(mpcx_co_setup(eng, 1_c_int32_t), "mpcx_co_setup")     ! main_pattern.f95:170-188
    ndraw = (3 + 2 * cfg%mpcnumcell + cfg%rxnnumcell) * ncoll_chunk       ! grnd() calls of the chunk
    allocate(draws(ndraw))
    do i = 1, nchunk
      call mpcx_check(mpcx_mtmod_fill(mt, mti, draws, ndraw), "mpcx_mtmod_fill")
      call mpcx_check(mpcx_co_run(eng, ncoll_chunk, draws), "mpcx_co_run")
      call mpcx_check(mpcx_co_get_state(eng, x, y, z, vx, vy, vz, flag, dimer12, scal), "mpcx_co_get_state")
    end do
    call mpcx_check(mpcx_co_destroy(eng), "mpcx_co_destroy")

    open(newunit=u, file="co_out.bin", access="stream", form="unformatted", status="replace")
    write(u) x, y, z, vx, vy, vz, flag, dimer12, scal, mt, mti
    close(u)
  end subroutine

  subroutine run_sk()
    type(mpcx_sk_config) :: cfg
    type(c_ptr) :: eng
    integer(c_int64_t) :: space, nchunk, ncoll_chunk, vol, ub, used, i, near
    integer(c_int64_t) :: scal(4)
    integer(c_int32_t) :: seed, val(624), cnt, val0(624), cnt0
    real(c_double), allocatable :: x(:), y(:), z(:), vx(:), vy(:), vz(:), fx(:), fy(:), fz(:), draws(:)
    integer(c_int64_t), allocatable :: flag(:)
    real(c_double) :: dm(2, 12), tpot
    integer :: u

    open(newunit=u, file="sk_in.bin", access="stream", form="unformatted", status="old")
    read(u) cfg
    read(u) nchunk, ncoll_chunk, seed
    space = cfg%space
    allocate(x(space), y(space), z(space), vx(space), vy(space), vz(space), fx(space), fy(space), fz(space))
    allocate(flag(space))
    read(u) x, y, z, vx, vy, vz, fx, fy, fz, flag, dm
    close(u)
    vol = nint(cfg%box(1)) * nint(cfg%box(2)) * nint(cfg%box(3))

    call mpcx_check(mpcx_sk_create(eng, cfg), "mpcx_sk_create")
    call mpcx_check(mpcx_sk_set_state(eng, 0_c_int64_t, cfg%NT, x, y, z, vx, vy, vz, fx, fy, fz, flag, dm), &
                    "mpcx_sk_set_state")
    call mpcx_check(mpcx_mt95_seed(seed, val, cnt), "mpcx_mt95_seed")
    do i = 1, nchunk
      call mpcx_check(mpcx_sk_get_state(eng, x, y, z, vx, vy, vz, fx, fy, fz, flag, dm, scal, tpot), "get")
      ! a generous buffer of the interval's draws (its conversion draws depend
      ! on the catalytic neighbourhood); the stream then advances by `used`
      near = max(scal(4), 1024_c_int64_t)
      ub = ncoll_chunk * (3 + 4 * vol + int(cfg%MPCtstep, c_int64_t) * 2 * near)
      if (allocated(draws)) deallocate(draws)
      allocate(draws(ub))
      val0 = val
      cnt0 = cnt
      call mpcx_check(mpcx_mt95_fill(val, cnt, draws, ub), "mpcx_mt95_fill")
      call mpcx_check(mpcx_sk_run(eng, ncoll_chunk, draws, ub, used), "mpcx_sk_run")
      val = val0
      cnt = cnt0
      call mpcx_check(mpcx_mt95_fill(val, cnt, draws, used), "mpcx_mt95_fill")
    end do
    call mpcx_check(mpcx_sk_get_state(eng, x, y, z, vx, vy, vz, fx, fy, fz, flag, dm, scal, tpot), "get")
    call mpcx_check(mpcx_sk_destroy(eng), "mpcx_sk_destroy")

    open(newunit=u, file="sk_out.bin", access="stream", form="unformatted", status="replace")
    write(u) x, y, z, vx, vy, vz, fx, fy, fz, flag, dm, scal, tpot, val, cnt
    close(u)
  end subroutine
end program
