! ISO_C_BINDING interfaces of libmpcx (include/mpcx.h) for Fortran callers.
! Integers map to c_int32_t/c_int64_t, reals to c_double; arrays are passed
! by address (explicit-size dummies), scalars by value as in the C prototypes.
! The uint32 Mersenne-Twister words travel as default (32-bit) integers, the
! type the collective_code driver keeps them in.
module mpcx_c
  use iso_c_binding
  implicit none
  private :: c_strlen

  ! bind(C) mirrors of the engines' configuration structs, component for
  ! component in the order and kinds of include/mpcx.h (neither has padding)
  type, bind(C) :: mpcx_co_config
    integer(c_int64_t) :: NT, numdims, vol, mpcnumcell, rxnnumcell, max_neighbour
    integer(c_int32_t) :: MPCtstep, sort_step, react_dir, rng_mode
    integer(c_int64_t) :: seed                  ! uint64 Philox key
    real(c_double) :: box(3)
    real(c_double) :: MDt, Ma, Mb, Ms, Mc, Mn, Dcn, cutoff, Rc, Rn
    real(c_double) :: EaC, EbC, EsC, EaN, EbN, EsN, Rw, Ew, Rm_delta, Em, RPab, RPba
    real(c_double) :: skinrad, rotang, a0, MPCt, k1
    real(c_double) :: sinrot, cosrot, wall_w6
    real(c_double) :: reaction_p(0:100)
    real(c_double) :: sphere_guard
  end type

  type, bind(C) :: mpcx_sk_config
    integer(c_int64_t) :: NT, space
    integer(c_int32_t) :: MPCtstep, sort_step, rng_mode, pad
    integer(c_int64_t) :: seed
    real(c_double) :: box(3)
    real(c_double) :: MDt, MPCt, Mf, Mg, Mi, Mc, Mn, Dcn, Rc, Rn
    real(c_double) :: EfC, EgC, EiC, EfN, EgN, EiN, RPfg, RPgf, A0
    real(c_double) :: k(6)
    real(c_double) :: sinrot, cosrot, sphere_guard
  end type

  interface
    ! ---------------------------------------------------------------- engines
    function mpcx_co_create(eng, cfg) bind(C, name="mpcx_co_create") result(rc)
      import :: c_int, c_ptr, mpcx_co_config
      type(c_ptr) :: eng
      type(mpcx_co_config) :: cfg
      integer(c_int) :: rc
    end function

    function mpcx_co_destroy(eng) bind(C, name="mpcx_co_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: eng
      integer(c_int) :: rc
    end function

    function mpcx_co_set_state(eng, cstep, x, y, z, vx, vy, vz, flag, xd, yd, zd, vxd, vyd, vzd, fxd, fyd, fzd, &
                               react_box, tour) bind(C, name="mpcx_co_set_state") result(rc)
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: eng
      integer(c_int32_t), value :: cstep
      real(c_double) :: x(*), y(*), z(*), vx(*), vy(*), vz(*)
      integer(c_int32_t) :: flag(*)
      real(c_double) :: xd(*), yd(*), zd(*), vxd(*), vyd(*), vzd(*), fxd(*), fyd(*), fzd(*)
      integer(c_int32_t) :: react_box(*)
      integer(c_int64_t) :: tour(*)
      integer(c_int) :: rc
    end function

    function mpcx_co_setup(eng, do_sort) bind(C, name="mpcx_co_setup") result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: eng
      integer(c_int32_t), value :: do_sort
      integer(c_int) :: rc
    end function

    ! draws: (3 + 2*mpcnumcell + rxnnumcell) per collision from the caller's
    ! grnd() stream in rng_mode 0; ignored (any one-element array) in rng_mode 1
    function mpcx_co_run(eng, ncoll, draws) bind(C, name="mpcx_co_run") result(rc)
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: eng
      integer(c_int64_t), value :: ncoll
      real(c_double) :: draws(*)
      integer(c_int) :: rc
    end function

    function mpcx_co_init_solvent(eng, Na0, Nb0, T, seed) bind(C, name="mpcx_co_init_solvent") result(rc)
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: eng
      integer(c_int64_t), value :: Na0, Nb0
      real(c_double), value :: T
      integer(c_int64_t), value :: seed
      integer(c_int) :: rc
    end function

    function mpcx_co_get_state(eng, x, y, z, vx, vy, vz, flag, dimer12, scalars) &
        bind(C, name="mpcx_co_get_state") result(rc)
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: eng
      real(c_double) :: x(*), y(*), z(*), vx(*), vy(*), vz(*)
      integer(c_int32_t) :: flag(*)
      real(c_double) :: dimer12(*)        ! (2*numdims, 12): x y z vx vy vz fx fy fz qx qy qz
      integer(c_int32_t) :: scalars(4)    ! cstep, time_count, time_penetrate, inside_num
      integer(c_int) :: rc
    end function

    function mpcx_co_set_near_mode(eng, mode) bind(C, name="mpcx_co_set_near_mode") result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: eng
      integer(c_int32_t), value :: mode
      integer(c_int) :: rc
    end function

    function mpcx_co_set_overlap(eng, on) bind(C, name="mpcx_co_set_overlap") result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: eng
      integer(c_int32_t), value :: on
      integer(c_int) :: rc
    end function

    function mpcx_co_set_precision(eng, split) bind(C, name="mpcx_co_set_precision") result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: eng
      integer(c_int32_t), value :: split
      integer(c_int) :: rc
    end function

    function mpcx_co_observables(eng, out16) bind(C, name="mpcx_co_observables") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: eng
      real(c_double) :: out16(16)
      integer(c_int) :: rc
    end function

    function mpcx_sk_create(eng, cfg) bind(C, name="mpcx_sk_create") result(rc)
      import :: c_int, c_ptr, mpcx_sk_config
      type(c_ptr) :: eng
      type(mpcx_sk_config) :: cfg
      integer(c_int) :: rc
    end function

    function mpcx_sk_destroy(eng) bind(C, name="mpcx_sk_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: eng
      integer(c_int) :: rc
    end function

    function mpcx_sk_set_state(eng, cstep, NT, x, y, z, vx, vy, vz, fx, fy, fz, flag, dm) &
        bind(C, name="mpcx_sk_set_state") result(rc)
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: eng
      integer(c_int64_t), value :: cstep, NT
      real(c_double) :: x(*), y(*), z(*), vx(*), vy(*), vz(*), fx(*), fy(*), fz(*)
      integer(c_int64_t) :: flag(*)
      real(c_double) :: dm(2, 12)
      integer(c_int) :: rc
    end function

    function mpcx_sk_run(eng, ncoll, draws, ndraws, used) bind(C, name="mpcx_sk_run") result(rc)
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: eng
      integer(c_int64_t), value :: ncoll
      real(c_double) :: draws(*)
      integer(c_int64_t), value :: ndraws
      integer(c_int64_t) :: used
      integer(c_int) :: rc
    end function

    function mpcx_sk_get_state(eng, x, y, z, vx, vy, vz, fx, fy, fz, flag, dm, scalars, tpotential) &
        bind(C, name="mpcx_sk_get_state") result(rc)
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: eng
      real(c_double) :: x(*), y(*), z(*), vx(*), vy(*), vz(*), fx(*), fy(*), fz(*)
      integer(c_int64_t) :: flag(*)
      real(c_double) :: dm(2, 12)
      integer(c_int64_t) :: scalars(4)   ! cstep, NT, draws used, near-set size
      real(c_double) :: tpotential
      integer(c_int) :: rc
    end function

    function mpcx_sk_observables(eng, out16) bind(C, name="mpcx_sk_observables") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: eng
      real(c_double) :: out16(16)
      integer(c_int) :: rc
    end function

    function mpcx_mt95_seed(seed, val, cnt) bind(C, name="mpcx_mt95_seed") result(rc)
      import :: c_int, c_int32_t
      integer(c_int32_t), value :: seed
      integer(c_int32_t) :: val(624), cnt
      integer(c_int) :: rc
    end function

    function mpcx_mt95_fill(val, cnt, out, n) bind(C, name="mpcx_mt95_fill") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t) :: val(624), cnt
      real(c_double) :: out(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function

    ! batches of replicas (engines made by mpcx_co_create)
    function mpcx_cob_create(b, R, engines, tile) bind(C, name="mpcx_cob_create") result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr) :: b
      integer(c_int32_t), value :: R
      type(c_ptr) :: engines(*)
      integer(c_int32_t), value :: tile
      integer(c_int) :: rc
    end function

    function mpcx_cob_destroy(b) bind(C, name="mpcx_cob_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int) :: rc
    end function

    function mpcx_cob_run(b, ncoll, use_graph) bind(C, name="mpcx_cob_run") result(rc)
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: b
      integer(c_int64_t), value :: ncoll
      integer(c_int32_t), value :: use_graph
      integer(c_int) :: rc
    end function

    function mpcx_cob_set_tile(b, tile) bind(C, name="mpcx_cob_set_tile") result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: b
      integer(c_int32_t), value :: tile
      integer(c_int) :: rc
    end function

    function mpcx_cob_set_bulk_slots(b, slots) bind(C, name="mpcx_cob_set_bulk_slots") result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: b
      integer(c_int32_t), value :: slots
      integer(c_int) :: rc
    end function

    function mpcx_cob_set_sort_step(b, steps) bind(C, name="mpcx_cob_set_sort_step") result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: b
      integer(c_int32_t), value :: steps
      integer(c_int) :: rc
    end function

    function mpcx_cob_set_inbox_cap(b, cap) bind(C, name="mpcx_cob_set_inbox_cap") result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: b
      integer(c_int32_t), value :: cap
      integer(c_int) :: rc
    end function

    function mpcx_shutdown() bind(C, name="mpcx_shutdown") result(rc)
      import :: c_int
      integer(c_int) :: rc
    end function

    ! ------------------------------------------------------ drop-ins and I/O
    function mpcx_last_error() bind(C, name="mpcx_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function

    function c_strlen(p) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: p
      integer(c_size_t) :: n
    end function

    function mpcx_device_count() bind(C, name="mpcx_device_count") result(n)
      import :: c_int
      integer(c_int) :: n
    end function

    function mpcx_mtmod_fill(mt, mti, out, n) bind(C, name="mpcx_mtmod_fill") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t) :: mt(*)
      integer(c_int32_t) :: mti
      real(c_double) :: out(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function

    function mpcx_mpc_rotation(NT, rotang, numcell, a0, box, x, y, z, vx, vy, vz, vxcm, vycm, vzcm, tcs, ncn, &
                               uvx, uvy, uvz, draws) bind(C, name="mpcx_mpc_rotation") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: NT
      real(c_double), value :: rotang
      integer(c_int64_t), value :: numcell
      real(c_double), value :: a0
      real(c_double) :: box(*), x(*), y(*), z(*), vx(*), vy(*), vz(*)
      real(c_double) :: vxcm(*), vycm(*), vzcm(*), tcs(*), uvx(*), uvy(*), uvz(*), draws(*)
      integer(c_int64_t) :: ncn(*)
      integer(c_int) :: rc
    end function

    function mpcx_reaction_simple(MPCt, numcell, k1, NT, flag, tB, sB, vol, draws) &
        bind(C, name="mpcx_reaction_simple") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double
      real(c_double), value :: MPCt
      integer(c_int64_t), value :: numcell
      real(c_double), value :: k1
      integer(c_int32_t), value :: NT
      integer(c_int32_t) :: flag(*)
      integer(c_int64_t) :: tB(*), sB(*)
      integer(c_int64_t), value :: vol
      real(c_double) :: draws(*)
      integer(c_int) :: rc
    end function

    function mpcx_neighbour_list(NT, box, cutoff, skinrad, x, y, z, xd, yd, zd, num, nb, stride, max_nb) &
        bind(C, name="mpcx_neighbour_list") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: NT
      real(c_double) :: box(*)
      real(c_double), value :: cutoff, skinrad
      real(c_double) :: x(*), y(*), z(*)
      real(c_double), value :: xd, yd, zd
      integer(c_int64_t) :: num
      integer(c_int64_t) :: nb(*)
      integer(c_int64_t), value :: stride, max_nb
      integer(c_int) :: rc
    end function

    function mpcx_inout_particle_list(NT, box, cutC, cutN, skinrad, x, y, z, xd, yd, zd, out_num, out_list, &
                                      in_num, in_list, in_cap, numdims) &
        bind(C, name="mpcx_inout_particle_list") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: NT
      real(c_double) :: box(*)
      real(c_double), value :: cutC, cutN, skinrad
      real(c_double) :: x(*), y(*), z(*), xd(*), yd(*), zd(*)
      integer(c_int64_t) :: out_num, in_num
      integer(c_int64_t) :: out_list(*), in_list(*)
      integer(c_int64_t), value :: in_cap
      integer(c_int32_t), value :: numdims
      integer(c_int) :: rc
    end function

    function mpcx_distribute(NT, numdims, flag, x, y, z, xd, yd, zd, box, vol, tA, tB, tS, sA, sB, sS, maxs, &
                             cutC2, cutN2, out_num, out_list, in_num, in_list) &
        bind(C, name="mpcx_distribute") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: NT, numdims
      integer(c_int32_t) :: flag(*)
      real(c_double) :: x(*), y(*), z(*), xd(*), yd(*), zd(*), box(*)
      integer(c_int64_t), value :: vol
      integer(c_int64_t) :: tA(*), tB(*), tS(*), sA(*), sB(*), sS(*)
      integer(c_int64_t), value :: maxs
      real(c_double), value :: cutC2, cutN2
      integer(c_int64_t), value :: out_num
      integer(c_int64_t) :: out_list(*)
      integer(c_int64_t), value :: in_num
      integer(c_int64_t) :: in_list(*)
      integer(c_int) :: rc
    end function

    function mpcx_sort_hilbert(vol, NT, box, tour, x, y, z, flag, vx, vy, vz) &
        bind(C, name="mpcx_sort_hilbert") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int64_t), value :: vol
      integer(c_int32_t), value :: NT
      real(c_double) :: box(*)
      integer(c_int64_t) :: tour(*)
      real(c_double) :: x(*), y(*), z(*), vx(*), vy(*), vz(*)
      integer(c_int32_t) :: flag(*)
      integer(c_int) :: rc
    end function

    ! restart files (main_pattern.f95:128-146,243-251; mtmod.f90:152-229;
    ! initial_mod.f95:118; restart_mod.f95:410-414); paths NUL-terminated
    function mpcx_timestep_write(path, cstep, n_a, n_b, n_s) bind(C, name="mpcx_timestep_write") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_char
      character(kind=c_char) :: path(*)
      integer(c_int32_t), value :: cstep
      integer(c_int64_t), value :: n_a, n_b, n_s
      integer(c_int) :: rc
    end function

    function mpcx_timestep_read(path, cstep, n_a, n_b, n_s) bind(C, name="mpcx_timestep_read") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_char
      character(kind=c_char) :: path(*)
      integer(c_int32_t) :: cstep
      integer(c_int64_t) :: n_a, n_b, n_s
      integer(c_int) :: rc
    end function

    function mpcx_mt_state_write(path, mt, mti, nstreams) bind(C, name="mpcx_mt_state_write") result(rc)
      import :: c_int, c_int32_t, c_char
      character(kind=c_char) :: path(*)
      integer(c_int32_t) :: mt(*), mti(*)
      integer(c_int32_t), value :: nstreams
      integer(c_int) :: rc
    end function

    function mpcx_mt_state_read(path, mt, mti, max_streams, nstreams) bind(C, name="mpcx_mt_state_read") result(rc)
      import :: c_int, c_int32_t, c_char
      character(kind=c_char) :: path(*)
      integer(c_int32_t) :: mt(*), mti(*)
      integer(c_int32_t), value :: max_streams
      integer(c_int32_t) :: nstreams
      integer(c_int) :: rc
    end function

    function mpcx_motor_positions_write(path, cm, nd) bind(C, name="mpcx_motor_positions_write") result(rc)
      import :: c_int, c_int32_t, c_double, c_char
      character(kind=c_char) :: path(*)
      real(c_double) :: cm(*)
      integer(c_int32_t), value :: nd
      integer(c_int) :: rc
    end function

    function mpcx_motor_positions_read(path, cm, nd) bind(C, name="mpcx_motor_positions_read") result(rc)
      import :: c_int, c_int32_t, c_double, c_char
      character(kind=c_char) :: path(*)
      real(c_double) :: cm(*)
      integer(c_int32_t), value :: nd
      integer(c_int) :: rc
    end function
  end interface

contains

  ! Stops the program with libmpcx's message when a call failed (the C ABI
  ! returns 0 on success and a negative code otherwise).
  subroutine mpcx_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    character(len=:), allocatable :: msg
    integer :: i, n
    if (rc == 0) return
    p = mpcx_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = s(i)
    end do
    write(0, '(a)') what // ': ' // msg
    error stop 1
  end subroutine
end module mpcx_c
