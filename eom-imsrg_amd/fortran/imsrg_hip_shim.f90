! Drop-in shim for n-parzuchowski/EOM-IMSRG: the reference's commutator,
! BCH / Magnus and flow routines with their own argument lists
! (commutators.f90:10-925, cross_coupled.f90:1190-1300,
! TS_commutators.f90:10-1849, IMSRG_MAGNUS.f90:131-146,389-464,
! IMSRG_MAGNUS.f90:15-234,620-691,754-764, generators.f90:4-374,
! basic_IMSRG.f90:4056-4199), running in libimsrg_hip.so through the
! bind(C) module imsrg_hip_c.  Compiled with the reference (it uses
! basic_IMSRG and cross_coupled); a caller swaps its use statement, e.g.
!     use imsrg_hip_shim, only: magnus_decouple => magnus_decouple_hip
! and calls hip_init(jbas) once after allocate_blocks.
!
! Term-level routines copy their operands to device mirrors and the result
! back on every call (PCIe-bound: for correctness checks and porting);
! magnus_decouple_hip keeps the whole flow on the device and copies H in
! and H(s), Omega out once.
module imsrg_hip_shim
  use basic_IMSRG
  use cross_coupled
  use imsrg_hip_c
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: hip_init, hip_finalize, to_device, from_device
  public :: commutator_110_hip, commutator_220_hip, commutator_111_hip, commutator_121_hip, commutator_122_hip
  public :: commutator_222_pp_hh_hip, commutator_221_hip, calculate_cross_coupled_hip, commutator_222_ph_hip
  public :: BCH_EXPAND_hip, MAGNUS_EXPAND_hip, euler_step_hip, magnus_decouple_hip
  public :: build_gs_white_hip, build_gs_atan_hip, build_gs_w2_hip, mbpt2_hip, mat_frob_norm_hip
  public :: TS_commutator_111_hip, TS_commutator_121_hip, TS_commutator_211_hip, TS_commutator_122_hip
  public :: TS_commutator_212_hip, TS_commutator_222_pp_hh_hip, TS_commutator_221_hip, TS_commutator_222_ph_hip
  public :: BCH_TENSOR_hip, transform_sq_op_BCH_hip

  integer(c_int), parameter :: SZ_OP = 5, SZ_CC = 7, SZ_WCC = 8, SZ_NCC = 6
  type(c_ptr), save :: hb = c_null_ptr      ! imsrg_basis
  type(c_ptr), save :: hf = c_null_ptr      ! imsrg_flow
  type(c_ptr), save :: dop(6) = c_null_ptr  ! device operator mirrors: L, R, RES, w1, w2, spare
  type(c_ptr), save :: dcc(3) = c_null_ptr  ! device cc_mat mirrors: LCC, RCC, WCC
  ! tensor operators: one library tensor type (rank, dpar) at a time
  integer(c_int), parameter :: TSZ_OP = 0, TSZ_CC = 3
  type(c_ptr), save :: ht = c_null_ptr      ! imsrg_tensor
  integer, save :: ht_rank = -1, ht_dpar = -1
  type(c_ptr), save :: dt(4) = c_null_ptr   ! device tensor mirrors: R, RES, w1, w2
  type(c_ptr), save :: dtcc = c_null_ptr    ! device pandya_mat mirror (CCX and CCR of every block)

  interface
    function hipMemcpy(dst, src, bytes, kind) bind(C, name="hipMemcpy")
      import
      integer(c_int) :: hipMemcpy
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: bytes
      integer(c_int), value :: kind
    end function hipMemcpy
  end interface
  integer(c_int), parameter :: H2D = 1, D2H = 2

contains

  subroutine hip_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(imsrg_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    print *, 'imsrg_hip (', where, '): ', msg(1:n)
    if (rc == 3) STOP 'clipping BCH'                 ! the reference's own text
    STOP 'imsrg_hip error'
  end subroutine hip_check

  ! basis and flow workspace of the library, once after allocate_blocks
  subroutine hip_init(jbas)
    type(spd), intent(in) :: jbas
    integer(c_int), allocatable :: n(:), l(:), jj(:), tz(:), hole(:)
    integer :: k
    if (c_associated(hb)) return
    n = jbas%nn; l = jbas%ll; jj = jbas%jj; tz = jbas%itzp; hole = jbas%con
    call hip_check(imsrg_basis_create(int(jbas%total_orbits, c_int), n, l, jj, tz, hole, hb), 'basis_create')
    call hip_check(imsrg_flow_create(hb, hf), 'flow_create')
    call hip_check(imsrg_flow_set_graphs(hf, 1_c_int), 'flow_set_graphs')
    do k = 1, 6
      call hip_check(imsrg_alloc(imsrg_basis_size(hb, SZ_OP), dop(k)), 'alloc')
    end do
    call hip_check(imsrg_alloc(imsrg_basis_size(hb, SZ_CC), dcc(1)), 'alloc')
    call hip_check(imsrg_alloc(imsrg_basis_size(hb, SZ_CC), dcc(2)), 'alloc')
    call hip_check(imsrg_alloc(imsrg_basis_size(hb, SZ_WCC), dcc(3)), 'alloc')
  end subroutine hip_init

  subroutine hip_finalize()
    integer :: k
    if (.not. c_associated(hb)) return
    do k = 1, 6
      call hip_check(imsrg_free(dop(k)), 'free')
    end do
    do k = 1, 3
      call hip_check(imsrg_free(dcc(k)), 'free')
    end do
    call tensor_release()
    call hip_check(imsrg_flow_destroy(hf), 'flow_destroy')
    call hip_check(imsrg_basis_destroy(hb), 'basis_destroy')
    hb = c_null_ptr; hf = c_null_ptr
  end subroutine hip_finalize

  ! host sq_op -> device mirror d (the library's flat layout, component by component)
  subroutine to_device(op, d)
    type(sq_op), intent(in) :: op
    type(c_ptr), intent(in) :: d
    real(c_double) :: e0(1)
    integer :: q, g
    e0(1) = op%E0
    call hip_check(imsrg_op_put(hb, d, -1_c_int, 0_c_int, e0), 'op_put')
    call hip_check(imsrg_op_put(hb, d, -1_c_int, 1_c_int, op%fhh), 'op_put')
    call hip_check(imsrg_op_put(hb, d, -1_c_int, 2_c_int, op%fph), 'op_put')
    call hip_check(imsrg_op_put(hb, d, -1_c_int, 3_c_int, op%fpp), 'op_put')
    do q = 1, op%nblocks
      do g = 1, 6
        if (size(op%mat(q)%gam(g)%X) > 0) &
          call hip_check(imsrg_op_put(hb, d, int(q - 1, c_int), int(g - 1, c_int), op%mat(q)%gam(g)%X), 'op_put')
      end do
    end do
  end subroutine to_device

  subroutine from_device(d, op)
    type(c_ptr), intent(in) :: d
    type(sq_op), intent(inout) :: op
    real(c_double) :: e0(1)
    integer :: q, g
    call hip_check(imsrg_op_get(hb, d, -1_c_int, 0_c_int, e0), 'op_get')
    op%E0 = e0(1)
    call hip_check(imsrg_op_get(hb, d, -1_c_int, 1_c_int, op%fhh), 'op_get')
    call hip_check(imsrg_op_get(hb, d, -1_c_int, 2_c_int, op%fph), 'op_get')
    call hip_check(imsrg_op_get(hb, d, -1_c_int, 3_c_int, op%fpp), 'op_get')
    do q = 1, op%nblocks
      do g = 1, 6
        if (size(op%mat(q)%gam(g)%X) > 0) &
          call hip_check(imsrg_op_get(hb, d, int(q - 1, c_int), int(g - 1, c_int), op%mat(q)%gam(g)%X), 'op_get')
      end do
    end do
  end subroutine from_device

  ! cc_mat <-> device: CCX(q)%X (rlen x nb, column-major) at the library's
  ! block offsets (cross_coupled.f90:61-177 and the same block order)
  subroutine cc_copy(CC, d, dir)
    type(cc_mat), intent(inout), target :: CC
    type(c_ptr), intent(in) :: d
    integer(c_int), intent(in) :: dir
    integer(c_int), allocatable :: rlen(:), ncol(:), jtp(:)
    integer(c_int64_t), allocatable :: ccoff(:), wcoff(:)
    integer :: q, m
    integer(c_size_t) :: bytes
    type(c_ptr) :: dev
    m = int(imsrg_basis_size(hb, SZ_NCC))
    allocate(rlen(m), ncol(m), jtp(3 * m), ccoff(m), wcoff(m))
    call hip_check(imsrg_basis_cc(hb, rlen, ncol, ccoff, wcoff, jtp), 'basis_cc')
    do q = 1, m
      bytes = int(rlen(q), c_size_t) * int(ncol(q), c_size_t) * 8_c_size_t
      if (bytes == 0) cycle
      dev = transfer(transfer(d, 0_c_intptr_t) + 8_c_intptr_t * ccoff(q), d)
      if (dir == H2D) then
        call hip_check(hipMemcpy(dev, c_loc(CC%CCX(q)%X), bytes, H2D), 'hipMemcpy')
      else
        call hip_check(hipMemcpy(c_loc(CC%CCX(q)%X), dev, bytes, D2H), 'hipMemcpy')
      end if
    end do
  end subroutine cc_copy

  ! ---------------------------------------------------------------- terms
  real(8) function commutator_110_hip(L, R, jbas)
    type(sq_op) :: L, R
    type(spd) :: jbas
    real(c_double) :: v(1)
    call to_device(L, dop(1)); call to_device(R, dop(2))
    call hip_check(imsrg_commutator_110(hb, dop(1), int(L%herm, c_int), dop(2), int(R%herm, c_int), v, c_null_ptr), &
                   'commutator_110')
    commutator_110_hip = v(1)
  end function commutator_110_hip

  real(8) function commutator_220_hip(L, R, jbas)
    type(sq_op) :: L, R
    type(spd) :: jbas
    real(c_double) :: v(1)
    call to_device(L, dop(1)); call to_device(R, dop(2))
    call hip_check(imsrg_commutator_220(hb, dop(1), int(L%herm, c_int), dop(2), int(R%herm, c_int), v, c_null_ptr), &
                   'commutator_220')
    commutator_220_hip = v(1)
  end function commutator_220_hip

  ! 111 and 122 assign their parts of RES, 121 accumulates: RES goes down and back
  subroutine commutator_111_hip(L, R, RES, jbas)
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call to_device(L, dop(1)); call to_device(R, dop(2)); call to_device(RES, dop(3))
    call hip_check(imsrg_commutator_111(hb, dop(1), int(L%herm, c_int), dop(2), int(R%herm, c_int), dop(3), &
                                        int(RES%herm, c_int), c_null_ptr), 'commutator_111')
    call from_device(dop(3), RES)
  end subroutine commutator_111_hip

  subroutine commutator_121_hip(L, R, RES, jbas)
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call to_device(L, dop(1)); call to_device(R, dop(2)); call to_device(RES, dop(3))
    call hip_check(imsrg_commutator_121(hb, dop(1), int(L%herm, c_int), dop(2), int(R%herm, c_int), dop(3), &
                                        int(RES%herm, c_int), c_null_ptr), 'commutator_121')
    call from_device(dop(3), RES)
  end subroutine commutator_121_hip

  subroutine commutator_122_hip(L, R, RES, jbas)
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call to_device(L, dop(1)); call to_device(R, dop(2)); call to_device(RES, dop(3))
    call hip_check(imsrg_commutator_122(hb, dop(1), int(L%herm, c_int), dop(2), int(R%herm, c_int), dop(3), &
                                        int(RES%herm, c_int), c_null_ptr), 'commutator_122')
    call from_device(dop(3), RES)
  end subroutine commutator_122_hip

  subroutine commutator_222_pp_hh_hip(L, R, RES, w1, w2, jbas)
    type(sq_op) :: L, R, RES, w1, w2
    type(spd) :: jbas
    call to_device(L, dop(1)); call to_device(R, dop(2)); call to_device(RES, dop(3))
    call hip_check(imsrg_commutator_222_pp_hh(hb, dop(1), int(L%herm, c_int), dop(2), int(R%herm, c_int), dop(3), &
                                              int(RES%herm, c_int), dop(4), dop(5), c_null_ptr), 'commutator_222_pp_hh')
    call from_device(dop(3), RES); call from_device(dop(4), w1); call from_device(dop(5), w2)
  end subroutine commutator_222_pp_hh_hip

  subroutine commutator_221_hip(L, R, RES, w1, w2, jbas)
    type(sq_op) :: L, R, RES, w1, w2
    type(spd) :: jbas
    call to_device(L, dop(1)); call to_device(R, dop(2)); call to_device(RES, dop(3))
    call to_device(w1, dop(4)); call to_device(w2, dop(5))
    call hip_check(imsrg_commutator_221(hb, dop(1), int(L%herm, c_int), dop(2), int(R%herm, c_int), dop(3), &
                                        int(RES%herm, c_int), dop(4), dop(5), c_null_ptr), 'commutator_221')
    call from_device(dop(3), RES)
  end subroutine commutator_221_hip

  subroutine calculate_cross_coupled_hip(HS, CCME, jbas)
    type(sq_op) :: HS
    type(cc_mat) :: CCME
    type(spd) :: jbas
    call to_device(HS, dop(1))
    call hip_check(imsrg_calculate_cross_coupled(hb, dop(1), int(HS%herm, c_int), dcc(1), c_null_ptr), &
                   'calculate_cross_coupled')
    call cc_copy(CCME, dcc(1), D2H)
  end subroutine calculate_cross_coupled_hip

  ! WCC is a workspace here as in the reference: it is not copied back
  subroutine commutator_222_ph_hip(LCC, RCC, RES, WCC, jbas)
    type(cc_mat) :: LCC, RCC, WCC
    type(sq_op) :: RES
    type(spd) :: jbas
    call cc_copy(LCC, dcc(1), H2D); call cc_copy(RCC, dcc(2), H2D)
    call to_device(RES, dop(3))
    call hip_check(imsrg_commutator_222_ph(hb, dcc(1), int(LCC%herm, c_int), dcc(2), int(RCC%herm, c_int), dop(3), &
                                           int(RES%herm, c_int), dcc(3), c_null_ptr), 'commutator_222_ph')
    call from_device(dop(3), RES)
  end subroutine commutator_222_ph_hip

  ! ------------------------------------------------------- BCH and Magnus
  subroutine BCH_EXPAND_hip(HS, G, H, jbas, quads)
    type(sq_op) :: HS, G, H
    type(spd) :: jbas
    character(1) :: quads
    real(c_double) :: advals(40)
    integer(c_int) :: nt(1)
    if (quads /= 'n') STOP 'imsrg_hip: BCH_EXPAND with quads /= n is not built'
    call to_device(G, dop(1)); call to_device(H, dop(2))
    call hip_check(imsrg_bch_expand(hf, dop(3), dop(1), int(G%herm, c_int), dop(2), int(H%herm, c_int), &
                                    dble(real(1e-6)), 40_c_int, nt, advals, c_null_ptr), 'bch_expand')
    HS%herm = H%herm
    call from_device(dop(3), HS)
  end subroutine BCH_EXPAND_hip

  subroutine MAGNUS_EXPAND_hip(DG, G, AD, jbas)
    type(sq_op) :: DG, G, AD
    type(spd) :: jbas
    integer(c_int) :: nc(1)
    call to_device(DG, dop(1)); call to_device(G, dop(2))
    call hip_check(imsrg_magnus_expand(hf, dop(1), int(DG%herm, c_int), dop(2), int(G%herm, c_int), &
                                       dble(real(1e-6)), nc, c_null_ptr), 'magnus_expand')
    call from_device(dop(1), DG)
  end subroutine MAGNUS_EXPAND_hip

  subroutine euler_step_hip(G, DG, s, stp)
    type(sq_op) :: G, DG
    real(8) :: s, stp
    call to_device(G, dop(1)); call to_device(DG, dop(2))
    call hip_check(imsrg_euler_step(hb, dop(1), dop(2), stp, c_null_ptr), 'euler_step')
    call from_device(dop(1), G)
    s = s + stp
  end subroutine euler_step_hip

  subroutine build_gs_white_hip(H, ETA, jbas)
    type(sq_op) :: H, ETA
    type(spd) :: jbas
    call to_device(H, dop(1)); call to_device(ETA, dop(2))
    call hip_check(imsrg_build_gs_white(hb, dop(1), dop(2), c_null_ptr), 'build_gs_white')
    call from_device(dop(2), ETA)
  end subroutine build_gs_white_hip

  subroutine build_gs_atan_hip(H, ETA, jbas)
    type(sq_op) :: H, ETA
    type(spd) :: jbas
    call to_device(H, dop(1)); call to_device(ETA, dop(2))
    call hip_check(imsrg_build_gs_atan(hb, dop(1), dop(2), c_null_ptr), 'build_gs_atan')
    call from_device(dop(2), ETA)
  end subroutine build_gs_atan_hip

  subroutine build_gs_w2_hip(H, ETA, jbas)
    type(sq_op) :: H, ETA
    type(spd) :: jbas
    call to_device(H, dop(1)); call to_device(ETA, dop(2))
    call hip_check(imsrg_build_gs_w2(hf, dop(1), dop(2), c_null_ptr), 'build_gs_w2')
    call from_device(dop(2), ETA)
  end subroutine build_gs_w2_hip

  real(8) function mbpt2_hip(H, jbas)
    type(sq_op) :: H
    type(spd) :: jbas
    real(c_double) :: v(1)
    call to_device(H, dop(1))
    call hip_check(imsrg_mbpt2(hb, dop(1), v, c_null_ptr), 'mbpt2')
    mbpt2_hip = v(1)
  end function mbpt2_hip

  real(8) function mat_frob_norm_hip(op)
    type(sq_op) :: op
    real(c_double) :: v(1)
    call to_device(op, dop(1))
    call hip_check(imsrg_mat_frob_norm(hb, dop(1), v, c_null_ptr), 'mat_frob_norm')
    mat_frob_norm_hip = v(1)
  end function mat_frob_norm_hip

  ! ------------------------------------------------------------ tensors
  subroutine tensor_release()
    integer :: k
    if (.not. c_associated(ht)) return
    do k = 1, 4
      call hip_check(imsrg_free(dt(k)), 'free')
    end do
    call hip_check(imsrg_free(dtcc), 'free')
    call hip_check(imsrg_tensor_destroy(ht), 'tensor_destroy')
    ht = c_null_ptr; ht_rank = -1; ht_dpar = -1
  end subroutine tensor_release

  ! the library's tensor layout of op's rank and parity (allocate_tensor)
  subroutine tensor_ready(op)
    type(sq_op), intent(in) :: op
    integer :: k
    if (c_associated(ht) .and. op%rank == ht_rank .and. op%dpar == ht_dpar) return
    call tensor_release()
    call hip_check(imsrg_tensor_create(hb, int(op%rank, c_int), int(op%dpar, c_int), ht), 'tensor_create')
    ht_rank = op%rank; ht_dpar = op%dpar
    do k = 1, 4
      call hip_check(imsrg_alloc(imsrg_tensor_size(ht, TSZ_OP), dt(k)), 'alloc')
    end do
    call hip_check(imsrg_alloc(max(imsrg_tensor_size(ht, TSZ_CC), 1_c_int64_t), dtcc), 'alloc')
  end subroutine tensor_ready

  ! host tensor sq_op -> device mirror (E0, fhh, fph, fpp, tblck(q)%tgam(g)%X)
  subroutine tensor_to_device(op, d)
    type(sq_op), intent(in) :: op
    type(c_ptr), intent(in) :: d
    real(c_double) :: e0(1)
    integer :: q, g
    e0(1) = op%E0
    call hip_check(imsrg_tensor_op_put(ht, d, -1_c_int, 0_c_int, e0), 'tensor_op_put')
    call hip_check(imsrg_tensor_op_put(ht, d, -1_c_int, 1_c_int, op%fhh), 'tensor_op_put')
    call hip_check(imsrg_tensor_op_put(ht, d, -1_c_int, 2_c_int, op%fph), 'tensor_op_put')
    call hip_check(imsrg_tensor_op_put(ht, d, -1_c_int, 3_c_int, op%fpp), 'tensor_op_put')
    do q = 1, op%nblocks
      do g = 1, 9
        if (size(op%tblck(q)%tgam(g)%X) > 0) call hip_check(imsrg_tensor_op_put(ht, d, int(q - 1, c_int), &
            int(g - 1, c_int), op%tblck(q)%tgam(g)%X), 'tensor_op_put')
      end do
    end do
  end subroutine tensor_to_device

  subroutine tensor_from_device(d, op)
    type(c_ptr), intent(in) :: d
    type(sq_op), intent(inout) :: op
    real(c_double) :: e0(1)
    integer :: q, g
    call hip_check(imsrg_tensor_op_get(ht, d, -1_c_int, 0_c_int, e0), 'tensor_op_get')
    op%E0 = e0(1)
    call hip_check(imsrg_tensor_op_get(ht, d, -1_c_int, 1_c_int, op%fhh), 'tensor_op_get')
    call hip_check(imsrg_tensor_op_get(ht, d, -1_c_int, 2_c_int, op%fph), 'tensor_op_get')
    call hip_check(imsrg_tensor_op_get(ht, d, -1_c_int, 3_c_int, op%fpp), 'tensor_op_get')
    do q = 1, op%nblocks
      do g = 1, 9
        if (size(op%tblck(q)%tgam(g)%X) > 0) call hip_check(imsrg_tensor_op_get(ht, d, int(q - 1, c_int), &
            int(g - 1, c_int), op%tblck(q)%tgam(g)%X), 'tensor_op_get')
      end do
    end do
  end subroutine tensor_from_device

  ! TS_commutators.f90: L scalar, R the tensor (the second argument), RES a tensor
  subroutine TS_commutator_111_hip(L, R, RES, jbas)
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call tensor_ready(R)
    call to_device(L, dop(1)); call tensor_to_device(R, dt(1)); call tensor_to_device(RES, dt(2))
    call hip_check(imsrg_ts_commutator_111(ht, dop(1), int(L%herm, c_int), dt(1), int(R%herm, c_int), dt(2), &
                                           int(RES%herm, c_int), c_null_ptr), 'ts_commutator_111')
    call tensor_from_device(dt(2), RES)
  end subroutine TS_commutator_111_hip

  subroutine TS_commutator_121_hip(L, R, RES, jbas)
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call tensor_ready(R)
    call to_device(L, dop(1)); call tensor_to_device(R, dt(1)); call tensor_to_device(RES, dt(2))
    call hip_check(imsrg_ts_commutator_121(ht, dop(1), int(L%herm, c_int), dt(1), int(R%herm, c_int), dt(2), &
                                           int(RES%herm, c_int), c_null_ptr), 'ts_commutator_121')
    call tensor_from_device(dt(2), RES)
  end subroutine TS_commutator_121_hip

  subroutine TS_commutator_211_hip(LCC, R, RES, jbas)
    type(cc_mat) :: LCC
    type(sq_op) :: R, RES
    type(spd) :: jbas
    call tensor_ready(R)
    call cc_copy(LCC, dcc(1), H2D)
    call tensor_to_device(R, dt(1)); call tensor_to_device(RES, dt(2))
    call hip_check(imsrg_ts_commutator_211(ht, dcc(1), int(LCC%herm, c_int), dt(1), int(R%herm, c_int), dt(2), &
                                           int(RES%herm, c_int), c_null_ptr), 'ts_commutator_211')
    call tensor_from_device(dt(2), RES)
  end subroutine TS_commutator_211_hip

  subroutine TS_commutator_122_hip(L, R, RES, jbas)
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call tensor_ready(R)
    call to_device(L, dop(1)); call tensor_to_device(R, dt(1)); call tensor_to_device(RES, dt(2))
    call hip_check(imsrg_ts_commutator_122(ht, dop(1), int(L%herm, c_int), dt(1), int(R%herm, c_int), dt(2), &
                                           int(RES%herm, c_int), c_null_ptr), 'ts_commutator_122')
    call tensor_from_device(dt(2), RES)
  end subroutine TS_commutator_122_hip

  subroutine TS_commutator_212_hip(L, R, RES, jbas)
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call tensor_ready(R)
    call to_device(L, dop(1)); call tensor_to_device(R, dt(1)); call tensor_to_device(RES, dt(2))
    call hip_check(imsrg_ts_commutator_212(ht, dop(1), int(L%herm, c_int), dt(1), int(R%herm, c_int), dt(2), &
                                           int(RES%herm, c_int), c_null_ptr), 'ts_commutator_212')
    call tensor_from_device(dt(2), RES)
  end subroutine TS_commutator_212_hip

  subroutine TS_commutator_222_pp_hh_hip(L, R, RES, w1, w2, jbas)
    type(sq_op) :: L, R, RES, w1, w2
    type(spd) :: jbas
    call tensor_ready(R)
    call to_device(L, dop(1)); call tensor_to_device(R, dt(1)); call tensor_to_device(RES, dt(2))
    call hip_check(imsrg_ts_commutator_222_pp_hh(ht, dop(1), int(L%herm, c_int), dt(1), int(R%herm, c_int), dt(2), &
                                                 int(RES%herm, c_int), dt(3), dt(4), c_null_ptr), 'ts_commutator_222_pp_hh')
    call tensor_from_device(dt(2), RES)
    call tensor_from_device(dt(3), w1)
    call tensor_from_device(dt(4), w2)
  end subroutine TS_commutator_222_pp_hh_hip

  subroutine TS_commutator_221_hip(w1, w2, pm, RES, jbas)
    type(sq_op) :: w1, w2, RES
    integer, intent(in) :: pm
    type(spd) :: jbas
    call tensor_ready(RES)
    call tensor_to_device(w1, dt(3)); call tensor_to_device(w2, dt(4)); call tensor_to_device(RES, dt(2))
    call hip_check(imsrg_ts_commutator_221(ht, dt(3), dt(4), int(pm, c_int), dt(2), int(RES%herm, c_int), &
                                           c_null_ptr), 'ts_commutator_221')
    call tensor_from_device(dt(2), RES)
  end subroutine TS_commutator_221_hip

  ! RCC (the reference's pandya_mat workspace) is filled on the device only
  subroutine TS_commutator_222_ph_hip(LCC, RCC, R, RES, jbas)
    type(cc_mat) :: LCC
    type(pandya_mat) :: RCC
    type(sq_op) :: R, RES
    type(spd) :: jbas
    call tensor_ready(R)
    call cc_copy(LCC, dcc(1), H2D)
    call tensor_to_device(R, dt(1)); call tensor_to_device(RES, dt(2))
    call hip_check(imsrg_ts_commutator_222_ph(ht, dcc(1), int(LCC%herm, c_int), dtcc, dt(1), int(R%herm, c_int), &
                                              dt(2), int(RES%herm, c_int), c_null_ptr), 'ts_commutator_222_ph')
    call tensor_from_device(dt(2), RES)
  end subroutine TS_commutator_222_ph_hip

  ! BCH_TENSOR, IMSRG_MAGNUS.f90:389-464: HS <- e^G HS e^-G for a tensor HS
  subroutine BCH_TENSOR_hip(G, HS, jbas, quads)
    type(sq_op) :: G, HS
    type(spd) :: jbas
    character(1) :: quads
    real(c_double) :: advals(30)
    integer(c_int) :: nt(1)
    call tensor_ready(HS)
    call to_device(G, dop(1)); call tensor_to_device(HS, dt(1))
    call hip_check(imsrg_bch_tensor(ht, dop(1), int(G%herm, c_int), dt(1), int(HS%herm, c_int), &
                                    dble(real(1e-6)), nt, advals, c_null_ptr), 'bch_tensor')
    call tensor_from_device(dt(1), HS)
  end subroutine BCH_TENSOR_hip

  ! transform_sq_op_BCH, IMSRG_MAGNUS.f90:131-146
  subroutine transform_sq_op_BCH_hip(Op, G, jbas, quads)
    type(sq_op) :: Op, G
    type(spd) :: jbas
    character(1) :: quads
    integer(c_int) :: nt(1)
    if (Op%rank > 0) then
      call BCH_TENSOR_hip(G, Op, jbas, quads)
      return
    end if
    if (quads /= 'n') STOP 'imsrg_hip: BCH_EXPAND with quads /= n is not built'
    call to_device(G, dop(1)); call to_device(Op, dop(2))
    call hip_check(imsrg_transform_sq_op_bch(hf, dop(2), int(Op%herm, c_int), dop(1), dble(real(1e-6)), 40_c_int, &
                                             nt, c_null_ptr), 'transform_sq_op_bch')
    call from_device(dop(2), Op)
  end subroutine transform_sq_op_BCH_hip

  ! ------------------------------------------------------------ the flow
  ! which of the library's generators build_generator is: applied once on
  ! the host to HS and compared with the device's white, atan and w2
  integer function identify_generator(HS, jbas, build_generator)
    type(sq_op) :: HS
    type(spd) :: jbas
    external :: build_generator
    type(sq_op) :: E_host, E_dev
    real(8) :: scale
    integer(c_int) :: k
    call duplicate_sq_op(HS, E_host)
    call duplicate_sq_op(HS, E_dev)
    call build_generator(HS, E_host, jbas)
    scale = max(maxval(abs(E_host%fph)), 1.d-300)
    call to_device(HS, dop(1))
    do k = 0, 2
      select case (k)
      case (0)
        call hip_check(imsrg_build_gs_white(hb, dop(1), dop(2), c_null_ptr), 'build_gs_white')
      case (1)
        call hip_check(imsrg_build_gs_atan(hb, dop(1), dop(2), c_null_ptr), 'build_gs_atan')
      case (2)
        call hip_check(imsrg_build_gs_w2(hf, dop(1), dop(2), c_null_ptr), 'build_gs_w2')
      end select
      call from_device(dop(2), E_dev)
      if (maxval(abs(E_dev%fph - E_host%fph)) <= 1.d-10 * scale) then
        identify_generator = k
        return
      end if
    end do
    STOP 'imsrg_hip: build_generator is not one of build_gs_white, build_gs_atan, build_gs_w2'
  end function identify_generator

  ! magnus_decouple, IMSRG_MAGNUS.f90:15-128: the whole flow on the device.
  ! Checkpointing (the module flag `checkpointing`) uses the reference's
  ! file names (scratch // prefix up to 'hw..' // '_s<s>.bin') and rules.
  subroutine magnus_decouple_hip(HS, G, jbas, quads, trips, build_generator)
    type(sq_op) :: HS, G
    type(spd) :: jbas
    character(1) :: quads, trips
    external :: build_generator
    integer, parameter :: maxsteps = 1000
    real(c_double), allocatable, target :: flog(:)
    real(c_double), target :: sx
    integer(c_int) :: steps(1)
    integer :: i, k, n0
    character(700) :: stem
    real(8) :: ds
    if (quads /= 'n' .or. trips /= 'n') STOP 'imsrg_hip: magnus_decouple with quads / trips is not built'
    ds = 0.5d0
    if (HS%lawson_beta >= 6.0) ds = 0.1d0
    if (.not. allocated(G%mat)) call duplicate_sq_op(HS, G)
    G%herm = -1
    call hip_check(imsrg_flow_set_generator(hf, int(identify_generator(HS, jbas, build_generator), c_int)), &
                   'flow_set_generator')
    allocate(flog(4 * maxsteps))
    call to_device(HS, dop(1))
    if (checkpointing .and. prefix(1:8) /= 'testcase') then
      i = index(prefix, 'hwHO0')
      if (i > 0) then
        i = i + 3
      else
        i = index(prefix, 'hw')
      end if
      stem = trim(scratch) // prefix(1:i + 3) // c_null_char
      call hip_check(imsrg_magnus_decouple_chk(hf, dop(1), dop(2), ds, dble(real(1e-6)), dble(real(1e-6)), 0_c_int, &
                                               dble(real(1e-6)), int(maxsteps, c_int), stem, 0_c_int, flog, steps, &
                                               c_loc(sx), c_null_ptr), 'magnus_decouple')
    else
      sx = -1.d0
      call hip_check(imsrg_magnus_decouple(hf, dop(1), dop(2), ds, dble(real(1e-6)), dble(real(1e-6)), 0_c_int, &
                                           dble(real(1e-6)), int(maxsteps, c_int), flog, steps, c_null_ptr), &
                     'magnus_decouple')
    end if
    call from_device(dop(1), HS)
    call from_device(dop(2), G)
    ! the reference's flow log: step, s, E0, E0 + mbpt2
    n0 = 0
    if (sx >= 0.d0) n0 = nint(sx / ds)
    do k = 1, steps(1) - n0
      write(*, '(I6,3(e15.7))') n0 + k, flog(4 * k - 3), flog(4 * k - 2), flog(4 * k - 1)
    end do
  end subroutine magnus_decouple_hip

end module imsrg_hip_shim
