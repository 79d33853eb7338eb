! ISO_C_BINDING interface of libgecco_hip.so for Fortran callers.
!
! GeCCo builds with 8-byte default INTEGER and LOGICAL; every argument below
! is integer(c_int64_t) (a LOGICAL(8) actual argument has the same storage:
! nonzero is true) and passed by reference, as the reference's own kernels
! take them.  The kernel-call entries keep the argument lists of
! contract/contr_blk1blk2_blocked_mm.f:2-20 and contract/
! contr_blk1blk2_wmaps_c.f:2-19, so an F77 caller with implicit interfaces
! (contract/contr_op1op2_wmaps_c.f:1899,1927) binds to them unchanged; the
! interfaces here serve callers that want them checked.  gecco_hip_check is
! the error path: the library records an error instead of calling quit
! (baserout/quit.F:36-50), and the check stops with its message.  The
! kernel-call interfaces are in the second module, gecco_hip_kernels_f.
module gecco_hip_f
  use iso_c_binding
  implicit none

  interface
    ! one binary contraction with contr_op1op2's semantics, its derived-type
    ! arguments flattened (include/gecco_hip_model.h)
    subroutine gecco_hip_contr_op1op2(xfac, update, iocc_op1, iocc_op2, iocc_op1op2, &
         iocc_cnt, tra_op1, tra_op2, tra_op1op2, use_tr, nsym, norb_h, igam_h, &
         norb_p, igam_p, xop1, xop2, xop1op2, ierr) bind(C, name='gecco_hip_contr_op1op2_')
      import :: c_double, c_int64_t
      real(c_double), intent(in) :: xfac, xop1(*), xop2(*)
      real(c_double), intent(inout) :: xop1op2(*)
      integer(c_int64_t), intent(in) :: update, iocc_op1(4, 2), iocc_op2(4, 2), &
           iocc_op1op2(4, 2), iocc_cnt(4, 2), tra_op1, tra_op2, tra_op1op2, use_tr, &
           nsym, norb_h, igam_h(*), norb_p, igam_p(*)
      integer(c_int64_t), intent(out) :: ierr
    end subroutine

    ! the same for operators with several vertices: merge maps, EX/CNT
    ! occupations and the reorder after the kernel calls (nreo 0 or 1)
    subroutine gecco_hip_contr_op1op2_n(xfac, update, njoined_op1, iocc_op1, njoined_op2, &
         iocc_op2, njoined_op1op2, iocc_op1op2, iocc_op1op2tmp, njoined_cnt, iocc_ex1, &
         iocc_ex2, iocc_cnt, merge_op1, merge_op2, merge_op1op2, merge_op2op1, nreo, &
         from_vtx, to_vtx, iocc_shift, sign_reo, use_tr, nsym, norb_h, igam_h, norb_p, &
         igam_p, xop1, xop2, xop1op2, ierr) bind(C, name='gecco_hip_contr_op1op2_n_')
      import :: c_double, c_int64_t
      real(c_double), intent(in) :: xfac, xop1(*), xop2(*)
      real(c_double), intent(inout) :: xop1op2(*)
      integer(c_int64_t), intent(in) :: update, njoined_op1, iocc_op1(4, 2, *), njoined_op2, &
           iocc_op2(4, 2, *), njoined_op1op2, iocc_op1op2(4, 2, *), iocc_op1op2tmp(4, 2, *), &
           njoined_cnt, iocc_ex1(4, 2, *), iocc_ex2(4, 2, *), iocc_cnt(4, 2, *), &
           merge_op1(*), merge_op2(*), merge_op1op2(*), merge_op2op1(*), nreo, from_vtx, &
           to_vtx, iocc_shift(4, 2), sign_reo, use_tr, nsym, norb_h, igam_h(*), norb_p, igam_p(*)
      integer(c_int64_t), intent(out) :: ierr
    end subroutine

    ! reorder of a whole list with reo_op_wmaps_c's semantics, flattened
    ! (include/gecco_hip_model.h); vertices numbered from 1
    subroutine gecco_hip_reo_op(fac, update, sign_reo, njoined_ori, iocc_ori, &
         njoined_reo, iocc_reo, from_vtx, to_vtx, iocc_shift, nsym, norb_h, igam_h, &
         norb_p, igam_p, xori, xreo, ierr) bind(C, name='gecco_hip_reo_op_')
      import :: c_double, c_int64_t
      real(c_double), intent(in) :: fac, xori(*)
      real(c_double), intent(inout) :: xreo(*)
      integer(c_int64_t), intent(in) :: update, sign_reo, njoined_ori, iocc_ori(4, 2, *), &
           njoined_reo, iocc_reo(4, 2, *), from_vtx, to_vtx, iocc_shift(4, 2), nsym, &
           norb_h, igam_h(*), norb_p, igam_p(*)
      integer(c_int64_t), intent(out) :: ierr
    end subroutine

    ! the allocation-free route-3 call (device operands only)
    subroutine contr_blk1blk2_blocked_mm_noalloc(xfac, xop1op2, xop1, xop2, &
         tra_op1, tra_op2, tra_op1op2, xscr, lenscr, lenblock, lenblock_cnt, &
         ncblk_op1, nablk_op1, ncblk_ex1, nablk_ex1, ncblk_op2, nablk_op2, &
         ncblk_ex2, nablk_ex2, ncblk_cnt, nablk_cnt, ncblk_op1op2, nablk_op1op2, &
         hpvxop1c, hpvxop1a, hpvxop2c, hpvxop2a, hpvxop1op2c, hpvxop1op2a, &
         lstrop1, lstrop2, lstrop1op2, lstr_ex1, lstr_ex2, lstr_cnt, &
         map_info_12c, map_info_12a, map_info_1c, map_info_1a, map_info_2c, map_info_2a, &
         map_ex1ex2c, map_ex1ex2a, map_ex1cntc, map_ex1cnta, map_ex2cntc, map_ex2cnta) &
         bind(C, name='contr_blk1blk2_blocked_mm_noalloc_')
      import :: c_double, c_int64_t, c_ptr
      type(c_ptr), value :: xop1op2, xop1, xop2
      real(c_double), intent(in) :: xfac
      real(c_double), intent(inout) :: xscr(*)
      integer(c_int64_t), intent(in) :: tra_op1, tra_op2, tra_op1op2, lenscr, lenblock, lenblock_cnt
      integer(c_int64_t), intent(in) :: ncblk_op1, nablk_op1, ncblk_ex1, nablk_ex1, &
           ncblk_op2, nablk_op2, ncblk_ex2, nablk_ex2, ncblk_cnt, nablk_cnt, &
           ncblk_op1op2, nablk_op1op2
      integer(c_int64_t), intent(in) :: hpvxop1c(*), hpvxop1a(*), hpvxop2c(*), hpvxop2a(*), &
           hpvxop1op2c(*), hpvxop1op2a(*), lstrop1(*), lstrop2(*), lstrop1op2(*), &
           lstr_ex1(*), lstr_ex2(*), lstr_cnt(*), map_info_12c(*), map_info_12a(*), &
           map_info_1c(*), map_info_1a(*), map_info_2c(*), map_info_2a(*), &
           map_ex1ex2c(*), map_ex1ex2a(*), map_ex1cntc(*), map_ex1cnta(*), &
           map_ex2cntc(*), map_ex2cnta(*)
    end subroutine

    integer(c_int) function gecco_hip_last_error(msg, msglen) bind(C, name='gecco_hip_last_error')
      import :: c_int, c_char, c_int64_t
      character(kind=c_char), intent(out) :: msg(*)
      integer(c_int64_t), value :: msglen
    end function

    integer(c_int) function gecco_hip_call_cache(entries, clear) bind(C, name='gecco_hip_call_cache')
      import :: c_int, c_int64_t
      integer(c_int64_t), intent(out) :: entries
      integer(c_int64_t), value :: clear
    end function

    integer(c_int) function gecco_hip_synchronize() bind(C, name='gecco_hip_synchronize')
      import :: c_int
    end function
  end interface

contains

  ! Stops with the library's message when the last call recorded an error
  ! (the role of quit() after a failed kernel call).
  subroutine gecco_hip_check(where)
    character(len=*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    character(len=512) :: text
    integer(c_int) :: code
    integer :: i
    code = gecco_hip_last_error(buf, int(size(buf), c_int64_t))
    if (code == 0) return
    text = ' '
    do i = 1, size(buf)
      if (buf(i) == c_null_char) exit
      text(i:i) = buf(i)
    end do
    write(*, '(a, a, a, i0, a, a)') 'gecco_hip error in ', where, ': code ', code, ': ', trim(text)
    error stop 'gecco_hip error'
  end subroutine

end module gecco_hip_f


! Checked interfaces of the kernel-call entries and of the route-3 collect
! and scatter steps, for callers that want them; an F77 caller with implicit interfaces (as GeCCo's enumerator) binds
! to the same symbols without this module, and a scope cannot use both.
module gecco_hip_kernels_f
  use iso_c_binding
  implicit none
  interface
    ! route-3 kernel call (contract/contr_blk1blk2_blocked_mm.f:2-20)
    subroutine contr_blk1blk2_blocked_mm(xfac, xop1op2, xop1, xop2, &
         tra_op1, tra_op2, tra_op1op2, xscr, lenscr, lenblock, lenblock_cnt, &
         ncblk_op1, nablk_op1, ncblk_ex1, nablk_ex1, ncblk_op2, nablk_op2, &
         ncblk_ex2, nablk_ex2, ncblk_cnt, nablk_cnt, ncblk_op1op2, nablk_op1op2, &
         hpvxop1c, hpvxop1a, hpvxop2c, hpvxop2a, hpvxop1op2c, hpvxop1op2a, &
         lstrop1, lstrop2, lstrop1op2, lstr_ex1, lstr_ex2, lstr_cnt, &
         map_info_12c, map_info_12a, map_info_1c, map_info_1a, map_info_2c, map_info_2a, &
         map_ex1ex2c, map_ex1ex2a, map_ex1cntc, map_ex1cnta, map_ex2cntc, map_ex2cnta) &
         bind(C, name='contr_blk1blk2_blocked_mm_')
      import :: c_double, c_int64_t
      real(c_double), intent(in) :: xfac, xop1(*), xop2(*)
      real(c_double), intent(inout) :: xop1op2(*), xscr(*)
      integer(c_int64_t), intent(in) :: tra_op1, tra_op2, tra_op1op2, lenscr, lenblock, lenblock_cnt
      integer(c_int64_t), intent(in) :: ncblk_op1, nablk_op1, ncblk_ex1, nablk_ex1, &
           ncblk_op2, nablk_op2, ncblk_ex2, nablk_ex2, ncblk_cnt, nablk_cnt, &
           ncblk_op1op2, nablk_op1op2
      integer(c_int64_t), intent(in) :: hpvxop1c(*), hpvxop1a(*), hpvxop2c(*), hpvxop2a(*), &
           hpvxop1op2c(*), hpvxop1op2a(*), lstrop1(*), lstrop2(*), lstrop1op2(*), &
           lstr_ex1(*), lstr_ex2(*), lstr_cnt(*), map_info_12c(*), map_info_12a(*), &
           map_info_1c(*), map_info_1a(*), map_info_2c(*), map_info_2a(*), &
           map_ex1ex2c(*), map_ex1ex2a(*), map_ex1cntc(*), map_ex1cnta(*), &
           map_ex2cntc(*), map_ex2cnta(*)
    end subroutine

    ! route-2 kernel call (contract/contr_blk1blk2_wmaps_c.f:2-19)
    subroutine contr_blk1blk2_wmaps_c(xfac, xop1op2, xop1, xop2, &
         tra_op1, tra_op2, tra_op1op2, &
         ncblk_op1, nablk_op1, ncblk_ex1, nablk_ex1, ncblk_op2, nablk_op2, &
         ncblk_ex2, nablk_ex2, ncblk_cnt, nablk_cnt, ncblk_op1op2, nablk_op1op2, &
         hpvxop1c, hpvxop1a, hpvxop2c, hpvxop2a, hpvxop1op2c, hpvxop1op2a, &
         lstrop1, lstrop2, lstrop1op2, lstr_ex1, lstr_ex2, lstr_cnt, &
         map_info_12c, map_info_12a, map_info_1c, map_info_1a, map_info_2c, map_info_2a, &
         map_ex1ex2c, map_ex1ex2a, map_ex1cntc, map_ex1cnta, map_ex2cntc, map_ex2cnta) &
         bind(C, name='contr_blk1blk2_wmaps_c_')
      import :: c_double, c_int64_t
      real(c_double), intent(in) :: xfac, xop1(*), xop2(*)
      real(c_double), intent(inout) :: xop1op2(*)
      integer(c_int64_t), intent(in) :: tra_op1, tra_op2, tra_op1op2
      integer(c_int64_t), intent(in) :: ncblk_op1, nablk_op1, ncblk_ex1, nablk_ex1, &
           ncblk_op2, nablk_op2, ncblk_ex2, nablk_ex2, ncblk_cnt, nablk_cnt, &
           ncblk_op1op2, nablk_op1op2
      integer(c_int64_t), intent(in) :: hpvxop1c(*), hpvxop1a(*), hpvxop2c(*), hpvxop2a(*), &
           hpvxop1op2c(*), hpvxop1op2a(*), lstrop1(*), lstrop2(*), lstrop1op2(*), &
           lstr_ex1(*), lstr_ex2(*), lstr_cnt(*), map_info_12c(*), map_info_12a(*), &
           map_info_1c(*), map_info_1a(*), map_info_2c(*), map_info_2a(*), &
           map_ex1ex2c(*), map_ex1ex2a(*), map_ex1cntc(*), map_ex1cnta(*), &
           map_ex2cntc(*), map_ex2cnta(*)
    end subroutine

    ! route-3 collect step (contract/collect_block_impr.f:2-17)
    subroutine collect_block_impr(xblk, xop, kc_st0, kc_nd0, kc_len, ka_st0, ka_nd0, ka_len, &
         k_len, kca_trp, lc_st0, lc_nd0, lc_len, la_st0, la_nd0, la_len, l_len, &
         map_kl_c, map_kl_a, ncblk_op, nablk_op, ldim_opc, ldim_opa, ncblk_k, nablk_k, &
         ncblk_l, nablk_l, lstr_kc, lstr_ka, lstr_lc, lstr_la, nstr_kc, nstr_ka, nstr_lc, &
         nstr_la, ireo_kc, ireo_ka, ireo_lc, ireo_la) bind(C, name='collect_block_impr_')
      import :: c_double, c_int64_t
      real(c_double), intent(out) :: xblk(*)
      real(c_double), intent(in) :: xop(*)
      integer(c_int64_t), intent(in) :: kc_st0, kc_nd0, kc_len, ka_st0, ka_nd0, ka_len, k_len, &
           kca_trp, lc_st0, lc_nd0, lc_len, la_st0, la_nd0, la_len, l_len, ncblk_op, nablk_op, &
           ncblk_k, nablk_k, ncblk_l, nablk_l
      integer(c_int64_t), intent(in) :: map_kl_c(*), map_kl_a(*), ldim_opc(*), ldim_opa(*), &
           lstr_kc(*), lstr_ka(*), lstr_lc(*), lstr_la(*), nstr_kc(*), nstr_ka(*), &
           nstr_lc(*), nstr_la(*), ireo_kc(*), ireo_ka(*), ireo_lc(*), ireo_la(*)
    end subroutine

    ! route-3 scatter step (contract/scatter_block.f:2-20)
    subroutine scatter_block(xop, xblk, kc_st0, kc_nd0, kc_len, ka_st0, ka_nd0, ka_len, &
         k_len, lc_st0, lc_nd0, lc_len, la_st0, la_nd0, la_len, l_len, &
         map_kl_c, map_kl_a, ncblk_op, nablk_op, ldim_opc, ldim_opa, ncblk_k, nablk_k, &
         ncblk_l, nablk_l, lstr_kc, lstr_ka, lstr_lc, lstr_la, nstr_kc, nstr_ka, nstr_lc, &
         nstr_la, ireo_kc, ireo_ka, ireo_lc, ireo_la, istr_kc, istr_ka, istr_lc, istr_la) &
         bind(C, name='scatter_block_')
      import :: c_double, c_int64_t
      real(c_double), intent(inout) :: xop(*)
      real(c_double), intent(in) :: xblk(*)
      integer(c_int64_t), intent(in) :: kc_st0, kc_nd0, kc_len, ka_st0, ka_nd0, ka_len, k_len, &
           lc_st0, lc_nd0, lc_len, la_st0, la_nd0, la_len, l_len, ncblk_op, nablk_op, &
           ncblk_k, nablk_k, ncblk_l, nablk_l
      integer(c_int64_t), intent(in) :: map_kl_c(*), map_kl_a(*), ldim_opc(*), ldim_opa(*), &
           lstr_kc(*), lstr_ka(*), lstr_lc(*), lstr_la(*), nstr_kc(*), nstr_ka(*), &
           nstr_lc(*), nstr_la(*), ireo_kc(*), ireo_ka(*), ireo_lc(*), ireo_la(*)
      integer(c_int64_t), intent(inout) :: istr_kc(*), istr_ka(*), istr_lc(*), istr_la(*)
    end subroutine
  end interface
end module gecco_hip_kernels_f
