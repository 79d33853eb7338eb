*----------------------------------------------------------------------*
*     Test caller of libgecco_hip.so in the style of GeCCo's block
*     enumerator (contract/contr_op1op2_wmaps_c.f:1899,1927): F77 implicit
*     interfaces, 8-byte default INTEGER and LOGICAL (-fdefault-integer-8),
*     arguments by reference, host arrays.
*
*     Reads from unit 5 (list-directed): the lengths and contents of the
*     lists op1, op2 and the result; the kernel calls of one contraction
*     (route, factor, transpose flags, block offsets, the 12 block counts
*     and the 24 integer arrays of contr_blk1blk2_blocked_mm); then the
*     flattened arguments of one contr_op1op2 call.  Calls every kernel
*     call with its route (3: contr_blk1blk2_blocked_mm, 2:
*     contr_blk1blk2_wmaps_c), each twice (the second call reuses the
*     library's kept plan), writes the result, then runs the contr_op1op2
*     call on a copy of the initial result and writes that.
*----------------------------------------------------------------------*
      program contr_driver
      use gecco_hip_f, only: gecco_hip_check, gecco_hip_contr_op1op2
      implicit none
      external contr_blk1blk2_blocked_mm, contr_blk1blk2_wmaps_c
      integer :: n1, n2, nr, ncall, icall, k, ntot, irep
      integer :: iroute, ioff1, ioff2, ioff12
      integer :: idim(12), ilen(24), ipos(24)
      real(8) :: xfac, xscr(16)
      logical :: tra1, tra2, tra12
      integer :: itra1, itra2, itra12
      real(8), allocatable :: x1(:), x2(:), xr0(:), xr(:)
      integer, allocatable :: iarr(:)
      integer :: iupd, iusetr, it1, it2, it12, nsym, norbh, norbp, ierr
      integer :: iocc1(4,2), iocc2(4,2), iocc12(4,2), icnt(4,2)
      integer, allocatable :: igamh(:), igamp(:)

      read(5,*) n1, n2, nr
      allocate(x1(n1), x2(n2), xr0(nr), xr(nr))
      read(5,*) x1
      read(5,*) x2
      read(5,*) xr0
      xr = xr0
      read(5,*) ncall
      do icall = 1, ncall
        read(5,*) iroute, xfac, itra1, itra2, itra12
        read(5,*) ioff1, ioff2, ioff12
        read(5,*) idim
        read(5,*) ilen
        ntot = 0
        do k = 1, 24
          ipos(k) = ntot + 1
          ntot = ntot + max(ilen(k), 1)
        end do
        allocate(iarr(ntot))
        iarr = 0
        do k = 1, 24
          if (ilen(k).gt.0)
     &         read(5,*) iarr(ipos(k):ipos(k)+ilen(k)-1)
        end do
        tra1 = itra1.ne.0
        tra2 = itra2.ne.0
        tra12 = itra12.ne.0
        do irep = 1, 2
          if (iroute.eq.3) then
            call contr_blk1blk2_blocked_mm(xfac,
     &           xr(ioff12+1), x1(ioff1+1), x2(ioff2+1),
     &           tra1, tra2, tra12, xscr, 16, 200, -1,
     &           idim(1), idim(2), idim(3), idim(4),
     &           idim(5), idim(6), idim(7), idim(8),
     &           idim(9), idim(10), idim(11), idim(12),
     &           iarr(ipos(1)), iarr(ipos(2)), iarr(ipos(3)),
     &           iarr(ipos(4)), iarr(ipos(5)), iarr(ipos(6)),
     &           iarr(ipos(7)), iarr(ipos(8)), iarr(ipos(9)),
     &           iarr(ipos(10)), iarr(ipos(11)), iarr(ipos(12)),
     &           iarr(ipos(13)), iarr(ipos(14)), iarr(ipos(15)),
     &           iarr(ipos(16)), iarr(ipos(17)), iarr(ipos(18)),
     &           iarr(ipos(19)), iarr(ipos(20)), iarr(ipos(21)),
     &           iarr(ipos(22)), iarr(ipos(23)), iarr(ipos(24)))
          else
            call contr_blk1blk2_wmaps_c(xfac,
     &           xr(ioff12+1), x1(ioff1+1), x2(ioff2+1),
     &           tra1, tra2, tra12,
     &           idim(1), idim(2), idim(3), idim(4),
     &           idim(5), idim(6), idim(7), idim(8),
     &           idim(9), idim(10), idim(11), idim(12),
     &           iarr(ipos(1)), iarr(ipos(2)), iarr(ipos(3)),
     &           iarr(ipos(4)), iarr(ipos(5)), iarr(ipos(6)),
     &           iarr(ipos(7)), iarr(ipos(8)), iarr(ipos(9)),
     &           iarr(ipos(10)), iarr(ipos(11)), iarr(ipos(12)),
     &           iarr(ipos(13)), iarr(ipos(14)), iarr(ipos(15)),
     &           iarr(ipos(16)), iarr(ipos(17)), iarr(ipos(18)),
     &           iarr(ipos(19)), iarr(ipos(20)), iarr(ipos(21)),
     &           iarr(ipos(22)), iarr(ipos(23)), iarr(ipos(24)))
          end if
          call gecco_hip_check('contr_blk1blk2')
        end do
        deallocate(iarr)
      end do
      write(6,'(a)') 'kernel calls'
      write(6,'(es25.17)') xr
      ! one contr_op1op2 call with its arguments flattened
      read(5,*) xfac, iupd, iusetr, it1, it2, it12
      read(5,*) iocc1, iocc2, iocc12, icnt
      read(5,*) nsym, norbh
      allocate(igamh(norbh))
      read(5,*) igamh
      read(5,*) norbp
      allocate(igamp(norbp))
      read(5,*) igamp
      xr = xr0
      call gecco_hip_contr_op1op2(xfac, iupd, iocc1, iocc2, iocc12,
     &     icnt, it1, it2, it12, iusetr, nsym, norbh, igamh,
     &     norbp, igamp, x1, x2, xr, ierr)
      if (ierr.ne.0) call gecco_hip_check('gecco_hip_contr_op1op2')
      write(6,'(a)') 'contr_op1op2'
      write(6,'(es25.17)') xr
      end program
