! Independent Fortran side of the CC response-vector file test: writes and
! reads the record sequence of a closed-shell singlet CCSD vector (header
! NWRITE, MODFIL; orbital part; singles; doubles; six empty records) with
! plain unformatted sequential I/O, and prints the file names of the
! 'CC' // list // symmetry // index rule for a few cases.
!
!   rsp_layout w PATH NALLAI NT1 NT2   write vec0 = 0, vec1(k) = 0.25 k - 1,
!                                      vec2(k) = 1 / k
!   rsp_layout r PATH NALLAI NT1 NT2   read and print NWRITE, MODFIL and
!                                      the sums of vec1, k * vec1, vec2, k * vec2
!   rsp_layout n                       print file names
!
! Compiled with the default INTEGER kind and with -fdefault-integer-8, as
! the two Dalton integer builds.
program rsp_layout
  implicit none
  character(len=256) :: mode, path, arg
  integer :: nallai, nt1, nt2, k, nwrite
  character(len=10) :: modfil
  double precision, allocatable :: v0(:), v1(:), v2(:)
  call get_command_argument(1, mode)
  if (mode(1:1) == 'n') then
    call name('R0 ', 1, 0)
    call name('L0 ', 1, 0)
    call name('R1 ', 1, 12)
    call name('R1 ', 2, 345)
    call name('PL1', 3, 7)
    call name('LE ', 1, 1234)
    stop
  end if
  call get_command_argument(2, path)
  call get_command_argument(3, arg)
  read(arg, *) nallai
  call get_command_argument(4, arg)
  read(arg, *) nt1
  call get_command_argument(5, arg)
  read(arg, *) nt2
  allocate(v0(2 * nallai), v1(nt1), v2(nt2))
  open(unit=21, file=trim(path), form='unformatted', access='sequential')
  if (mode(1:1) == 'w') then
    v0 = 0d0
    do k = 1, nt1
      v1(k) = 0.25d0 * k - 1d0
    end do
    do k = 1, nt2
      v2(k) = 1d0 / k
    end do
    nwrite = 3
    modfil = 'CCSD'
    write(21) nwrite, modfil
    write(21) (v0(k), k = 1, 2 * nallai)
    write(21) (v1(k), k = 1, nt1)
    write(21) (v2(k), k = 1, nt2)
    do k = 1, 6
      write(21)
    end do
  else
    read(21) nwrite, modfil
    read(21) (v0(k), k = 1, 2 * nallai)
    read(21) (v1(k), k = 1, nt1)
    read(21) (v2(k), k = 1, nt2)
    do k = 1, 6
      read(21)
    end do
    print '(I3,1X,A10)', nwrite, modfil
    print '(4ES24.16)', sum(v1), sum([(k * v1(k), k = 1, nt1)]), sum(v2), sum([(k * v2(k), k = 1, nt2)])
  end if
  close(21)
contains
  subroutine name(list, isym, idx)
    character(len=3), intent(in) :: list
    integer, intent(in) :: isym, idx
    character(len=4) :: l4
    character(len=10) :: f
    integer :: i
    l4 = list
    if (l4(3:3) == ' ') l4(3:3) = '_'
    write(l4(4:4), '(I1)') isym
    if (idx <= 999) then
      write(f, '(A2,A4,1X,I3)') 'CC', l4, idx
      do i = 1, 10
        if (f(i:i) == ' ') f(i:i) = '_'
      end do
    else
      write(f, '(A2,A4,I4)') 'CC', l4, idx
    end if
    print '(A)', f
  end subroutine
end program
