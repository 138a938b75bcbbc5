! Golden-vector driver: the reference's own DIAGV_FO (diagv_fo.F) called through its module data, linked against
! the reference objects that oracle/build_nrlmol.py compiles from the reference sources.  Reads m, n and
! HAM(1:m,1:m) column by column from standard input; writes the sweep count line printed by the reference and
! PHIRES(1:m,1:m) column by column to phires.txt.
program diagv_fo_driver
  use for_diag1, only : HAM
  use diagv1, only : NORB, PHIRES
  use global_inputs, only : INBAS
  include 'PARAMA2'
  integer :: m, n, i, j
  read(*,*) m, n
  allocate(HAM(NDH, NDH))
  HAM = 0.0d0
  do j = 1, m
    read(*,*) (HAM(i, j), i = 1, m)
  end do
  allocate(PHIRES(m, m))
  INBAS = m
  NORB = n
  call DIAGV_FO(0)
  open(21, file='phires.txt')
  do j = 1, m
    write(21,'(ES26.17E3)') (PHIRES(i, j), i = 1, m)
  end do
  close(21)
end program diagv_fo_driver
