! Command-line driver in the spirit of `quip E F V` (src/Programs/quip.f95:330-352)
! for one structure: reads a plain text structure file
!   line 1: XML potential path
!   line 2: N
!   line 3: lattice(1:3,1) lattice(1:3,2) lattice(1:3,3)
!   line 4: pbc as three 0/1
!   N lines: Z x y z
! and prints energy, forces and virial through IPModel_GAPX.
program gapx_calc_driver
  use IPModel_GAPX_module
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  type(IPModel_GAPX) :: pot
  character(len=1024) :: infile, xmlpath, xmldir
  character(len=:), allocatable :: xml
  integer :: n, i, u, ipbc(3), err, sz, slash
  real(dp) :: lattice(3, 3), e, virial(3, 3)
  real(dp), allocatable :: pos(:, :), f(:, :)
  integer, allocatable :: Z(:)
  call get_command_argument(1, infile)
  open(newunit=u, file=trim(infile), status='old', action='read')
  read(u, '(a)') xmlpath
  read(u, *) n
  read(u, *) lattice
  read(u, *) ipbc
  allocate(pos(3, n), Z(n), f(3, n))
  do i = 1, n
     read(u, *) Z(i), pos(:, i)
  end do
  close(u)
  inquire(file=trim(xmlpath), size=sz)
  allocate(character(len=sz) :: xml)
  open(newunit=u, file=trim(xmlpath), access='stream', form='unformatted', status='old', action='read')
  read(u) xml
  close(u)
  slash = index(xmlpath, '/', back=.true.)
  xmldir = '.'
  if (slash > 0) xmldir = xmlpath(1:slash - 1)
  call IPModel_GAPX_Initialise_str(pot, 'IP GAP', xml, trim(xmldir), error=err)
  if (err /= 0) stop 2
  call IPModel_GAPX_Print(pot)
  call IPModel_GAPX_Calc(pot, n, pos, Z, lattice, ipbc /= 0, e=e, f=f, virial=virial, error=err)
  if (err /= 0) stop 3
  write(*, '(a,es25.17)') 'ENERGY ', e
  do i = 1, n
     write(*, '(a,i8,3es25.17)') 'FORCE ', i, f(:, i)
  end do
  do i = 1, 3
     write(*, '(a,3es25.17)') 'VIRIAL ', virial(i, :)
  end do
  call IPModel_GAPX_Finalise(pot)
end program gapx_calc_driver
