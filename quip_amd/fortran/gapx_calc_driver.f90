! Command-line driver in the spirit of `quip E F V` (src/Programs/quip.f95:330-352)
! for one structure: reads a plain text structure file
!   line 1: XML potential path
!   line 2: N
!   line 3: lattice(1:3,1) lattice(1:3,2) lattice(1:3,3)
!   line 4: pbc as three 0/1
!   N lines: Z x y z
! and prints energy, forces and virial through IPModel_GAPX.  An optional second
! argument is the calc args_str; with atom_mask_name in it each atom line carries a
! fifth column (0/1, the mask), and energy_per_coordinate / local_gap_variance add
! ENERGY_PER_COORDINATE, LOCAL_GAP_VARIANCE and GAP_VARIANCE_GRADIENT lines.
program gapx_calc_driver
  use IPModel_GAPX_module
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  type(IPModel_GAPX) :: pot
  character(len=1024) :: infile, xmlpath, xmldir, args
  character(len=:), allocatable :: xml
  integer :: n, i, u, ipbc(3), err, sz, slash, im
  real(dp) :: lattice(3, 3), e, virial(3, 3)
  real(dp), allocatable :: pos(:, :), f(:, :), epc(:), lgv(:), gvg(:, :)
  integer, allocatable :: Z(:)
  logical, allocatable :: mask(:)
  logical :: has_mask
  call get_command_argument(1, infile)
  args = ''
  if (command_argument_count() >= 2) call get_command_argument(2, args)
  has_mask = index(args, 'atom_mask_name=') > 0
  open(newunit=u, file=trim(infile), status='old', action='read')
  read(u, '(a)') xmlpath
  read(u, *) n
  read(u, *) lattice
  read(u, *) ipbc
  allocate(pos(3, n), Z(n), f(3, n), mask(n), lgv(n), gvg(3, n), epc(16))
  mask = .true.
  do i = 1, n
     if (has_mask) then
        read(u, *) Z(i), pos(:, i), im
        mask(i) = im /= 0
     else
        read(u, *) Z(i), pos(:, i)
     end if
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
  if (len_trim(args) > 0) then
     deallocate(epc)
     allocate(epc(IPModel_GAPX_n_coordinates(pot)))
     call IPModel_GAPX_Calc(pot, n, pos, Z, lattice, ipbc /= 0, e=e, f=f, virial=virial, args_str=trim(args), &
          error=err, atom_mask=mask, energy_per_coordinate=epc, local_gap_variance=lgv, gap_variance_gradient=gvg)
  else
     call IPModel_GAPX_Calc(pot, n, pos, Z, lattice, ipbc /= 0, e=e, f=f, virial=virial, error=err)
  end if
  if (err /= 0) stop 3
  write(*, '(a,es25.17)') 'ENERGY ', e
  do i = 1, n
     write(*, '(a,i8,3es25.17)') 'FORCE ', i, f(:, i)
  end do
  do i = 1, 3
     write(*, '(a,3es25.17)') 'VIRIAL ', virial(i, :)
  end do
  if (index(args, 'energy_per_coordinate=') > 0) then
     do i = 1, size(epc)
        write(*, '(a,i8,es25.17)') 'ENERGY_PER_COORDINATE ', i, epc(i)
     end do
  end if
  if (index(args, 'local_gap_variance=') > 0) then
     do i = 1, n
        write(*, '(a,i8,es25.17)') 'LOCAL_GAP_VARIANCE ', i, lgv(i)
        write(*, '(a,i8,3es25.17)') 'GAP_VARIANCE_GRADIENT ', i, gvg(:, i)
     end do
  end if
  call IPModel_GAPX_Finalise(pot)
end program gapx_calc_driver
