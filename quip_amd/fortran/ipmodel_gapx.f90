! Fortran host module for libgapx.so (ISO_C_BINDING), shaped like QUIP's
! IPModel_GAP (src/Potentials/IPModel_GAP.f95): Initialise from the args
! string and the XML text of a GAP potential, Calc with caller-owned optional
! outputs e, local_e(N), f(3,N), virial(3,3), local_virial(9,N), an optional
! error status instead of an abort, Finalise and Print.  Arrays are passed
! straight through: QUIP's column-major pos(3,N), lattice(3,3) (columns are
! lattice vectors) and f(3,N) have exactly the C-ABI's memory layout.
module gapx_c_binding
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: GAPX_WANT_E = 1, GAPX_WANT_F = 2, GAPX_WANT_V = 4, &
       GAPX_WANT_LOCAL_E = 8, GAPX_WANT_LOCAL_V = 16, GAPX_WANT_VARIANCE = 32

  interface
     integer(c_int) function gapx_api_version() bind(c, name="gapx_api_version")
       import :: c_int
     end function gapx_api_version

     subroutine gapx_last_error(buf, len) bind(c, name="gapx_last_error")
       import :: c_char, c_int
       character(kind=c_char), dimension(*) :: buf
       integer(c_int), value :: len
     end subroutine gapx_last_error

     integer(c_int) function gapx_model_load(xml_text, xml_len, label, xml_dir, device, model) &
          bind(c, name="gapx_model_load")
       import :: c_char, c_int, c_int64_t, c_ptr
       character(kind=c_char), dimension(*) :: xml_text
       integer(c_int64_t), value :: xml_len
       type(c_ptr), value :: label
       character(kind=c_char), dimension(*) :: xml_dir
       integer(c_int), value :: device
       type(c_ptr) :: model
     end function gapx_model_load

     integer(c_int) function gapx_model_free(model) bind(c, name="gapx_model_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: model
     end function gapx_model_free

     integer(c_int) function gapx_model_cutoff(model, cutoff) bind(c, name="gapx_model_cutoff")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: model
       real(c_double) :: cutoff
     end function gapx_model_cutoff

     integer(c_int) function gapx_model_set_e_scale(model, e_scale) bind(c, name="gapx_model_set_e_scale")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: model
       real(c_double), value :: e_scale
     end function gapx_model_set_e_scale

     integer(c_int) function gapx_model_set_precision(model, prec) bind(c, name="gapx_model_set_precision")
       import :: c_int, c_ptr
       type(c_ptr), value :: model
       integer(c_int), value :: prec
     end function gapx_model_set_precision

     integer(c_int) function gapx_calc(model, n, pos, z, lattice, pbc, nl_cutoff, want, e, local_e, f, &
          virial, local_virial) bind(c, name="gapx_calc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: model
       integer(c_int), value :: n
       real(c_double), dimension(*) :: pos
       integer(c_int), dimension(*) :: z
       real(c_double), dimension(*) :: lattice
       integer(c_int), dimension(*) :: pbc
       real(c_double), value :: nl_cutoff
       integer(c_int), value :: want
       type(c_ptr), value :: e, local_e, f, virial, local_virial
     end function gapx_calc

     integer(c_int) function gapx_calc_ex(model, n, pos, z, lattice, pbc, nl_cutoff, want, atom_mask, &
          only_descriptor, reg, e, local_e, f, virial, local_virial, epc, lgv, gvg) bind(c, name="gapx_calc_ex")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: model
       integer(c_int), value :: n
       real(c_double), dimension(*) :: pos
       integer(c_int), dimension(*) :: z
       real(c_double), dimension(*) :: lattice
       integer(c_int), dimension(*) :: pbc
       real(c_double), value :: nl_cutoff
       integer(c_int), value :: want
       type(c_ptr), value :: atom_mask
       integer(c_int), value :: only_descriptor
       real(c_double), value :: reg
       type(c_ptr), value :: e, local_e, f, virial, local_virial, epc, lgv, gvg
     end function gapx_calc_ex

     integer(c_int) function gapx_descriptor_variances(model, coord, max_entries, variance, cutoff_sq, atoms, &
          n_entries) bind(c, name="gapx_descriptor_variances")
       import :: c_int, c_ptr
       type(c_ptr), value :: model
       integer(c_int), value :: coord, max_entries
       type(c_ptr), value :: variance, cutoff_sq, atoms
       integer(c_int) :: n_entries
     end function gapx_descriptor_variances

     integer(c_int) function gapx_model_info(model, n_coord, coord, dim, n_sparse) bind(c, name="gapx_model_info")
       import :: c_int, c_ptr
       type(c_ptr), value :: model
       integer(c_int) :: n_coord
       integer(c_int), value :: coord
       type(c_ptr), value :: dim, n_sparse
     end function gapx_model_info
  end interface
end module gapx_c_binding


module IPModel_GAPX_module
  use iso_c_binding
  use gapx_c_binding
  implicit none
  private

  integer, parameter :: dp = kind(1.0d0)

  public :: IPModel_GAPX, IPModel_GAPX_Initialise_str, IPModel_GAPX_Finalise, IPModel_GAPX_Calc, &
       IPModel_GAPX_Print, IPModel_GAPX_cutoff, IPModel_GAPX_last_error, IPModel_GAPX_n_coordinates

  type IPModel_GAPX
     type(c_ptr) :: model = c_null_ptr
     real(dp) :: cutoff = 0.0_dp
     real(dp) :: E_scale = 1.0_dp
     character(len=256) :: label = ''
     character(len=16) :: precision = 'fp64'
  end type IPModel_GAPX

contains

  ! value of key=value in a QUIP-style args string ('' when absent)
  function arg_value(args_str, key) result(val)
    character(len=*), intent(in) :: args_str, key
    character(len=256) :: val
    integer :: i, j
    val = ''
    i = index(' '//args_str, ' '//trim(key)//'=')
    if (i == 0) return
    i = i + len_trim(key) + 1
    j = index(args_str(i:), ' ')
    if (j == 0) then
       val = args_str(i:)
    else
       val = args_str(i:i + j - 2)
    end if
  end function arg_value

  function IPModel_GAPX_last_error() result(msg)
    character(len=512) :: msg
    character(kind=c_char) :: buf(512)
    integer :: k
    call gapx_last_error(buf, 512_c_int)
    msg = ''
    do k = 1, 512
       if (buf(k) == c_null_char) exit
       msg(k:k) = buf(k)
    end do
  end function IPModel_GAPX_last_error

  ! IPModel_GAP_Initialise_str (IPModel_GAP.f95:149-192): args_str "IP GAP label=... E_scale=..."
  ! param_str is the whole XML text; xml_dir resolves the sparseX files.
  subroutine IPModel_GAPX_Initialise_str(this, args_str, param_str, xml_dir, error, device)
    type(IPModel_GAPX), intent(inout) :: this
    character(len=*), intent(in) :: args_str, param_str, xml_dir
    integer, intent(out), optional :: error
    integer, intent(in), optional :: device
    character(kind=c_char), allocatable, target :: lab(:)
    character(len=256) :: v
    integer(c_int) :: st, dev
    type(c_ptr) :: labp
    integer :: k
    if (present(error)) error = 0
    call IPModel_GAPX_Finalise(this)
    this%label = arg_value(args_str, 'label')
    labp = c_null_ptr
    if (len_trim(this%label) > 0) then
       allocate(lab(len_trim(this%label) + 1))
       do k = 1, len_trim(this%label)
          lab(k) = this%label(k:k)
       end do
       lab(len_trim(this%label) + 1) = c_null_char
       labp = c_loc(lab)
    end if
    dev = 0
    if (present(device)) dev = device
    st = gapx_model_load(param_str, int(len(param_str), c_int64_t), labp, trim(xml_dir)//c_null_char, dev, &
         this%model)
    if (st /= 0) then
       call raise(st, 'IPModel_GAPX_Initialise_str: '//trim(IPModel_GAPX_last_error()), error)
       return
    end if
    st = gapx_model_cutoff(this%model, this%cutoff)
    v = arg_value(args_str, 'E_scale')
    if (len_trim(v) > 0) then
       read(v, *) this%E_scale
       st = gapx_model_set_e_scale(this%model, this%E_scale)
    end if
    v = arg_value(args_str, 'precision')
    if (len_trim(v) > 0) then
       this%precision = v
       select case (trim(v))
       case ('bf16x3')
          st = gapx_model_set_precision(this%model, 1_c_int)
       case ('fp16x3')
          st = gapx_model_set_precision(this%model, 2_c_int)
       case default
          st = gapx_model_set_precision(this%model, 0_c_int)
       end select
    end if
  end subroutine IPModel_GAPX_Initialise_str

  subroutine IPModel_GAPX_Finalise(this)
    type(IPModel_GAPX), intent(inout) :: this
    integer(c_int) :: st
    if (c_associated(this%model)) st = gapx_model_free(this%model)
    this%model = c_null_ptr
    this%cutoff = 0.0_dp
  end subroutine IPModel_GAPX_Finalise

  function IPModel_GAPX_cutoff(this) result(c)
    type(IPModel_GAPX), intent(in) :: this
    real(dp) :: c
    c = this%cutoff
  end function IPModel_GAPX_cutoff

  function IPModel_GAPX_n_coordinates(this) result(nc)
    type(IPModel_GAPX), intent(in) :: this
    integer :: nc
    integer(c_int) :: n, st
    n = 0
    st = gapx_model_info(this%model, n, -1_c_int, c_null_ptr, c_null_ptr)
    nc = int(n)
  end function IPModel_GAPX_n_coordinates

  subroutine IPModel_GAPX_Print(this)
    type(IPModel_GAPX), intent(in) :: this
    write(*, '(a,a,a,f10.5,a,a)') 'IPModel_GAPX : label=', trim(this%label), ' cutoff=', this%cutoff, &
         ' precision=', trim(this%precision)
  end subroutine IPModel_GAPX_Print

  subroutine raise(code, msg, error)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: msg
    integer, intent(out), optional :: error
    if (present(error)) then
       error = int(code)
       write(0, '(a)') trim(msg)
    else
       write(0, '(a)') trim(msg)
       error stop 1
    end if
  end subroutine raise

  ! IPModel_GAP_Calc (IPModel_GAP.f95:233-605) with the Atoms object unpacked:
  ! pos(3,n), Z(n), lattice(3,3), pbc(3).  Outputs are optional and
  ! caller-allocated; wrong sizes are reported like check_size (:283-298).
  ! The args_str keys of the reference that change what is computed (:324-335) take
  ! their data through optional arguments in place of Atoms properties:
  !   atom_mask_name=NAME        needs atom_mask(n) (the logical property NAME)
  !   only_descriptor=k          coordinate k alone when there are several
  !   energy_per_coordinate=NAME needs energy_per_coordinate(n_coordinates)
  !   local_gap_variance=NAME    needs local_gap_variance(n); gap_variance_gradient(3,n)
  !                              is filled when present and forces or virials are asked for
  !   print_gap_variance=T       prints the per-instance variances (:506-516)
  !   gap_variance_regularisation=x (default 0.001)
  subroutine IPModel_GAPX_Calc(this, n, pos, Z, lattice, pbc, e, local_e, f, virial, local_virial, &
       args_str, error, atom_mask, energy_per_coordinate, local_gap_variance, gap_variance_gradient)
    type(IPModel_GAPX), intent(inout) :: this
    integer, intent(in) :: n
    real(dp), intent(in), target :: pos(3, n), lattice(3, 3)
    integer, intent(in) :: Z(n)
    logical, intent(in) :: pbc(3)
    real(dp), intent(out), optional, target :: e, local_e(:), f(:, :), virial(3, 3), local_virial(:, :)
    character(len=*), intent(in), optional :: args_str
    integer, intent(out), optional :: error
    logical, intent(in), optional :: atom_mask(:)
    real(dp), intent(out), optional, target :: energy_per_coordinate(:), local_gap_variance(:), &
         gap_variance_gradient(:, :)
    integer(c_int), allocatable :: zc(:)
    integer(c_int) :: pbcc(3), want, st, only, ncoord
    real(c_double) :: rc, reg
    real(dp), target :: e_dummy
    character(kind=c_char), allocatable, target :: maskc(:)
    type(c_ptr) :: pe, ple, pf, pv, plv, pmask, pepc, plgv, pgvg
    character(len=256) :: v
    logical :: print_var, grad
    if (present(error)) error = 0
    if (.not. c_associated(this%model)) then
       call raise(1_c_int, 'IPModel_GAPX_Calc: model not initialised', error)
       return
    end if
    want = GAPX_WANT_E
    pe = c_loc(e_dummy)
    ple = c_null_ptr; pf = c_null_ptr; pv = c_null_ptr; plv = c_null_ptr
    pmask = c_null_ptr; pepc = c_null_ptr; plgv = c_null_ptr; pgvg = c_null_ptr
    if (present(e)) pe = c_loc(e)
    if (present(local_e)) then
       if (size(local_e) /= n) then
          call raise(1_c_int, 'IPModel_GAPX_Calc: local_e has the wrong size', error)
          return
       end if
       want = ior(want, GAPX_WANT_LOCAL_E); ple = c_loc(local_e)
    end if
    if (present(f)) then
       if (size(f, 1) /= 3 .or. size(f, 2) /= n) then
          call raise(1_c_int, 'IPModel_GAPX_Calc: force has the wrong shape', error)
          return
       end if
       want = ior(want, GAPX_WANT_F); pf = c_loc(f)
    end if
    if (present(virial)) then
       want = ior(want, GAPX_WANT_V); pv = c_loc(virial)
    end if
    if (present(local_virial)) then
       if (size(local_virial, 1) /= 9 .or. size(local_virial, 2) /= n) then
          call raise(1_c_int, 'IPModel_GAPX_Calc: local_virial has the wrong shape', error)
          return
       end if
       want = ior(want, GAPX_WANT_LOCAL_V); plv = c_loc(local_virial)
    end if
    grad = present(f) .or. present(virial) .or. present(local_virial)
    rc = 0.0_dp
    only = 0
    reg = 1.0e-3_dp
    print_var = .false.
    if (present(args_str)) then
       v = arg_value(args_str, 'cutoff')
       if (len_trim(v) > 0) read(v, *) rc
       if (len_trim(arg_value(args_str, 'E_scale')) > 0 .or. len_trim(arg_value(args_str, 'r_scale')) > 0) then
          call raise(1_c_int, 'IPModel_GAPX_Calc: r_scale / E_scale cannot be set at calc time', error)
          return
       end if
       v = arg_value(args_str, 'atom_mask_name')
       if (len_trim(v) > 0) then
          if (.not. present(atom_mask)) then
             call raise(1_c_int, 'IPModel_GAPX_Calc did not find '//trim(v)//' property (pass atom_mask)', error)
             return
          end if
          if (size(atom_mask) /= n) then
             call raise(1_c_int, 'IPModel_GAPX_Calc: atom_mask has the wrong size', error)
             return
          end if
          allocate(maskc(n))
          maskc = merge(achar(1, c_char), achar(0, c_char), atom_mask)
          pmask = c_loc(maskc)
       end if
       v = arg_value(args_str, 'only_descriptor')
       if (len_trim(v) > 0) read(v, *) only
       v = arg_value(args_str, 'gap_variance_regularisation')
       if (len_trim(v) > 0) read(v, *) reg
       v = arg_value(args_str, 'print_gap_variance')
       print_var = len_trim(v) > 0 .and. (v(1:1) == 'T' .or. v(1:1) == 't' .or. v(1:2) == '.T' .or. v(1:2) == '.t')
       if (print_var) want = ior(want, GAPX_WANT_VARIANCE)
       if (len_trim(arg_value(args_str, 'energy_per_coordinate')) > 0) then
          st = gapx_model_info(this%model, ncoord, -1_c_int, c_null_ptr, c_null_ptr)
          if (.not. present(energy_per_coordinate)) then
             call raise(1_c_int, 'IPModel_GAPX_Calc: energy_per_coordinate needs its output array', error)
             return
          end if
          if (size(energy_per_coordinate) /= ncoord) then
             call raise(1_c_int, 'IPModel_GAPX_Calc: energy_per_coordinate has the wrong size', error)
             return
          end if
          pepc = c_loc(energy_per_coordinate)
       end if
       if (len_trim(arg_value(args_str, 'local_gap_variance')) > 0) then
          if (.not. present(local_gap_variance)) then
             call raise(1_c_int, 'IPModel_GAPX_Calc: local_gap_variance needs its output array', error)
             return
          end if
          if (size(local_gap_variance) /= n) then
             call raise(1_c_int, 'IPModel_GAPX_Calc: local_gap_variance has the wrong size', error)
             return
          end if
          plgv = c_loc(local_gap_variance)
          if (present(gap_variance_gradient) .and. grad) then
             if (size(gap_variance_gradient, 1) /= 3 .or. size(gap_variance_gradient, 2) /= n) then
                call raise(1_c_int, 'IPModel_GAPX_Calc: gap_variance_gradient has the wrong shape', error)
                return
             end if
             pgvg = c_loc(gap_variance_gradient)
          end if
       end if
    end if
    allocate(zc(n))
    zc = Z
    pbcc = merge(1, 0, pbc)
    st = gapx_calc_ex(this%model, int(n, c_int), pos, zc, lattice, pbcc, rc, want, pmask, only, reg, pe, ple, pf, &
         pv, plv, pepc, plgv, pgvg)
    if (st /= 0) then
       call raise(st, 'IPModel_GAPX_Calc: '//trim(IPModel_GAPX_last_error()), error)
       return
    end if
    if (print_var) call print_variances(this, only)
  end subroutine IPModel_GAPX_Calc

  ! print_gap_variance (IPModel_GAP.f95:401-404, 506-516)
  subroutine print_variances(this, only)
    type(IPModel_GAPX), intent(in) :: this
    integer(c_int), intent(in) :: only
    integer(c_int) :: ncoord, k, ne, st
    real(c_double), allocatable, target :: var(:), c2(:)
    integer(c_int), allocatable, target :: at2(:, :)
    integer :: i
    st = gapx_model_info(this%model, ncoord, -1_c_int, c_null_ptr, c_null_ptr)
    write(*, '(a,a,a,i0,a)') 'GAP_VARIANCE potential ', trim(this%label), ' calculating for ', ncoord, ' descriptors'
    do k = 0, ncoord - 1
       if (only > 0 .and. ncoord > 1 .and. k /= only - 1) cycle
       st = gapx_descriptor_variances(this%model, k, 0_c_int, c_null_ptr, c_null_ptr, c_null_ptr, ne)
       if (st /= 0) return
       if (ne == 0) then
          write(*, '(a,a,a,i0,a)') 'GAP_VARIANCE potential ', trim(this%label), ' descriptor ', k + 1, ' not found'
          cycle
       end if
       allocate(var(ne), c2(ne), at2(2, ne))
       st = gapx_descriptor_variances(this%model, k, ne, c_loc(var), c_loc(c2), c_loc(at2), ne)
       do i = 1, ne
          write(*, '(a,a,a,i0,a,i0,a,es25.16,a,es25.16,a)') 'GAP_VARIANCE potential ', trim(this%label), &
               ' descriptor ', k + 1, ' var( ', i, ' ) = ', var(i), ' * ', c2(i), ' cutoff'
          if (at2(2, i) >= 0) then
             write(*, '(a,a,a,i0,a,i0,a,2i8)') 'GAP_VARIANCE potential ', trim(this%label), ' descriptor ', k + 1, &
                  ' ii( ', i, ' ) = ', at2(1, i) + 1, at2(2, i) + 1
          else
             write(*, '(a,a,a,i0,a,i0,a,i8)') 'GAP_VARIANCE potential ', trim(this%label), ' descriptor ', k + 1, &
                  ' ii( ', i, ' ) = ', at2(1, i) + 1
          end if
       end do
       deallocate(var, c2, at2)
    end do
  end subroutine print_variances

end module IPModel_GAPX_module
