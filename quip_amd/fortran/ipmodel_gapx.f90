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
       GAPX_WANT_LOCAL_E = 8, GAPX_WANT_LOCAL_V = 16

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
  end interface
end module gapx_c_binding


module IPModel_GAPX_module
  use iso_c_binding
  use gapx_c_binding
  implicit none
  private

  integer, parameter :: dp = kind(1.0d0)

  public :: IPModel_GAPX, IPModel_GAPX_Initialise_str, IPModel_GAPX_Finalise, IPModel_GAPX_Calc, &
       IPModel_GAPX_Print, IPModel_GAPX_cutoff, IPModel_GAPX_last_error

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
  subroutine IPModel_GAPX_Calc(this, n, pos, Z, lattice, pbc, e, local_e, f, virial, local_virial, &
       args_str, error)
    type(IPModel_GAPX), intent(inout) :: this
    integer, intent(in) :: n
    real(dp), intent(in), target :: pos(3, n), lattice(3, 3)
    integer, intent(in) :: Z(n)
    logical, intent(in) :: pbc(3)
    real(dp), intent(out), optional, target :: e, local_e(:), f(:, :), virial(3, 3), local_virial(:, :)
    character(len=*), intent(in), optional :: args_str
    integer, intent(out), optional :: error
    integer(c_int), allocatable :: zc(:)
    integer(c_int) :: pbcc(3), want, st
    real(c_double) :: rc
    real(dp), target :: e_dummy
    type(c_ptr) :: pe, ple, pf, pv, plv
    character(len=256) :: v
    if (present(error)) error = 0
    if (.not. c_associated(this%model)) then
       call raise(1_c_int, 'IPModel_GAPX_Calc: model not initialised', error)
       return
    end if
    want = GAPX_WANT_E
    pe = c_loc(e_dummy)
    ple = c_null_ptr; pf = c_null_ptr; pv = c_null_ptr; plv = c_null_ptr
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
    rc = 0.0_dp
    if (present(args_str)) then
       v = arg_value(args_str, 'cutoff')
       if (len_trim(v) > 0) read(v, *) rc
       if (len_trim(arg_value(args_str, 'E_scale')) > 0 .or. len_trim(arg_value(args_str, 'r_scale')) > 0) then
          call raise(1_c_int, 'IPModel_GAPX_Calc: r_scale / E_scale cannot be set at calc time', error)
          return
       end if
    end if
    allocate(zc(n))
    zc = Z
    pbcc = merge(1, 0, pbc)
    st = gapx_calc(this%model, int(n, c_int), pos, zc, lattice, pbcc, rc, want, pe, ple, pf, pv, plv)
    if (st /= 0) call raise(st, 'IPModel_GAPX_Calc: '//trim(IPModel_GAPX_last_error()), error)
  end subroutine IPModel_GAPX_Calc

end module IPModel_GAPX_module
