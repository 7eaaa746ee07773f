! ISO_C_BINDING interfaces to the radg C ABI (include/radg.h).
!
! A NEMESIS build keeps calling
!     CALL cirsradg_wave(dist,inormal,iray,delh,nlayer,npath,ngas,press,...)
! through its implicit interface: the library exports the link name
! cirsradg_wave_ with the same 37 by-reference arguments.  This module adds
! the entry points with no Fortran counterpart: registering the k-tables and
! CIA table that the reference keeps in /INTERPK/ and /CIATABLE/, choosing the
! arithmetic form, the whole-spectrum call and the status query.
module radg
  use iso_c_binding
  implicit none
  private
  integer(c_int), parameter, public :: RADG_MODE_FP64 = 0, RADG_MODE_MIXED = 1
  integer(c_int), parameter, public :: RADG_OK = 0

  public :: radg_set_mode, radg_read_klist, radg_register_cia, radg_cirsradg_status
  public :: radg_spectrum_r4, radg_error_message, radg_read_klist_f

  interface
    integer(c_int) function radg_set_mode(mode) bind(C, name="radg_set_mode")
      import :: c_int
      integer(c_int), value :: mode
    end function

    integer(c_int) function radg_read_klist(path, recl) bind(C, name="radg_read_klist")
      import :: c_int, c_char
      character(kind=c_char), dimension(*) :: path
      integer(c_int), value :: recl
    end function

    integer(c_int) function radg_register_cia(dnu, temps, kcia) bind(C, name="radg_register_cia")
      import :: c_int, c_double
      real(c_double), value :: dnu
      real(c_double) :: temps(25), kcia(9, 25, 1501)
    end function

    integer(c_int) function radg_cirsradg_status() bind(C, name="radg_cirsradg_status")
      import :: c_int
    end function

    type(c_ptr) function radg_last_error_c() bind(C, name="radg_last_error")
      import :: c_ptr
    end function

    ! The wavenumber loop of cirsrtfg_wave.f:361-448 in one call; outputs
    ! output(npath,nwave) ... in Fortran order: see include/radg.h.
    integer(c_int) function radg_spectrum_r4(dist, inormal, iray, delh, nlayer, npath, ngas, &
        press, temp, pp, amount, ispace, nlayin, layinc, scale, imod, idgas, isogas, emtemp, &
        itype, nem, vem, emissivity, tsurf, flagh2p, nsw, isw, nwave, vwave, output, gtsurf, &
        doutputdq) bind(C, name="radg_spectrum_r4")
      import :: c_int, c_float
      real(c_float) :: dist, delh(*), press(*), temp(*), pp(*), amount(*), scale(*)
      real(c_float) :: emtemp(*), vem(*), emissivity(*), tsurf, vwave(*)
      real(c_float) :: output(*), gtsurf(*), doutputdq(*)
      integer(c_int) :: inormal, iray, nlayer, npath, ngas, ispace, nlayin(*), layinc(*)
      integer(c_int) :: imod(*), idgas(*), isogas(*), itype, nem, flagh2p, nsw, isw(*)
      integer(c_int), value :: nwave
    end function
  end interface

contains

  ! Register the .kta files listed in a .kls file (read_klist.f).
  integer function radg_read_klist_f(path, recl)
    character(len=*), intent(in) :: path
    integer, intent(in) :: recl
    radg_read_klist_f = radg_read_klist(trim(path) // c_null_char, int(recl, c_int))
  end function

  function radg_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = radg_last_error_c()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [512])
    do i = 1, 512
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    end do
  end function
end module radg
