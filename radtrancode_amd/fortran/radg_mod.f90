! ISO_C_BINDING interfaces to the radg C ABI (include/radg.h).
!
! A NEMESIS build keeps calling
!     CALL cirsradg_wave(dist,inormal,iray,delh,nlayer,npath,ngas,press,...)
! through its implicit interface: the library exports the link name
! cirsradg_wave_ with the same 37 by-reference arguments.  This module adds
! the entry points with no Fortran counterpart: registering the k-tables, CIA
! table and dust cross-sections that the reference keeps in /INTERPK/,
! /CIATABLE/ and /DUST/, choosing the arithmetic form, the whole-spectrum call
! and the status query.
module radg
  use iso_c_binding
  implicit none
  private
  integer(c_int), parameter, public :: RADG_MODE_FP64 = 0, RADG_MODE_MIXED = 1, RADG_MODE_BF16 = 2, RADG_MODE_F16 = 3
  integer(c_int), parameter, public :: RADG_OK = 0

  public :: radg_set_mode, radg_read_klist, radg_register_cia, radg_cirsradg_status
  public :: radg_register_dust, radg_set_nlte, radg_set_iform, radg_register_parah2
  public :: radg_register_e3, radg_register_n2nir, radg_register_solar, radg_register_gascon
  public :: radg_spectrum_r4, radg_error_message, radg_read_klist_f
  public :: radg_batch_create, radg_batch_launch, radg_batch_sync, radg_batch_read
  public :: radg_batch_destroy, radg_batch_set_statemaps, radg_batch_read_mapped
  public :: radg_batch_device_outputs
  public :: radg_oe_create, radg_oe_cost, radg_oe_propose, radg_oe_destroy
  public :: radg_cirsconv, radg_read_cia, radg_read_cia_f
  public :: radg_read_klbllist, radg_read_klbllist_f, radg_set_ilbl

  ! radg_request of include/radg.h: sizes, then C addresses (c_loc) of the
  ! caller's arrays in the compact 0-based layout the header documents.
  type, bind(C), public :: radg_request
    integer(c_int) :: nlayer, npath, ngas, nsw, nwave, nem
    integer(c_int) :: inormal, iray, ispace, itype, flagh2p
    real(c_double) :: tsurf
    type(c_ptr) :: press, temp, delh, pp, amount
    type(c_ptr) :: idgas, isogas, nlayin, imod, layinc
    type(c_ptr) :: scale, emtemp
    type(c_ptr) :: isw
    type(c_ptr) :: vwave, vem, emissivity, path_weight
    integer(c_int) :: ncont, nsec            ! dust: 0 = none
    type(c_ptr) :: cont, vsec, xsec          ! cont(ncont,nlayer), vsec(nsec), xsec(nsec,ncont)
    integer(c_int) :: inlte                  ! /NLTE_FLAGS/ of calc_nlte_t.f
    real(c_double) :: nlte_k1, nlte_a
    integer(c_int) :: iform                  ! /SOLARDAT/ iform
    real(c_double) :: radius1, radextra      ! km
    type(c_ptr) :: hfp                       ! hfp(nlayer) when flagh2p = 1
    real(c_double) :: dist                   ! AU from the star (iform 4)
  end type

  ! radg_gascon of include/radg.h: the /GASCONDEF/ flags and the tables
  ! ngascon.f reads; a NEMESIS caller passes c_loc of its own /CH2OS0/
  ! H2OS96, /CH2OS1/ H2OS60 and /CH2OF0/ H2OF96 arrays (BLOCK DATA of
  ! ngascon.f) and of the karkoschka.dat / lutzowen.dat columns
  type, bind(C), public :: radg_gascon
    integer(c_int) :: ih2o, ich4, io3, inh3, iuvscat
    integer(c_int) :: nh2o
    real(c_double) :: h2o_v0, h2o_dv, h2o_c2, h2o_ts, h2o_ps
    type(c_ptr) :: h2o_s296, h2o_s260, h2o_f296
    integer(c_int) :: nch4, nnh3
    type(c_ptr) :: ch4_wl, ch4_k, nh3_wl, nh3_k
  end type

  ! radg_conv of include/radg.h: cirsconv.f's fwhm, FWHMEXIST table and
  ! (fwhm < 0) the .fil filter functions
  type, bind(C), public :: radg_conv
    real(c_double) :: fwhm
    integer(c_int) :: nfwhm
    type(c_ptr) :: vfwhm, xfwhm
    integer(c_int) :: nfil
    type(c_ptr) :: fil_n, fil_vcen, fil_v, fil_f
  end type

  ! radg_statemap: the /LGRAD/ DAM, DTE and DCO and map2xvec's xmap of one
  ! request (DAM(nlayer,npro) in C order is dam(npro,nlayer) in Fortran order,
  ! and xmap(nv,nparam,npro) in C order is xmap(npro,nparam,nv)); dco may be
  ! c_null_ptr for a request without dust.
  type, bind(C), public :: radg_statemap
    integer(c_int) :: nv, npro
    type(c_ptr) :: dam, dte, xmap
    type(c_ptr) :: dco
    type(c_ptr) :: dfp                       ! /LGRAD/ DFP when flagh2p = 1
  end type

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

    ! /INTERPKLBL/ from the .lta files an .lls list names (read_klbllist.f)
    integer(c_int) function radg_read_klbllist(path, recl) bind(C, name="radg_read_klbllist")
      import :: c_int, c_char
      character(kind=c_char), dimension(*) :: path
      integer(c_int), value :: recl
    end function

    ! /LBLTABLE/ ILBL: 0 k-tables, 2 line-by-line tables
    integer(c_int) function radg_set_ilbl(ilbl) bind(C, name="radg_set_ilbl")
      import :: c_int
      integer(c_int), value :: ilbl
    end function

    integer(c_int) function radg_read_cia(path, dnu) bind(C, name="radg_read_cia")
      import :: c_int, c_char, c_double
      character(kind=c_char), dimension(*) :: path
      real(c_double), value :: dnu
    end function

    integer(c_int) function radg_register_cia(dnu, temps, kcia) bind(C, name="radg_register_cia")
      import :: c_int, c_double
      real(c_double), value :: dnu
      real(c_double) :: temps(25), kcia(9, 25, 1501)
    end function

    ! /DUST/: xsec(nsec,ncont) extinction cross-sections (cm2) on vsec(nsec);
    ! ncont = 0 clears it
    integer(c_int) function radg_register_dust(nsec, ncont, vsec, xsec) &
        bind(C, name="radg_register_dust")
      import :: c_int, c_double
      integer(c_int), value :: nsec, ncont
      real(c_double) :: vsec(*), xsec(*)
    end function

    ! /FPTABLE/: kh2h2/kh2he(npara,25,1501), kcia5(5,25,1501); npara = 0 clears
    integer(c_int) function radg_register_parah2(dnu, temps, npara, fp, kh2h2, kh2he, kcia5) &
        bind(C, name="radg_register_parah2")
      import :: c_int, c_double
      real(c_double), value :: dnu
      integer(c_int), value :: npara
      real(c_double) :: temps(25), fp(*), kh2h2(*), kh2he(*), kcia5(*)
    end function

    ! /SOLARDAT/ stellar spectrum (iread = 999): npt wavelengths and powers; 0 clears
    integer(c_int) function radg_register_solar(npt, wave, rad) bind(C, name="radg_register_solar")
      import :: c_int, c_double
      integer(c_int), value :: npt
      real(c_double) :: wave(*), rad(*)
    end function

    ! N2's near-infrared CIA: n2n2cia/n2h2cia values at their knots; 0, 0 clears
    integer(c_int) function radg_register_n2nir(n1, wn1, k1, n2, wn2, k2) &
        bind(C, name="radg_register_n2nir")
      import :: c_int, c_double
      integer(c_int), value :: n1, n2
      real(c_double) :: wn1(*), k1(*), wn2(*), k2(*)
    end function

    ! /E3TABLE/ (imod 18): n = 1001 transmissions and E3 values; n = 0 clears
    integer(c_int) function radg_register_e3(n, trans, e3) bind(C, name="radg_register_e3")
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double) :: trans(*), e3(*)
    end function

    ! /GASCONDEF/ and ngascon.f's tables (every flag 0 and no table: none)
    integer(c_int) function radg_register_gascon(g) bind(C, name="radg_register_gascon")
      import :: c_int, radg_gascon
      type(radg_gascon), intent(in) :: g
    end function

    ! cirsconv.f for ncol spectra at once: y(ncol,nwave) -> yout(ncol,nconv),
    ! both device arrays (c_loc of memory the caller holds on the GPU, e.g. a
    ! batch's radg_batch_device_outputs)
    integer(c_int) function radg_cirsconv(nwave, vwave, nconv, vconv, conv, ncol, y, yout, &
        stream) bind(C, name="radg_cirsconv")
      import :: c_int, c_double, c_ptr, radg_conv
      integer(c_int), value :: nwave, nconv, ncol
      real(c_double) :: vwave(*), vconv(*)
      type(radg_conv), intent(in) :: conv
      type(c_ptr), value :: y, yout, stream
    end function

    ! /SOLARDAT/: iform, iread (iread = 999 reads the registered spectrum)
    integer(c_int) function radg_set_iform(iform, iread) bind(C, name="radg_set_iform")
      import :: c_int
      integer(c_int), value :: iform, iread
    end function

    ! /NLTE_FLAGS/: inlte_flag, nlte_k1, nlte_A
    integer(c_int) function radg_set_nlte(flag, k1, a) bind(C, name="radg_set_nlte")
      import :: c_int, c_double
      integer(c_int), value :: flag
      real(c_double), value :: k1, a
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
        press, temp, pp, amount, ispace, nlayin, layinc, cont, scale, imod, idgas, isogas, &
        emtemp, itype, nem, vem, emissivity, tsurf, radius1, radextra, flagh2p, hfp, nsw, isw, &
        nwave, vwave, output, gtsurf, doutputdq) bind(C, name="radg_spectrum_r4")
      import :: c_int, c_float
      real(c_float) :: dist, delh(*), press(*), temp(*), pp(*), amount(*), scale(*), cont(*)
      real(c_float) :: hfp(*)
      real(c_float) :: emtemp(*), vem(*), emissivity(*), tsurf, vwave(*), radius1, radextra
      real(c_float) :: output(*), gtsurf(*), doutputdq(*)
      integer(c_int) :: inormal, iray, nlayer, npath, ngas, ispace, nlayin(*), layinc(*)
      integer(c_int) :: imod(*), idgas(*), isogas(*), itype, nem, flagh2p, nsw, isw(*)
      integer(c_int), value :: nwave
    end function

    ! ---- device-resident batches (include/radg.h)
    integer(c_int) function radg_batch_create(nreq, reqs, mode, budget, b) &
        bind(C, name="radg_batch_create")
      import :: c_int, c_size_t, c_ptr, radg_request
      integer(c_int), value :: nreq, mode
      type(radg_request) :: reqs(*)
      integer(c_size_t), value :: budget
      type(c_ptr) :: b
    end function

    integer(c_int) function radg_batch_launch(b, stream) bind(C, name="radg_batch_launch")
      import :: c_int, c_ptr
      type(c_ptr), value :: b, stream
    end function

    integer(c_int) function radg_batch_sync(b) bind(C, name="radg_batch_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function

    integer(c_int) function radg_batch_read(b, r, output, gtsurf, doutputdq, wout, wgts, wjac) &
        bind(C, name="radg_batch_read")
      import :: c_int, c_ptr
      type(c_ptr), value :: b, output, gtsurf, doutputdq, wout, wgts, wjac
      integer(c_int), value :: r
    end function

    integer(c_int) function radg_batch_destroy(b) bind(C, name="radg_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function

    ! ---- map2pro + map2xvec on the device (cirsrtfg_wave.f:403-425)
    integer(c_int) function radg_batch_set_statemaps(b, maps, per_path) &
        bind(C, name="radg_batch_set_statemaps")
      import :: c_int, c_ptr, radg_statemap
      type(c_ptr), value :: b
      type(radg_statemap) :: maps(*)
      integer(c_int), value :: per_path
    end function

    integer(c_int) function radg_batch_read_mapped(b, r, doutdx, kk) &
        bind(C, name="radg_batch_read_mapped")
      import :: c_int, c_ptr
      type(c_ptr), value :: b, doutdx, kk
      integer(c_int), value :: r
    end function

    integer(c_int) function radg_batch_device_outputs(b, r, wout, kk) &
        bind(C, name="radg_batch_device_outputs")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: r
      type(c_ptr) :: wout, kk
    end function

    ! ---- optimal estimation for many retrievals (coreret.f:554-1068).
    ! Fortran order: y(ny,nret), se(ny,nret), xa(nx,nret), sa(nx,nx[,nret]),
    ! xn(nx,nret); yn_dev and kk_dev hold device addresses per retrieval
    ! (radg_batch_device_outputs).
    integer(c_int) function radg_oe_create(nret, nx, ny, y, se, xa, sa, sa_shared, h) &
        bind(C, name="radg_oe_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nret, nx, ny, sa_shared
      real(c_double) :: y(*), se(*), xa(*), sa(*)
      type(c_ptr) :: h
    end function

    integer(c_int) function radg_oe_cost(h, xn, yn_dev, phi, chisq) bind(C, name="radg_oe_cost")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double) :: xn(*), phi(*), chisq(*)
      type(c_ptr) :: yn_dev(*)
    end function

    integer(c_int) function radg_oe_propose(h, xn, yn_dev, kk_dev, lambda, xn1) &
        bind(C, name="radg_oe_propose")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double) :: xn(*), lambda(*), xn1(*)
      type(c_ptr) :: yn_dev(*), kk_dev(*)
    end function

    integer(c_int) function radg_oe_destroy(h) bind(C, name="radg_oe_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
  end interface

contains

  ! Register the .kta files listed in a .kls file (read_klist.f).
  integer function radg_read_klist_f(path, recl)
    character(len=*), intent(in) :: path
    integer, intent(in) :: recl
    radg_read_klist_f = radg_read_klist(trim(path) // c_null_char, int(recl, c_int))
  end function

  ! Register the .lta files listed in an .lls file (read_klbllist.f).
  integer function radg_read_klbllist_f(path, recl)
    character(len=*), intent(in) :: path
    integer, intent(in) :: recl
    radg_read_klbllist_f = radg_read_klbllist(trim(path) // c_null_char, int(recl, c_int))
  end function

  integer function radg_read_cia_f(path, dnu)
    character(len=*), intent(in) :: path
    real, intent(in) :: dnu
    radg_read_cia_f = radg_read_cia(trim(path) // c_null_char, real(dnu, c_double))
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
