! A Fortran host for the drop-in: it calls cirsradg_wave once per
! wavenumber with the reference's argument list and fixed leading dimensions,
! exactly as radtran/cirsradg/cirsrtfg_wave.f:397-401 does, then the
! whole-spectrum entry radg_spectrum_r4 once.
!
! Usage: radg_host <input.bin> <output.bin>
! The input file (stream access) holds the k-list path, the CIA table and
! the argument arrays; the output file holds output/gtsurf/doutputdq per
! wavenumber from both calls.
program radg_host
  use radg
  use iso_c_binding
  implicit none
  integer, parameter :: maxlay = 400, maxpat = 200, maxgas = 30, maxcon = 10
  integer, parameter :: maxsec = 16000, nparmax = maxgas + 2 + maxcon
  character(len=512) :: fin, fout, klist
  integer(c_int) :: klen, recl, nlayer, npath, ngas, nwave, nem, nsw, inormal, iray, ispace
  integer(c_int) :: itype, flagh2p, amform, iwave, nparam, rc, mode, nsec, ncont, inlte, iform
  real(c_double) :: nlte_k1, nlte_a, pdnu, ptemps(25)
  integer(c_int) :: npara, ne3, nn1, nn2
  real(c_double), allocatable :: pfp(:), pkhh(:), pkhe(:), pk5(:)
  real(c_double) :: e3t(1001), e3v(1001)
  real(c_double), allocatable :: wn1(:), k1(:), wn2(:), k2(:)
  real(c_double), allocatable :: vsec(:), xsec(:)
  real(c_double) :: dnu, ctemps(25)
  real(c_double), allocatable :: kcia(:, :, :)
  real(c_float) :: dist, tsurf, radius1, radextra, vv
  real(c_float), allocatable :: delh(:), vwave(:)
  real(c_float) :: press(maxlay), temp(maxlay), hfp(maxlay)
  real(c_float) :: pp(maxlay, maxgas), amount(maxlay, maxgas), cont(maxcon, maxlay)
  real(c_float) :: scale(maxlay, maxpat), emtemp(maxlay, maxpat)
  real(c_float) :: vem(maxsec), emissivity(maxsec)
  real(c_float) :: output(maxpat), gtsurf(maxpat)
  real(c_float), allocatable :: doutputdq(:, :, :), sout(:), sgts(:), sdq(:)
  integer(c_int) :: nlayin(maxpat), layinc(maxlay, maxpat), imod(maxpat)
  integer(c_int) :: idgas(maxgas), isogas(maxgas), isw(nparmax)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (10, file=trim(fin), access='stream', form='unformatted', status='old')
  read (10) klen
  klist = ' '
  read (10) klist(1:klen)
  read (10) recl, mode
  allocate (kcia(9, 25, 1501))
  read (10) dnu, ctemps, kcia
  read (10) nlayer, npath, ngas, nwave, nem, nsw, inormal, iray, ispace, itype, flagh2p, amform
  read (10) dist, tsurf, radius1, radextra
  allocate (delh(nlayer), vwave(nwave))
  read (10) delh, press, temp, pp, amount, nlayin, layinc, cont, scale, imod, idgas, isogas
  read (10) emtemp, vem, emissivity, hfp, isw, vwave
  ! /DUST/: nsec, ncont, then vsec(nsec) and xsec(nsec,ncont) when ncont > 0
  read (10) nsec, ncont
  allocate (vsec(max(nsec, 1)), xsec(max(nsec * ncont, 1)))
  if (ncont > 0) read (10) vsec(1:nsec), xsec(1:nsec * ncont)
  ! /NLTE_FLAGS/
  read (10) inlte, nlte_k1, nlte_a
  ! /SOLARDAT/ iform
  read (10) iform
  ! /FPTABLE/: npara (0 = none), then dnu, temps(25), fp(npara), kh2h2, kh2he, kcia5
  read (10) npara
  if (npara > 0) then
    allocate (pfp(npara), pkhh(npara * 25 * 1501), pkhe(npara * 25 * 1501), pk5(5 * 25 * 1501))
    read (10) pdnu, ptemps, pfp, pkhh, pkhe, pk5
  end if
  ! /E3TABLE/: ne3 (0 = none or 1001), then trans(1001), e3(1001)
  read (10) ne3
  if (ne3 > 0) read (10) e3t, e3v
  ! N2's near-infrared CIA: nn1, nn2 (0, 0 = none), then the two knot tables
  read (10) nn1, nn2
  allocate (wn1(max(nn1, 1)), k1(max(nn1, 1)), wn2(max(nn2, 1)), k2(max(nn2, 1)))
  if (nn1 > 0) read (10) wn1, k1, wn2, k2
  close (10)

  if (radg_read_klist_f(klist(1:klen), recl) /= RADG_OK) call die()
  if (radg_register_cia(dnu, ctemps, kcia) /= RADG_OK) call die()
  if (radg_set_mode(mode) /= RADG_OK) call die()
  if (radg_register_dust(nsec, ncont, vsec, xsec) /= RADG_OK) call die()
  if (radg_set_nlte(inlte, nlte_k1, nlte_a) /= RADG_OK) call die()
  if (radg_set_iform(iform, 0) /= RADG_OK) call die()
  if (npara > 0) then
    if (radg_register_parah2(pdnu, ptemps, npara, pfp, pkhh, pkhe, pk5) /= RADG_OK) call die()
  end if
  if (ne3 > 0) then
    if (radg_register_e3(ne3, e3t, e3v) /= RADG_OK) call die()
  end if
  if (nn1 > 0) then
    if (radg_register_n2nir(nn1, wn1, k1, nn2, wn2, k2) /= RADG_OK) call die()
  end if

  nparam = ngas + 1 + ncont
  if (flagh2p == 1) nparam = nparam + 1
  allocate (doutputdq(maxpat, maxlay, nparmax))
  open (11, file=trim(fout), access='stream', form='unformatted', status='replace')
  do iwave = 1, nwave
    vv = vwave(iwave)
    call cirsradg_wave(dist, inormal, iray, delh, nlayer, npath, ngas, press, temp, pp, amount, &
                       iwave, ispace, amform, vv, nlayin, layinc, cont, scale, imod, idgas, &
                       isogas, emtemp, itype, nem, vem, emissivity, tsurf, gtsurf, radius1, &
                       flagh2p, hfp, radextra, nsw, isw, output, doutputdq)
    if (radg_cirsradg_status() /= RADG_OK) call die()
    write (11) output(1:npath), gtsurf(1:npath), doutputdq(1:npath, 1:nlayer, 1:nparam)
  end do
  allocate (sout(nwave * npath), sgts(nwave * npath), sdq(nwave * npath * nlayer * nparam))
  rc = radg_spectrum_r4(dist, inormal, iray, delh, nlayer, npath, ngas, press, temp, pp, amount, &
                        ispace, nlayin, layinc, cont, scale, imod, idgas, isogas, emtemp, itype, nem, &
                        vem, emissivity, tsurf, radius1, radextra, flagh2p, hfp, nsw, isw, nwave, vwave, &
                        sout, sgts, sdq)
  if (rc /= RADG_OK) call die()
  write (11) sout, sgts, sdq
  close (11)

contains

  subroutine die()
    write (0, '(a)') 'radg error: ' // trim(radg_error_message())
    stop 1
  end subroutine
end program radg_host
