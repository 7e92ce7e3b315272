! Drop-in check of the Fortran wa_module over libsht_hip.so (runs on a GPU).
! T21 on 64 x 32 (dcpam_APEI07Couple_T21L16.conf:56-66), 16 levels.  Closed forms
! of the real orthonormal harmonics and their gradients, a round trip of
! random spectra, the global mean, and the SPML index helpers.  Prints PASS
! and stops 0 when every check holds.
program test_wa_module
  use wa_module
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  integer, parameter :: nm = 21, im = 64, jm = 32, km = 16
  real(dp) :: xya(0:im-1, 1:jm, km), g(0:im-1, 1:jm, km)
  real(dp) :: wa((nm+1)**2, km), wa2((nm+1)**2, km)
  real(dp) :: mu(1:jm), cl(1:jm), err, s3
  integer :: i, j, k, nbad
  integer :: pair(2)

  nbad = 0
  call wa_Initial(nm, im, jm, km)
  s3 = sqrt(3.0_dp)
  mu = sin(y_Lat)
  cl = cos(y_Lat)
  if (abs(sum(y_Lat_Weight) - 2.0_dp) > 1e-14_dp) call bad('sum of latitude weights')
  if (any(y_Lat(2:jm) <= y_Lat(1:jm-1))) call bad('latitudes not south to north')

  ! layer 1: sqrt3 cos(lat) cos(lon) = real harmonic (1, 1); layer 2: sqrt3 sin(lat) = (1, 0);
  ! layer 3: sqrt3 cos(lat) sin(lon) = (1, -1); layer 4: 1 = (0, 0); others: zero
  xya = 0.0_dp
  do j = 1, jm
    do i = 0, im - 1
      xya(i, j, 1) = s3 * cl(j) * cos(x_Lon(i))
      xya(i, j, 2) = s3 * mu(j)
      xya(i, j, 3) = s3 * cl(j) * sin(x_Lon(i))
      xya(i, j, 4) = 1.0_dp
    end do
  end do
  wa = wa_xya(xya)
  wa2 = 0.0_dp
  wa2(l_nm(1, 1), 1) = 1.0_dp
  wa2(l_nm(1, 0), 2) = 1.0_dp
  wa2(l_nm(1, -1), 3) = 1.0_dp
  wa2(l_nm(0, 0), 4) = 1.0_dp
  err = maxval(abs(wa - wa2))
  if (err > 1e-13_dp) call bad('wa_xya of closed-form harmonics')

  ! gradients of layer 1: (1/cos) d/dlon = -sqrt3 sin(lon); d/dlat = -sqrt3 sin(lat) cos(lon)
  g = xya_GradLon_wa(wa)
  err = 0.0_dp
  do j = 1, jm
    do i = 0, im - 1
      err = max(err, abs(g(i, j, 1) + s3 * sin(x_Lon(i))))
    end do
  end do
  g = xya_GradLat_wa(wa)
  do j = 1, jm
    do i = 0, im - 1
      err = max(err, abs(g(i, j, 1) + s3 * mu(j) * cos(x_Lon(i))))
      err = max(err, abs(g(i, j, 2) - s3 * cl(j)))
    end do
  end do
  if (err > 1e-12_dp) call bad('gradients')

  ! round trip of arbitrary spectra
  do k = 1, km
    do i = 1, (nm+1)**2
      pair = nm_l(i)
      wa2(i, k) = sin(0.37_dp * i + 1.3_dp * k) / (1.0_dp + pair(1))**1.5_dp
    end do
  end do
  xya = xya_wa(wa2)
  wa = wa_xya(xya)
  err = maxval(abs(wa - wa2)) / maxval(abs(wa2))
  if (err > 1e-12_dp) call bad('round trip')
  if (abs(AvrLonLat_xy(xya(:, :, 5)) - wa2(l_nm(0, 0), 5)) > 1e-13_dp) call bad('global mean')

  ! flux divergence: DivLat(cos lat) = -2 sin lat = -(2/sqrt3) harmonic (1, 0);
  ! div(grad f) = Laplacian f; LaplaInv(Lapla f) = f without its mean
  do k = 1, km
    do j = 1, jm
      xya(:, j, k) = cl(j)
    end do
  end do
  wa = wa_DivLat_xya(xya)
  wa2 = 0.0_dp
  wa2(l_nm(1, 0), :) = -2.0_dp / s3
  err = maxval(abs(wa - wa2))
  if (err > 1e-13_dp) call bad('wa_DivLat_xya closed form')
  do k = 1, km
    do i = 1, (nm+1)**2
      pair = nm_l(i)
      wa2(i, k) = cos(0.21_dp * i + 0.7_dp * k) / (1.0_dp + pair(1))**1.5_dp
    end do
  end do
  wa = wa_DivLon_xya(xya_GradLon_wa(wa2)) + wa_DivLat_xya(xya_GradLat_wa(wa2))
  err = maxval(abs(wa - wa_Lapla_wa(wa2))) / maxval(abs(wa_Lapla_wa(wa2)))
  if (err > 1e-12_dp) call bad('div grad = Laplacian')
  wa = wa_LaplaInv_wa(wa_Lapla_wa(wa2))
  wa2(l_nm(0, 0), :) = 0.0_dp
  err = maxval(abs(wa - wa2))
  if (err > 1e-13_dp) call bad('LaplaInv(Lapla)')

  call wa_Finalize()
  if (nbad == 0) then
    print '(a)', 'PASS wa_module drop-in (T21, 64x32, 16 levels; transforms, gradients, divergence, Laplacian)'
  else
    print '(a, i0, a)', 'FAIL: ', nbad, ' checks'
    error stop 1
  end if

contains

  subroutine bad(what)
    character(*), intent(in) :: what
    print '(a, a, a, es10.3)', 'check failed: ', what, ' err=', err
    nbad = nbad + 1
  end subroutine bad

end program test_wa_module
