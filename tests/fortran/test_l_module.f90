! Drop-in check of the Fortran l_module over libsht_hip.so (runs on a GPU), at
! the energy-balance model's own size: l_initial(1024, 1024)
! (exp/APEI07Couple/ebm/EBM_main.f90:5-8, :64).  With no ice (x_s = 1) the
! absorbed insolation H = S(mu) * coalbedo(mu) is a degree-4 polynomial, so
! sqrt(2n+1) l_y(H) equals calc_Hn_exact's closed forms (EBM_main.f90:250-299)
! and vanishes elsewhere; y_l(l_y(H)) returns H.  Prints PASS and stops 0.
program test_l_module
  use l_module
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  integer, parameter :: jm = 1024, nm = jm, lm = nm + 1
  real(dp), parameter :: C0 = 0.697_dp, C2 = -0.0779_dp, CI = 0.38_dp, SR2 = -0.482_dp
  real(dp) :: x(jm), P2(jm), H(jm), l_H(lm), lc(lm), y2(jm, 3), l2(0:nm, 3), err, xs
  real(dp) :: G002, G022, G222, G004, G024, G224, e1, e3, e5
  integer :: l, nbad

  nbad = 0
  call l_Initial(nm, jm)
  if (abs(sum(y_Lat_Weight) - 2.0_dp) > 1e-13_dp) call bad('sum of latitude weights', 1.0_dp)
  if (any(y_Lat(2:jm) <= y_Lat(1:jm-1))) call bad('latitudes not south to north', 1.0_dp)
  x = sin(y_Lat)
  P2 = (3.0_dp * x**2 - 1.0_dp) * 0.5_dp
  H = (1.0_dp + SR2 * P2) * (C0 + C2 * P2)
  lc = l_y(H)
  do l = 1, lm
    l_H(l) = sqrt(2.0_dp * (l - 1) + 1.0_dp) * lc(l)
  end do
  xs = 1.0_dp
  G002 = 0.5_dp * (xs**3 - xs)
  G022 = 0.25_dp * (9.0_dp / 5.0_dp * xs**5 - 2.0_dp * xs**3 + xs)
  G222 = 0.125_dp * (-xs + 3.0_dp * xs**3 - 0.2_dp * (27.0_dp * xs**5) + (27.0_dp * xs**7) / 7.0_dp)
  G004 = 0.125_dp * (7.0_dp * xs**5 - 10.0_dp * xs**3 + 3.0_dp * xs)
  G024 = 0.0625_dp * (15.0_dp * xs**7 - 25.0_dp * xs**5 + 13.0_dp * xs**3 - 3.0_dp * xs)
  G224 = xs * (35.0_dp / 32.0_dp * xs**8 - 15.0_dp / 7.0_dp * xs**6 + 121.0_dp / 80.0_dp * xs**4 &
       & - 0.5_dp * xs**2 + 3.0_dp / 32.0_dp)
  e1 = (C0 - CI) * (xs + SR2 * G002) + C2 * (G002 + SR2 * G022) + CI
  e3 = 5.0_dp * (C0 - CI) * (G002 + SR2 * G022) + 5.0_dp * C2 * (G022 + SR2 * G222) + SR2 * CI
  e5 = 9.0_dp * (C0 - CI) * (G004 + SR2 * G024) + 9.0_dp * C2 * (G024 + SR2 * G224)
  err = max(abs(l_H(1) - e1), abs(l_H(3) - e3), abs(l_H(5) - e5))
  if (err > 1e-14_dp) call bad('calc_Hn_exact coefficients', err)
  lc(1) = 0.0_dp; lc(3) = 0.0_dp; lc(5) = 0.0_dp
  err = maxval(abs(lc))
  if (err > 2e-14_dp) call bad('vanishing coefficients', err)
  err = maxval(abs(y_l(l_y(H)) - H))
  if (err > 5e-12_dp) call bad('y_l(l_y(H)) = H', err)
  ! multi-layer forms agree with the single-layer ones
  y2(:, 1) = H; y2(:, 2) = x; y2(:, 3) = P2
  l2 = la_ya(y2)
  err = maxval(abs(l2(:, 1) - l_y(H))) + abs(l2(1, 2) - 1.0_dp / sqrt(3.0_dp)) + abs(l2(2, 3) - 1.0_dp / sqrt(5.0_dp))
  if (err > 1e-14_dp) call bad('la_ya', err)
  err = maxval(abs(ya_la(l2) - y2))
  if (err > 5e-12_dp) call bad('ya_la', err)
  call l_Finalize()
  if (nbad == 0) then
    print '(a)', 'PASS l_module drop-in (l_initial(1024, 1024); calc_Hn_exact, round trip, multi-layer)'
  else
    print '(a, i0, a)', 'FAIL: ', nbad, ' checks'
    error stop 1
  end if

contains

  subroutine bad(what, e)
    character(*), intent(in) :: what
    real(dp), intent(in) :: e
    print '(a, a, a, es10.3)', 'check failed: ', what, ' err=', e
    nbad = nbad + 1
  end subroutine bad

end program test_l_module
