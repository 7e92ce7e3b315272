! wa_module.f90 -- SPML-compatible multi-layer spectral transform module over
! the MI355X C ABI of libsht_hip.so (include/sht.h), through ISO_C_BINDING.
!
! Drop-in for SPML's wa_module as DCPAM's dynamical core uses it (calls inside
! DynamicsHSplVAS83, atm/dcpam_main_mod.f90:2173-2180; initialisation inside
! DynamicsHSplVAS83Init, :5036): the same public names and array shapes,
!   xya(0:im-1, 1:jm, 1:k)   real(8) grid, latitudes south to north
!   wa((nm+1)**2, 1:k)       real(8) spectra, index l_nm(n, m) = n(n+1)+m+1
! with one global plan per process as in SPML.  A failed call aborts with the
! library's message, as SPML aborts through MessageNotify.
! Linking: replace -lspml-omp -lisp-fort (Mkinclude:83-84) by this module's
! object and -lsht_hip.
module wa_module
  use iso_c_binding
  implicit none
  private

  public :: wa_Initial, wa_Finalize
  public :: xya_wa, wa_xya, xya_GradLon_wa, xya_GradLat_wa
  public :: xy_wa, wa_xy
  public :: wa_DivLon_xya, wa_DivLat_xya, wa_Lapla_wa, wa_LaplaInv_wa
  public :: l_nm, nm_l
  public :: AvrLonLat_xy, IntLonLat_xy
  public :: x_Lon, y_Lat, x_Lon_Weight, y_Lat_Weight, xy_Lon, xy_Lat

  integer, parameter :: dp = c_double

  type, bind(C) :: sht_desc
    integer(c_int) :: nmax, im, jm, kmax, precision, flags
  end type sht_desc

  integer(c_int), parameter :: SHT_FLAG_GRAD = 1

  interface
    integer(c_int) function c_plan_create(desc, plan) bind(C, name="sht_plan_create")
      import :: c_int, c_ptr, sht_desc
      type(sht_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: plan
    end function
    subroutine c_plan_destroy(plan) bind(C, name="sht_plan_destroy")
      import :: c_ptr
      type(c_ptr), value :: plan
    end subroutine
    type(c_ptr) function c_last_error() bind(C, name="sht_last_error")
      import :: c_ptr
    end function
    integer(c_int) function c_plan_grid(plan, lon, lat, lonw, latw) bind(C, name="sht_plan_grid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(out) :: lon(*), lat(*), lonw(*), latw(*)
    end function
    integer(c_int) function c_xya_wa(plan, wa, xya, nlev) bind(C, name="sht_xya_wa")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: wa(*)
      real(c_double), intent(out) :: xya(*)
      integer(c_int), value :: nlev
    end function
    integer(c_int) function c_wa_xya(plan, xya, wa, nlev) bind(C, name="sht_wa_xya")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: xya(*)
      real(c_double), intent(out) :: wa(*)
      integer(c_int), value :: nlev
    end function
    integer(c_int) function c_gradlon(plan, wa, xya, nlev) bind(C, name="sht_xya_gradlon_wa")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: wa(*)
      real(c_double), intent(out) :: xya(*)
      integer(c_int), value :: nlev
    end function
    integer(c_int) function c_gradlat(plan, wa, xya, nlev) bind(C, name="sht_xya_gradlat_wa")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: wa(*)
      real(c_double), intent(out) :: xya(*)
      integer(c_int), value :: nlev
    end function
    integer(c_int) function c_divlon(plan, xya, wa, nlev) bind(C, name="sht_wa_divlon_xya")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: xya(*)
      real(c_double), intent(out) :: wa(*)
      integer(c_int), value :: nlev
    end function
    integer(c_int) function c_divlat(plan, xya, wa, nlev) bind(C, name="sht_wa_divlat_xya")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: xya(*)
      real(c_double), intent(out) :: wa(*)
      integer(c_int), value :: nlev
    end function
    integer(c_int) function c_lapla(plan, wa, wout, nlev) bind(C, name="sht_wa_lapla_wa")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: wa(*)
      real(c_double), intent(out) :: wout(*)
      integer(c_int), value :: nlev
    end function
    integer(c_int) function c_laplainv(plan, wa, wout, nlev) bind(C, name="sht_wa_laplainv_wa")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: wa(*)
      real(c_double), intent(out) :: wout(*)
      integer(c_int), value :: nlev
    end function
    integer(c_int) function c_l_nm(n, m) bind(C, name="sht_l_nm")
      import :: c_int
      integer(c_int), value :: n, m
    end function
    integer(c_int) function c_nm_l(l, n, m) bind(C, name="sht_nm_l")
      import :: c_int
      integer(c_int), value :: l
      integer(c_int), intent(out) :: n, m
    end function
  end interface

  type(c_ptr), save :: plan = c_null_ptr
  integer, save :: im = 0, jm = 0, nm = 0, km = 0
  real(dp), allocatable, save :: x_Lon(:), y_Lat(:), x_Lon_Weight(:), y_Lat_Weight(:)
  real(dp), allocatable, save :: xy_Lon(:, :), xy_Lat(:, :)

contains

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(c_last_error(), msg, [512])
    text = ''
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(*, '(a, a, a, i0, a, a)') 'wa_module: ', where, ' failed (', rc, '): ', trim(text)
    error stop 1
  end subroutine check

  ! SPML wa_Initial(n_in, i_in, j_in, k_in [, np_in]); np_in (OpenMP threads in
  ! SPML) is accepted and unused.  prec_in selects the Legendre arithmetic
  ! (SHT_PREC_*, default FP64).
  subroutine wa_Initial(n_in, i_in, j_in, k_in, np_in, prec_in)
    integer, intent(in) :: n_in, i_in, j_in, k_in
    integer, intent(in), optional :: np_in, prec_in
    type(sht_desc) :: d
    integer :: i
    if (c_associated(plan)) call wa_Finalize()
    d = sht_desc(n_in, i_in, j_in, k_in, 0, SHT_FLAG_GRAD)
    if (present(prec_in)) d%precision = prec_in
    call check(c_plan_create(d, plan), 'wa_Initial')
    nm = n_in; im = i_in; jm = j_in; km = k_in
    allocate(x_Lon(0:im-1), x_Lon_Weight(0:im-1), y_Lat(1:jm), y_Lat_Weight(1:jm))
    call check(c_plan_grid(plan, x_Lon, y_Lat, x_Lon_Weight, y_Lat_Weight), 'wa_Initial')
    allocate(xy_Lon(0:im-1, 1:jm), xy_Lat(0:im-1, 1:jm))
    do i = 0, im - 1
      xy_Lon(i, :) = x_Lon(i)
      xy_Lat(i, :) = y_Lat
    end do
    if (present(np_in)) continue
  end subroutine wa_Initial

  subroutine wa_Finalize()
    if (c_associated(plan)) call c_plan_destroy(plan)
    plan = c_null_ptr
    if (allocated(x_Lon)) deallocate(x_Lon, x_Lon_Weight, y_Lat, y_Lat_Weight, xy_Lon, xy_Lat)
  end subroutine wa_Finalize

  function xya_wa(wa) result(xya)
    real(dp), intent(in) :: wa(:, :)
    real(dp) :: xya(0:im-1, 1:jm, size(wa, 2))
    call check(c_xya_wa(plan, wa, xya, int(size(wa, 2), c_int)), 'xya_wa')
  end function xya_wa

  function wa_xya(xya) result(wa)
    real(dp), intent(in) :: xya(0:, :, :)
    real(dp) :: wa((nm + 1)**2, size(xya, 3))
    call check(c_wa_xya(plan, xya, wa, int(size(xya, 3), c_int)), 'wa_xya')
  end function wa_xya

  ! (1/cos lat) d/dlon
  function xya_GradLon_wa(wa) result(xya)
    real(dp), intent(in) :: wa(:, :)
    real(dp) :: xya(0:im-1, 1:jm, size(wa, 2))
    call check(c_gradlon(plan, wa, xya, int(size(wa, 2), c_int)), 'xya_GradLon_wa')
  end function xya_GradLon_wa

  ! d/dlat
  function xya_GradLat_wa(wa) result(xya)
    real(dp), intent(in) :: wa(:, :)
    real(dp) :: xya(0:im-1, 1:jm, size(wa, 2))
    call check(c_gradlat(plan, wa, xya, int(size(wa, 2), c_int)), 'xya_GradLat_wa')
  end function xya_GradLat_wa

  ! single-layer forms (w_module's xy_w / w_xy over the same plan)
  function xy_wa(w) result(xy)
    real(dp), intent(in) :: w(:)
    real(dp) :: xy(0:im-1, 1:jm)
    call check(c_xya_wa(plan, w, xy, 1_c_int), 'xy_wa')
  end function xy_wa

  function wa_xy(xy) result(w)
    real(dp), intent(in) :: xy(0:, :)
    real(dp) :: w((nm + 1)**2)
    call check(c_wa_xya(plan, xy, w, 1_c_int), 'wa_xy')
  end function wa_xy

  ! (1/cos lat) dA/dlon to spectra
  function wa_DivLon_xya(xya) result(wa)
    real(dp), intent(in) :: xya(0:, :, :)
    real(dp) :: wa((nm + 1)**2, size(xya, 3))
    call check(c_divlon(plan, xya, wa, int(size(xya, 3), c_int)), 'wa_DivLon_xya')
  end function wa_DivLon_xya

  ! (1/cos lat) d(A cos lat)/dlat to spectra
  function wa_DivLat_xya(xya) result(wa)
    real(dp), intent(in) :: xya(0:, :, :)
    real(dp) :: wa((nm + 1)**2, size(xya, 3))
    call check(c_divlat(plan, xya, wa, int(size(xya, 3), c_int)), 'wa_DivLat_xya')
  end function wa_DivLat_xya

  ! -n(n+1) wa (unit sphere)
  function wa_Lapla_wa(wa) result(wout)
    real(dp), intent(in) :: wa(:, :)
    real(dp) :: wout(size(wa, 1), size(wa, 2))
    call check(c_lapla(plan, wa, wout, int(size(wa, 2), c_int)), 'wa_Lapla_wa')
  end function wa_Lapla_wa

  ! wa / (-n(n+1)), zero for n = 0
  function wa_LaplaInv_wa(wa) result(wout)
    real(dp), intent(in) :: wa(:, :)
    real(dp) :: wout(size(wa, 1), size(wa, 2))
    call check(c_laplainv(plan, wa, wout, int(size(wa, 2), c_int)), 'wa_LaplaInv_wa')
  end function wa_LaplaInv_wa

  integer function l_nm(n, m)
    integer, intent(in) :: n, m
    l_nm = c_l_nm(n, m)
    if (l_nm < 1) call check(int(l_nm, c_int), 'l_nm')
  end function l_nm

  function nm_l(l) result(nm_pair)
    integer, intent(in) :: l
    integer :: nm_pair(2)
    integer(c_int) :: n, m
    call check(c_nm_l(l, n, m), 'nm_l')
    nm_pair = [int(n), int(m)]
  end function nm_l

  ! IntLonLat_xy / AvrLonLat_xy (atm/dccm_atm_mod.f90:668-669, :774):
  ! sum_ij (2 pi/im) w_j f_ij and that over 4 pi.
  real(dp) function IntLonLat_xy(xy)
    real(dp), intent(in) :: xy(0:, :)
    integer :: j
    IntLonLat_xy = 0.0_dp
    do j = 1, jm
      IntLonLat_xy = IntLonLat_xy + y_Lat_Weight(j) * sum(x_Lon_Weight * xy(:, j))
    end do
  end function IntLonLat_xy

  real(dp) function AvrLonLat_xy(xy)
    real(dp), intent(in) :: xy(0:, :)
    AvrLonLat_xy = IntLonLat_xy(xy) / (4.0_dp * acos(-1.0_dp))
  end function AvrLonLat_xy

end module wa_module
