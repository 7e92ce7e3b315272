! l_module.f90 -- SPML-compatible zonal (m = 0) Legendre transform module over
! the MI355X C ABI of libsht_hip.so (include/sht.h, sht_zonal_*), through
! ISO_C_BINDING.
!
! Drop-in for SPML's l_module as the energy-balance model uses it
! (exp/APEI07Couple/ebm/EBM_main.f90: l_initial(nm, jm) at :64, l_finalize at
! :68, y_Lat and y_Lat_Weight at :153 and :200-204, l_y at :170 and :242):
!   y(1:jm)       real(8) grid, Gauss latitudes south to north
!   l(0:nm)       real(8) coefficients of Pt_n = sqrt(2n+1) P_n,
!                 l_n = (1/2) sum_j w_j Pt_n(mu_j) y_j;  y_l is its inverse
! la_ya / ya_la are the multi-layer forms, y(1:jm, 1:k) and l(0:nm, 1:k).
! With one global plan per process as in SPML; a failed call aborts with the
! library's message, as SPML aborts through MessageNotify.
module l_module
  use iso_c_binding
  implicit none
  private

  public :: l_Initial, l_Finalize, l_y, y_l, la_ya, ya_la
  public :: y_Lat, y_Lat_Weight

  integer, parameter :: dp = c_double

  interface
    integer(c_int) function c_zonal_create(nmax, jm, kmax, z) bind(C, name="sht_zonal_create")
      import :: c_int, c_ptr
      integer(c_int), value :: nmax, jm, kmax
      type(c_ptr), intent(out) :: z
    end function
    subroutine c_zonal_destroy(z) bind(C, name="sht_zonal_destroy")
      import :: c_ptr
      type(c_ptr), value :: z
    end subroutine
    integer(c_int) function c_zonal_grid(z, lat, latw) bind(C, name="sht_zonal_grid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: z
      real(c_double), intent(out) :: lat(*), latw(*)
    end function
    integer(c_int) function c_l_y(z, y, l, nlev) bind(C, name="sht_l_y")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: z
      real(c_double), intent(in) :: y(*)
      real(c_double), intent(out) :: l(*)
      integer(c_int), value :: nlev
    end function
    integer(c_int) function c_y_l(z, l, y, nlev) bind(C, name="sht_y_l")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: z
      real(c_double), intent(in) :: l(*)
      real(c_double), intent(out) :: y(*)
      integer(c_int), value :: nlev
    end function
    type(c_ptr) function c_last_error() bind(C, name="sht_last_error")
      import :: c_ptr
    end function
  end interface

  type(c_ptr), save :: zplan = c_null_ptr
  integer, save :: nm = 0, jm = 0, km = 0
  real(dp), allocatable, save :: y_Lat(:), y_Lat_Weight(:)

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
    write(*, '(a, a, a, i0, a, a)') 'l_module: ', where, ' failed (', rc, '): ', trim(text)
    error stop 1
  end subroutine check

  ! SPML l_Initial(n_in, j_in); k_in (optional) is the most layers per call of
  ! la_ya / ya_la (default 64).
  subroutine l_Initial(n_in, j_in, k_in)
    integer, intent(in) :: n_in, j_in
    integer, intent(in), optional :: k_in
    if (c_associated(zplan)) call l_Finalize()
    km = 64
    if (present(k_in)) km = k_in
    call check(c_zonal_create(n_in, j_in, km, zplan), 'l_Initial')
    nm = n_in; jm = j_in
    allocate(y_Lat(1:jm), y_Lat_Weight(1:jm))
    call check(c_zonal_grid(zplan, y_Lat, y_Lat_Weight), 'l_Initial')
  end subroutine l_Initial

  subroutine l_Finalize()
    if (c_associated(zplan)) call c_zonal_destroy(zplan)
    zplan = c_null_ptr
    if (allocated(y_Lat)) deallocate(y_Lat, y_Lat_Weight)
  end subroutine l_Finalize

  function l_y(y) result(l)
    real(dp), intent(in) :: y(:)
    real(dp) :: l(0:nm)
    call check(c_l_y(zplan, y, l, 1_c_int), 'l_y')
  end function l_y

  function y_l(l) result(y)
    real(dp), intent(in) :: l(0:)
    real(dp) :: y(1:jm)
    call check(c_y_l(zplan, l, y, 1_c_int), 'y_l')
  end function y_l

  function la_ya(y) result(l)
    real(dp), intent(in) :: y(:, :)
    real(dp) :: l(0:nm, size(y, 2))
    call check(c_l_y(zplan, y, l, int(size(y, 2), c_int)), 'la_ya')
  end function la_ya

  function ya_la(l) result(y)
    real(dp), intent(in) :: l(0:, :)
    real(dp) :: y(1:jm, size(l, 2))
    call check(c_y_l(zplan, l, y, int(size(l, 2), c_int)), 'ya_la')
  end function ya_la

end module l_module
