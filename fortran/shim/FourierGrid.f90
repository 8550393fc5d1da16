! FourierGrid.f90 -- drop-in for YAMC's FourierGrid module
! (FourierGrid.f90:1-189): the half-space k list, computed by the library
! (yamc_ws_grid) and kept in the same fields.
module FourierGrid
   use iso_c_binding
   implicit none

   type :: TFourierGrid
      integer :: kmax
      integer :: nk
      integer :: nalloc
      integer, dimension(:), allocatable :: ip
      integer, dimension(:), allocatable :: iq
      integer, dimension(:), allocatable :: kx, ky, kz
      real(8), dimension(:), allocatable :: xk, xk2
   end type TFourierGrid

contains

   ! FourierGrid.f90:37-77: the number of k vectors of |k| <= kmax.
   subroutine FourierGrid_calcNalloc(kmax, nalloc)
      use yamc_ewald
      use yamc_shim_core
      integer, intent(in) :: kmax
      integer :: nalloc
      integer(c_int32_t) :: nk
      call shim_check(yamc_ws_grid(shim_ws(kmax), nk, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                                   c_null_ptr, c_null_ptr), 'FourierGrid_calcNalloc')
      nalloc = nk
   end subroutine

   ! FourierGrid.f90:80-99
   subroutine FourierGrid_alloc(this, nalloc)
      type(TFourierGrid) :: this
      integer, intent(in) :: nalloc
      allocate(this % kx(nalloc), this % ky(nalloc), this % kz(nalloc))
      allocate(this % ip(nalloc), this % iq(nalloc))
      allocate(this % xk(nalloc), this % xk2(nalloc))
      this % nalloc = nalloc
      this % nk = 0
   end subroutine

   ! FourierGrid.f90:101-116
   subroutine FourierGrid_dealloc(this)
      type(TFourierGrid) :: this
      if (allocated(this % kx)) deallocate(this % kx, this % ky, this % kz, this % ip, this % iq)
      if (allocated(this % xk)) deallocate(this % xk, this % xk2)
      this % nalloc = 0
      this % nk = 0
   end subroutine

   ! FourierGrid.f90:118-189: kx, ky, kz, xk2, xk, iq, ip of every k.
   subroutine FourierGrid_init(this, kmax)
      use yamc_ewald
      use yamc_shim_core
      use error
      type(TFourierGrid) :: this
      integer, intent(in) :: kmax
      integer(c_int32_t) :: nk
      integer(c_int32_t), allocatable, target :: kx(:), ky(:), kz(:), ip(:), iq(:)
      real(c_double), allocatable, target :: xk(:), xk2(:)
      type(c_ptr) :: ws
      ws = shim_ws(kmax)
      call shim_check(yamc_ws_grid(ws, nk, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                                   c_null_ptr), 'FourierGrid_init')
      if (nk > this % nalloc) then
         write(error_message, *) 'FourierGrid_init: k exceeds allocated limit ', this % nalloc
         call error_throw(ERROR_LIMITS)
         return
      end if
      allocate(kx(nk), ky(nk), kz(nk), ip(nk), iq(nk), xk(nk), xk2(nk))
      call shim_check(yamc_ws_grid(ws, nk, c_loc(kx), c_loc(ky), c_loc(kz), c_loc(ip), c_loc(iq), c_loc(xk), &
                                   c_loc(xk2)), 'FourierGrid_init')
      this % kmax = kmax
      this % nk = nk
      this % kx(1:nk) = kx
      this % ky(1:nk) = ky
      this % kz(1:nk) = kz
      this % ip(1:nk) = ip
      this % iq(1:nk) = iq
      this % xk(1:nk) = xk
      this % xk2(1:nk) = xk2
   end subroutine

end module FourierGrid
