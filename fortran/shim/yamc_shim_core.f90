! yamc_shim_core.f90 -- shared state of the drop-in modules in this directory.
!
! The modules FourierGrid, SumSinCosKR, RhoSquared, EwaldSumKSpace,
! ForceKSpace, RealSumLocal, EwaldSumRealSpace and EwaldSumTails here keep the
! names, derived types, fields and procedure signatures of the YAMC modules of
! the same names and forward their arithmetic to libyamc_ewald.so (the
! per-procedure surface of include/yamc_ewald.h).  Fields stay ordinary host
! arrays, coherent after every call, so callers that read or write them
! directly (MC.f90, mc_calc_forces.f90) build and run unchanged.
!
! This module holds one device workspace per kmax (its k grid and device
! buffers) and turns a nonzero C status into the reference's error_throw
! (error.f90:112-130), which stops unless the caller catches the code.
module yamc_shim_core
   use iso_c_binding
   use yamc_ewald
   implicit none
   private
   public :: shim_ws, shim_check, shim_release

   integer, parameter :: MAX_KMAX = 40  ! YAMC_MAX_KMAX
   type(c_ptr), save :: ws_by_kmax(MAX_KMAX) = c_null_ptr

contains

   ! The workspace of the k grid of kmax (created on first use).
   function shim_ws(kmax) result(ws)
      integer, intent(in) :: kmax
      type(c_ptr) :: ws
      if (kmax < 1 .or. kmax > MAX_KMAX) then
         call shim_check(int(YAMC_ERROR_LIMITS, c_int), 'kmax outside 1..40')
      end if
      if (.not. c_associated(ws_by_kmax(kmax))) then
         call shim_check(yamc_ws_create(int(kmax, c_int32_t), ws_by_kmax(kmax)), 'yamc_ws_create')
      end if
      ws = ws_by_kmax(kmax)
   end function

   ! Frees every workspace (optional; the process end frees them too).
   subroutine shim_release()
      integer :: k, rc
      do k = 1, MAX_KMAX
         if (c_associated(ws_by_kmax(k))) then
            rc = yamc_ws_destroy(ws_by_kmax(k))
            ws_by_kmax(k) = c_null_ptr
         end if
      end do
   end subroutine

   subroutine shim_check(rc, where)
      use error, only : error_message, error_throw
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: where
      if (rc == YAMC_OK) return
      error_message = where // ': ' // yamc_last_error()
      call error_throw(int(rc))
   end subroutine

end module yamc_shim_core
