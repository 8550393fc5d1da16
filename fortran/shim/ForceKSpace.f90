! ForceKSpace.f90 -- drop-in for YAMC's ForceKSpace module
! (ForceKSpace.f90:1-305): the k-space Coulomb and LJ force on one atom from
! its sin/cos and the total sums, computed by the library.
module ForceKSpace
   use iso_c_binding
   use SumSinCosKR
   use FourierGrid
   use SystemSettings, only : TRealArrayPointer
   implicit none

contains

   ! ForceKSpace.f90:59-169 (beta already holds the (1 + sgn(kx)) factor)
   subroutine ForceKSpace_coulomb(beta, sumsincos_coulomb, sincos_kr_coulomb, fx_tot, fy_tot, fz_tot)
      use yamc_ewald
      use yamc_shim_core
      real(8), dimension(:), intent(in) :: beta
      type(TSumSinCosKR), intent(in), target :: sumsincos_coulomb
      type(TSinCosKR), intent(in), target :: sincos_kr_coulomb
      real(8) :: fx_tot, fy_tot, fz_tot
      type(c_ptr) :: ps(8), pk(8)
      real(c_double) :: f3(3)
      call sumsincos_ptrs(sumsincos_coulomb, ps)
      call sincoskr_ptrs(sincos_kr_coulomb, pk)
      call shim_check(yamc_op_force_coulomb(shim_ws(sincos_kr_coulomb % grid % kmax), beta, ps, pk, f3), &
                      'ForceKSpace_coulomb')
      fx_tot = f3(1); fy_tot = f3(2); fz_tot = f3(3)
   end subroutine

   ! ForceKSpace.f90:171-302: the LJ force on an atom of type ityp.
   subroutine ForceKSpace_LJ(beta_LJ, sumsincos_LJ, ntype, ityp, sincos_kr_ityp, fx_tot, fy_tot, fz_tot)
      use yamc_ewald
      use yamc_shim_core
      type(TRealArrayPointer), dimension(:), intent(in), target :: beta_LJ
      type(TSumSinCosKR), dimension(:), intent(in), target :: sumsincos_LJ
      integer, intent(in) :: ntype, ityp
      type(TSinCosKR), intent(in), target :: sincos_kr_ityp
      real(8) :: fx_tot, fy_tot, fz_tot
      type(c_ptr), allocatable :: pb(:), ps(:)
      type(c_ptr) :: pk(8)
      real(c_double) :: f3(3)
      integer :: t
      allocate(pb(max(ntype * ntype, 1)), ps(8 * max(ntype, 1)))
      do t = 1, ntype * ntype
         pb(t) = c_loc(beta_LJ(t) % ptr(1))
      end do
      do t = 1, ntype
         call sumsincos_ptrs(sumsincos_LJ(t), ps(8 * t - 7:8 * t))
      end do
      call sincoskr_ptrs(sincos_kr_ityp, pk)
      call shim_check(yamc_op_force_lj(shim_ws(sincos_kr_ityp % grid % kmax), int(ntype, c_int32_t), &
                                       int(ityp, c_int32_t), pb, ps, pk, f3), 'ForceKSpace_LJ')
      fx_tot = f3(1); fy_tot = f3(2); fz_tot = f3(3)
   end subroutine

end module ForceKSpace
