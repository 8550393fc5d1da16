! EwaldSumTails.f90 -- drop-in for YAMC's EwaldSumTails module
! (EwaldSumTails.f90:1-130): the position-independent Ewald terms, computed
! by the library in the reference's summation order.
module EwaldSumTails
   implicit none

contains

   ! EwaldSumTails.f90:18-38: -xclb alpha/sqrt(pi) SUM q_i^2
   function ewald_sum_coulomb_tail(charge, natom)
      use iso_c_binding
      use parameters, only : alpha, xclb, kmax
      use yamc_ewald
      use yamc_shim_core
      real(8) :: ewald_sum_coulomb_tail
      real(8), dimension(:), intent(in) :: charge
      integer, intent(in) :: natom
      real(c_double) :: self_c, l6, l12, z(1)
      integer(c_int32_t) :: c(1)
      z = 0.0d0
      c = 0
      call shim_check(yamc_op_tails(shim_ws(kmax), int(natom, c_int32_t), charge, 0_c_int32_t, z, z, c, alpha, xclb, &
                                    self_c, l6, l12), 'ewald_sum_coulomb_tail')
      ewald_sum_coulomb_tail = self_c
   end function

   ! EwaldSumTails.f90:45-127: the LJ6 and LJ12 terms.  As the reference
   ! does, the count of LJ type t is comp % mol_numbers(t).
   subroutine ewald_sum_lj_tails(comp, lj_types, lj6_tail, lj12_tail)
      use iso_c_binding
      use composition, only : TComposition
      use LJTypes, only : TLJTypes
      use parameters, only : alpha, xclb, kmax
      use yamc_ewald
      use yamc_shim_core
      type(TComposition), intent(in) :: comp
      type(TLJTypes), intent(in) :: lj_types
      real(8), intent(out) :: lj6_tail, lj12_tail
      real(c_double) :: self_c, q(1)
      integer(c_int32_t), allocatable :: counts(:)
      integer :: t, nt
      nt = lj_types % NType
      allocate(counts(max(nt, 1)))
      counts = 0
      do t = 1, nt
         counts(t) = comp % mol_numbers(t)
      end do
      q = 0.0d0
      call shim_check(yamc_op_tails(shim_ws(kmax), 0_c_int32_t, q, int(nt, c_int32_t), lj_types % LJ6Tab, &
                                    lj_types % LJ12Tab, counts, alpha, xclb, self_c, lj6_tail, lj12_tail), &
                      'ewald_sum_lj_tails')
   end subroutine

end module EwaldSumTails
