! RealSumLocal.f90 -- drop-in for YAMC's RealSumLocal module
! (RealSumLocal.f90:1-170): one pair's real-space Ewald terms in four steps.
! Every step evaluates the pair on the device (yamc_op_pair) from what the
! earlier steps recorded, and the fields hold the library's values.  Three
! components past the reference's fields carry the pair between the steps.
module RealSumLocal
   use iso_c_binding
   implicit none

   type :: TRealSumLocal
      real(8) :: r, r2
      real(8) :: a, a2, a6
      real(8) :: exp_minus_a2
      real(8) :: erfc_a
      real(8) :: C6, C12
      real(8) :: four_epsilon
      real(8) :: sigma_over_r_6
      real(8) :: potew
      real(8) :: potlj12, potlj6
      real(8) :: pot
      real(8) :: FR
      ! the pair as the steps so far gave it
      logical :: same_molecule = .false.
      real(8) :: qi = 0.0d0, qj = 0.0d0, sigma6 = 0.0d0
   end type TRealSumLocal

contains

   subroutine pair_eval(this, lj, out)
      use parameters, only : alpha, xclb, kmax
      use yamc_ewald
      use yamc_shim_core
      type(TRealSumLocal) :: this
      integer, intent(in) :: lj
      real(c_double), intent(out) :: out(14)
      call shim_check(yamc_op_pair(shim_ws(kmax), alpha, xclb, this % r2, merge(1, 0, this % same_molecule), &
                                   this % qi, this % qj, int(lj, c_int32_t), this % sigma6, this % four_epsilon, out), &
                      'RealSumLocal')
   end subroutine

   ! RealSumLocal.f90:63-88: r, a, exp(-a^2) and erfc(a) (-erf(a) within a
   ! molecule); pot = 0.
   subroutine RealSumLocal_init_electrostatics(this, r2, sameMolecule)
      type(TRealSumLocal) :: this
      real(8), intent(in) :: r2
      logical, intent(in) :: sameMolecule
      real(c_double) :: out(14)
      this % r2 = r2
      this % same_molecule = sameMolecule
      this % qi = 0.0d0
      this % qj = 0.0d0
      call pair_eval(this, 0, out)
      this % r = out(6); this % a = out(7); this % a2 = out(8)
      this % exp_minus_a2 = out(9); this % erfc_a = out(10)
      this % pot = 0
   end subroutine

   ! RealSumLocal.f90:90-110: potew, pot = potew and the Coulomb FR.
   subroutine RealSumLocal_calc_electrostatics(this, qi, qj)
      type(TRealSumLocal) :: this
      real(8), intent(in) :: qi, qj
      real(c_double) :: out(14)
      this % qi = qi
      this % qj = qj
      call pair_eval(this, 0, out)
      this % potew = out(1)
      this % pot = out(4)
      this % FR = out(5)
   end subroutine

   ! RealSumLocal.f90:112-132: (sigma/r)^6, C6, a^6, C12.
   subroutine RealSumLocal_init_LJ(this, sigma6)
      type(TRealSumLocal) :: this
      real(8), intent(in) :: sigma6
      real(c_double) :: out(14)
      this % sigma6 = sigma6
      call pair_eval(this, 0, out)
      this % C6 = out(11); this % C12 = out(12); this % a6 = out(13); this % sigma_over_r_6 = out(14)
   end subroutine

   ! RealSumLocal.f90:134-168: potlj6, potlj12; pot and FR gain the LJ terms.
   subroutine RealSumLocal_calc_LJ(this, four_epsilon)
      type(TRealSumLocal) :: this
      real(8) :: four_epsilon
      real(c_double) :: out(14)
      this % four_epsilon = four_epsilon
      call pair_eval(this, 1, out)
      this % potlj6 = out(2)
      this % potlj12 = out(3)
      this % pot = out(4)
      this % FR = out(5)
   end subroutine

end module RealSumLocal
