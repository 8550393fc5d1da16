! EwaldSumKSpace.f90 -- drop-in for YAMC's EwaldSumKSpace module
! (EwaldSumKSpace.f90:1-797): the k-space Coulomb and LJ Ewald energies.
! beta, beta6, beta12 and beta_LJ, the total sums and the energy fields stay
! host arrays of the reference's type; initBeta, the sums, rho^2, the energies
! and the dU of a move are computed by the library.
module EwaldSumKSpace
   use iso_c_binding
   use FourierGrid
   use SumSinCosKR
   use LJTypes
   use AtomicData
   use SystemSettings, only : TRealArrayPointer
   use RhoSquared
   implicit none

   type :: TEwaldSumKSpace
      type(TFourierGrid) :: grid
      type(TRhoSquared) :: rho_squared_total
      type(TAtomicData), pointer :: atomic_data
      type(TLJTypes), pointer :: lj_types
      real(8), dimension(:), allocatable :: beta, beta6, beta12
      ! beta_LJ((t1-1)*NType+t2) % ptr; (t2,t1) points at the array of (t1,t2)
      type(TRealArrayPointer), dimension(:), allocatable :: beta_LJ
      real(8) :: energy, energy_coulomb, energy_LJ
      real(8), dimension(:), allocatable :: twoC_plus_dC_ppp, twoC_plus_dC_ppm, twoC_plus_dC_pmp, twoC_plus_dC_pmm
      real(8), dimension(:), allocatable :: twoS_plus_dS_ppp, twoS_plus_dS_ppm, twoS_plus_dS_pmp, twoS_plus_dS_pmm
   end type TEwaldSumKSpace

contains

   function kws(this) result(ws)
      use yamc_shim_core
      type(TEwaldSumKSpace), intent(in) :: this
      type(c_ptr) :: ws
      ws = shim_ws(this % grid % kmax)
   end function

   subroutine beta_lj_ptrs(this, p)
      type(TEwaldSumKSpace), intent(in), target :: this
      type(c_ptr), allocatable :: p(:)
      integer :: t, nt
      nt = this % lj_types % NType
      allocate(p(max(nt * nt, 1)))
      do t = 1, nt * nt
         p(t) = c_loc(this % beta_LJ(t) % ptr(1))
      end do
   end subroutine

   ! EwaldSumKSpace.f90:88-149
   subroutine EwaldSumKSpace_alloc(this, atomic_data, lj_types)
      use parameters, only : kmax
      type(TEwaldSumKSpace) :: this
      type(TAtomicData), intent(in), target :: atomic_data
      type(TLJTypes), intent(in), target :: lj_types
      integer :: nk, ntypes, t1, t2
      this % atomic_data => atomic_data
      this % lj_types => lj_types
      call FourierGrid_calcNalloc(kmax, nk)
      call FourierGrid_alloc(this % grid, nk)
      call FourierGrid_init(this % grid, kmax)
      allocate(this % beta(nk), this % beta6(nk), this % beta12(nk))
      ntypes = lj_types % NType
      allocate(this % beta_LJ(ntypes ** 2))
      do t1 = 1, ntypes
         do t2 = t1, ntypes
            allocate(this % beta_LJ((t1 - 1) * ntypes + t2) % ptr(nk))
            if (t1 /= t2) this % beta_LJ((t2 - 1) * ntypes + t1) % ptr => this % beta_LJ((t1 - 1) * ntypes + t2) % ptr
         end do
      end do
      call RhoSquared_alloc(this % rho_squared_total, this % grid, lj_types)
      allocate(this % twoC_plus_dC_ppp(nk), this % twoC_plus_dC_pmp(nk), this % twoC_plus_dC_ppm(nk), &
               this % twoC_plus_dC_pmm(nk))
      allocate(this % twoS_plus_dS_ppp(nk), this % twoS_plus_dS_pmp(nk), this % twoS_plus_dS_ppm(nk), &
               this % twoS_plus_dS_pmm(nk))
   end subroutine

   ! EwaldSumKSpace.f90:151-194
   subroutine EwaldSumKSpace_dealloc(this)
      type(TEwaldSumKSpace) :: this
      integer :: t1, t2, ntypes
      deallocate(this % beta, this % beta6, this % beta12)
      ntypes = this % lj_types % NType
      do t1 = 1, ntypes
         do t2 = t1, ntypes
            deallocate(this % beta_LJ((t1 - 1) * ntypes + t2) % ptr)
         end do
      end do
      deallocate(this % beta_LJ)
      call RhoSquared_dealloc(this % rho_squared_total)
      deallocate(this % twoC_plus_dC_ppp, this % twoC_plus_dC_pmp, this % twoC_plus_dC_ppm, this % twoC_plus_dC_pmm)
      deallocate(this % twoS_plus_dS_ppp, this % twoS_plus_dS_pmp, this % twoS_plus_dS_ppm, this % twoS_plus_dS_pmm)
   end subroutine

   ! EwaldSumKSpace.f90:197-223: beta scaled by scale_beta (volume steps) or
   ! recomputed, then the total sums and energies.
   subroutine EwaldSumKSpace_init(this, scale_beta)
      type(TEwaldSumKSpace) :: this
      real(8), intent(in), optional :: scale_beta
      if (present(scale_beta)) then
         if (scale_beta > 1e-9) then
            this % beta(:) = this % beta(:) * scale_beta
         else
            call EwaldSumKSpace_initBeta(this)
            call EwaldSumKSpace_initBetaLJ(this)
         end if
      else
         call EwaldSumKSpace_initBeta(this)
         call EwaldSumKSpace_initBetaLJ(this)
      end if
      call EwaldSumKSpace_calc_total_energy(this)
   end subroutine

   ! EwaldSumKSpace.f90:225-260
   subroutine EwaldSumKSpace_calc_total_energy(this)
      type(TEwaldSumKSpace), target :: this
      call RhoSquared_nulify(this % rho_squared_total)
      call RhoSquared_addAtoms(this % rho_squared_total, this % atomic_data % xx, this % atomic_data % yy, &
                               this % atomic_data % zz, this % atomic_data % charge, &
                               this % lj_types % type_by_index, this % atomic_data % natom)
      call RhoSquared_calc_rho_squared(this % rho_squared_total)
      this % energy_coulomb = EwaldSumKSpace_calc_coulomb_energy(this)
      this % energy_LJ = EwaldSumKSpace_calc_LJ_energy(this)
      this % energy = this % energy_coulomb + this % energy_LJ
   end subroutine

   ! EwaldSumKSpace.f90:263-350
   subroutine EwaldSumKSpace_initBeta(this)
      use parameters, only : alpha, xclb
      use yamc_ewald
      use yamc_shim_core
      type(TEwaldSumKSpace), target :: this
      call shim_check(yamc_op_beta(kws(this), alpha, xclb, this % beta, this % beta6, this % beta12), &
                      'EwaldSumKSpace_initBeta')
   end subroutine

   ! EwaldSumKSpace.f90:352-386
   subroutine EwaldSumKSpace_initBetaLJ(this)
      use yamc_ewald
      use yamc_shim_core
      type(TEwaldSumKSpace), target :: this
      integer :: t1, t2, nt, off
      nt = this % lj_types % NType
      do t1 = 1, nt
         do t2 = t1, nt
            off = (t1 - 1) * nt + t2
            call shim_check(yamc_op_beta_lj(kws(this), this % lj_types % LJ6Tab(off), this % lj_types % LJ12Tab(off), &
                                            this % beta6, this % beta12, c_loc(this % beta_LJ(off) % ptr(1))), &
                            'EwaldSumKSpace_initBetaLJ')
         end do
      end do
   end subroutine

   ! EwaldSumKSpace.f90:389-416
   function EwaldSumKSpace_calc_coulomb_energy(this)
      use yamc_ewald
      use yamc_shim_core
      real(8) :: EwaldSumKSpace_calc_coulomb_energy
      type(TEwaldSumKSpace), intent(in), target :: this
      real(c_double) :: e
      call shim_check(yamc_op_energy_coulomb(kws(this), this % beta, this % rho_squared_total % rho_squared, e), &
                      'EwaldSumKSpace_calc_coulomb_energy')
      EwaldSumKSpace_calc_coulomb_energy = e
   end function

   ! EwaldSumKSpace.f90:418-526
   function EwaldSumKSpace_calc_LJ_energy(this)
      use yamc_ewald
      use yamc_shim_core
      real(8) :: EwaldSumKSpace_calc_LJ_energy
      type(TEwaldSumKSpace), intent(in), target :: this
      type(c_ptr), allocatable :: pb(:), ps(:)
      real(c_double) :: e
      call beta_lj_ptrs(this, pb)
      call lj_ptrs(this % rho_squared_total, ps)
      call shim_check(yamc_op_energy_lj(kws(this), int(this % lj_types % NType, c_int32_t), pb, ps, e), &
                      'EwaldSumKSpace_calc_LJ_energy')
      EwaldSumKSpace_calc_LJ_energy = e
   end function

   ! EwaldSumKSpace.f90:528-620: the Coulomb energy change when the total
   ! sums change by delta_sumsincos_coulomb.
   function EwaldSumKSpace_calc_dU_coulomb(this, delta_sumsincos_coulomb)
      use yamc_ewald
      use yamc_shim_core
      real(8) :: EwaldSumKSpace_calc_dU_coulomb
      type(TEwaldSumKSpace), target :: this
      type(TSumSinCosKR), target :: delta_sumsincos_coulomb
      type(c_ptr) :: ps(8), pd(8)
      real(c_double) :: du
      call sumsincos_ptrs(this % rho_squared_total % sumsincos_coulomb, ps)
      call sumsincos_ptrs(delta_sumsincos_coulomb, pd)
      call shim_check(yamc_op_du_coulomb(kws(this), this % beta, ps, pd, du), 'EwaldSumKSpace_calc_dU_coulomb')
      EwaldSumKSpace_calc_dU_coulomb = du
   end function

   ! EwaldSumKSpace.f90:623-753: the LJ energy change; types absent from
   ! the delta (type_is_present false) carry no delta terms.
   function EwaldSumKSpace_calc_dU_LJ(this, delta_sumsincos_LJ, type_is_present)
      use yamc_ewald
      use yamc_shim_core
      real(8) :: EwaldSumKSpace_calc_dU_LJ
      type(TEwaldSumKSpace), target :: this
      type(TSumSinCosKR), dimension(:), intent(in), target :: delta_sumsincos_LJ
      logical, dimension(:) :: type_is_present
      type(c_ptr), allocatable :: pb(:), ps(:), pd(:)
      integer(c_int32_t), allocatable :: pres(:)
      real(c_double) :: du
      integer :: t, nt
      nt = this % lj_types % NType
      call beta_lj_ptrs(this, pb)
      call lj_ptrs(this % rho_squared_total, ps)
      allocate(pd(8 * max(nt, 1)), pres(max(nt, 1)))
      do t = 1, nt
         call sumsincos_ptrs(delta_sumsincos_LJ(t), pd(8 * t - 7:8 * t))
         pres(t) = merge(1, 0, type_is_present(t))
      end do
      call shim_check(yamc_op_du_lj(kws(this), int(nt, c_int32_t), pb, ps, pd, pres, du), 'EwaldSumKSpace_calc_dU_LJ')
      EwaldSumKSpace_calc_dU_LJ = du
   end function

end module EwaldSumKSpace
