! RhoSquared.f90 -- drop-in for YAMC's RhoSquared module (RhoSquared.f90:9-341):
! the Coulomb S(k) and one S(k) per LJ type.  The sin/cos of an added atom,
! the sums of many atoms and rho^2 come from the library; the sums
! themselves are the host arrays of the fields.
module RhoSquared
   use iso_c_binding
   use SumSinCosKR
   use FourierGrid
   use LJTypes
   implicit none

   type :: TRhoSquared
      type(TLJTypes), pointer :: lj_types
      type(TFourierGrid), pointer :: grid
      type(TSumSinCosKR) :: sumsincos_coulomb
      real(8), dimension(:), allocatable :: rho_squared
      type(TSumSinCosKR), dimension(:), allocatable :: sumsincos_LJ
      logical, dimension(:), allocatable :: type_is_present
      type(TSinCosXYZ) :: sincos_xyz
   end type TRhoSquared

contains

   ! The 8 * (1 + NType) arrays of all sets: Coulomb, then LJ type 1 .. NType.
   subroutine rho_ptrs(this, p)
      type(TRhoSquared), target :: this
      type(c_ptr), allocatable :: p(:)
      integer :: t, nt
      nt = this % lj_types % NType
      allocate(p(8 * (1 + nt)))
      call sumsincos_ptrs(this % sumsincos_coulomb, p(1:8))
      do t = 1, nt
         call sumsincos_ptrs(this % sumsincos_LJ(t), p(8 * t + 1:8 * t + 8))
      end do
   end subroutine

   subroutine lj_ptrs(this, p)
      type(TRhoSquared), target :: this
      type(c_ptr), allocatable :: p(:)
      integer :: t, nt
      nt = this % lj_types % NType
      allocate(p(8 * max(nt, 1)))
      do t = 1, nt
         call sumsincos_ptrs(this % sumsincos_LJ(t), p(8 * t - 7:8 * t))
      end do
   end subroutine

   ! RhoSquared.f90:49-83
   subroutine RhoSquared_alloc(this, grid, lj_types)
      type(TRhoSquared) :: this
      type(TFourierGrid), target :: grid
      type(TLJTypes), target :: lj_types
      integer :: ntype, i
      this % lj_types => lj_types
      this % grid => grid
      call SinCosXYZ_alloc(this % sincos_xyz, grid % kmax)
      call SumSinCosKR_alloc(this % sumsincos_coulomb, grid)
      ntype = lj_types % NType
      allocate(this % sumsincos_LJ(ntype))
      do i = 1, ntype
         call SumSinCosKR_alloc(this % sumsincos_LJ(i), grid)
      end do
      allocate(this % rho_squared(grid % nk))
      allocate(this % type_is_present(ntype))
   end subroutine

   ! RhoSquared.f90:85-115
   subroutine RhoSquared_dealloc(this)
      type(TRhoSquared), target :: this
      integer :: ntype, i
      ntype = this % lj_types % NType
      nullify(this % lj_types)
      nullify(this % grid)
      call SinCosXYZ_dealloc(this % sincos_xyz)
      call SumSinCosKR_dealloc(this % sumsincos_coulomb)
      do i = 1, ntype
         call SumSinCosKR_dealloc(this % sumsincos_LJ(i))
      end do
      deallocate(this % sumsincos_LJ, this % rho_squared, this % type_is_present)
   end subroutine

   ! RhoSquared.f90:117-131
   subroutine RhoSquared_nulify(this)
      type(TRhoSquared) :: this
      integer :: i
      call SumSinCosKR_nulify(this % sumsincos_coulomb)
      do i = 1, this % lj_types % NType
         call SumSinCosKR_nulify(this % sumsincos_LJ(i))
      end do
      this % type_is_present(:) = .FALSE.
   end subroutine

   ! RhoSquared.f90:133-186: the atom's q sin/cos into the Coulomb sums (when
   ! charged), its LJ sin/cos (the Coulomb ones times 1/q for a charged atom)
   ! into the sums of its LJ type; both are returned in the sincos arguments.
   subroutine RhoSquared_addAtom(this, xx, yy, zz, charge, lj_type, sincos_kr_coulomb, sincos_kr_LJ)
      type(TRhoSquared) :: this
      real(8), intent(in) :: xx, yy, zz, charge
      integer, intent(in) :: lj_type
      type(TSinCosKR) :: sincos_kr_coulomb, sincos_kr_LJ
      if (charge /= 0.0) then
         call SinCosKR_fill_atom(sincos_kr_coulomb, xx, yy, zz, charge)
         call SumSinCosKR_addAtom(this % sumsincos_coulomb, sincos_kr_coulomb)
      else
         call SinCosKR_nulify(sincos_kr_coulomb)
      end if
      if (lj_type > 0) then
         if (charge /= 0.0) then
            call SinCosKR_copy(sincos_kr_LJ, sincos_kr_coulomb)
            call SinCosKR_mulScalar(sincos_kr_LJ, 1.0d0 / charge)
         else
            call SinCosKR_fill_atom(sincos_kr_LJ, xx, yy, zz, 1.0d0)
         end if
         call SumSinCosKR_addAtom(this % sumsincos_LJ(lj_type), sincos_kr_LJ)
         this % type_is_present(lj_type) = .TRUE.
      end if
   end subroutine

   ! RhoSquared.f90:188-243: the natom atoms' sums added to the sums.
   subroutine RhoSquared_addAtoms(this, xx, yy, zz, charge, type_by_index, natom)
      use yamc_ewald
      use yamc_shim_core
      type(TRhoSquared), target :: this
      real(8), dimension(:), intent(in) :: xx, yy, zz, charge
      integer, dimension(:), intent(in) :: type_by_index
      integer, intent(in) :: natom
      type(TRhoSquared), target :: part
      type(c_ptr), allocatable :: p(:)
      integer(c_int32_t), allocatable :: present(:), typ(:)
      integer :: nt, t
      nt = this % lj_types % NType
      call RhoSquared_alloc(part, this % grid, this % lj_types)
      call rho_ptrs(part, p)
      allocate(present(max(nt, 1)), typ(max(natom, 1)))
      typ(1:natom) = type_by_index(1:natom)
      call shim_check(yamc_op_sums(shim_ws(this % grid % kmax), int(natom, c_int32_t), xx, yy, zz, charge, typ, &
                                   int(nt, c_int32_t), p, present), 'RhoSquared_addAtoms')
      call SumSinCosKR_addSum(this % sumsincos_coulomb, part % sumsincos_coulomb)
      do t = 1, nt
         call SumSinCosKR_addSum(this % sumsincos_LJ(t), part % sumsincos_LJ(t))
         if (present(t) /= 0) this % type_is_present(t) = .TRUE.
      end do
      call RhoSquared_dealloc(part)
   end subroutine

   ! RhoSquared.f90:245-273
   subroutine RhoSquared_calc_rho_squared(this)
      use yamc_ewald
      use yamc_shim_core
      type(TRhoSquared), target :: this
      type(c_ptr) :: p(8)
      call sumsincos_ptrs(this % sumsincos_coulomb, p)
      call shim_check(yamc_op_rho_squared(shim_ws(this % grid % kmax), p, this % rho_squared), &
                      'RhoSquared_calc_rho_squared')
   end subroutine

   ! RhoSquared.f90:275-341 (copy, add and sub leave type_is_present alone)
   subroutine RhoSquared_copy(dst, src)
      type(TRhoSquared) :: dst, src
      integer :: i
      call SumSinCosKR_copy(dst % sumsincos_coulomb, src % sumsincos_coulomb)
      do i = 1, dst % lj_types % NType
         call SumSinCosKR_copy(dst % sumsincos_LJ(i), src % sumsincos_LJ(i))
      end do
   end subroutine

   subroutine RhoSquared_add(this, rho_squared)
      type(TRhoSquared) :: this, rho_squared
      integer :: i
      call SumSinCosKR_addSum(this % sumsincos_coulomb, rho_squared % sumsincos_coulomb)
      do i = 1, this % lj_types % NType
         call SumSinCosKR_addSum(this % sumsincos_LJ(i), rho_squared % sumsincos_LJ(i))
      end do
   end subroutine

   subroutine RhoSquared_sub(this, rho_squared)
      type(TRhoSquared) :: this, rho_squared
      integer :: i
      call SumSinCosKR_subSum(this % sumsincos_coulomb, rho_squared % sumsincos_coulomb)
      do i = 1, this % lj_types % NType
         call SumSinCosKR_subSum(this % sumsincos_LJ(i), rho_squared % sumsincos_LJ(i))
      end do
   end subroutine

   ! RhoSquared.f90:310-339: the sums per k as text.
   subroutine RhoSquared_save_to_file(this, filename)
      use io, only : io_open, io_close
      use SystemSettings, only : SYSTEM_STRING_LENGTH
      type(TRhoSquared), intent(in), target :: this
      character(SYSTEM_STRING_LENGTH), intent(in) :: filename
      type(TSumSinCosKR), pointer :: ssc
      integer :: hfile, i, k
      hfile = io_open(filename, 'w')
      do k = 1, this % grid % nk
         ssc => this % sumsincos_coulomb
         write(hfile, '(F20.10,F20.10,F20.10,F20.10,$)') ssc % sumcos_ppp(k), ssc % sumcos_pmp(k), &
                                                          ssc % sumcos_ppm(k), ssc % sumcos_pmm(k)
         write(hfile, '(F20.10,F20.10,F20.10,F20.10,$)') ssc % sumsin_ppp(k), ssc % sumsin_pmp(k), &
                                                          ssc % sumsin_ppm(k), ssc % sumsin_pmm(k)
         do i = 1, this % lj_types % NType
            ssc => this % sumsincos_LJ(i)
            write(hfile, '(F20.10,F20.10,F20.10,F20.10,$)') ssc % sumcos_ppp(k), ssc % sumcos_pmp(k), &
                                                             ssc % sumcos_ppm(k), ssc % sumcos_pmm(k)
            write(hfile, '(F20.10,F20.10,F20.10,F20.10,$)') ssc % sumsin_ppp(k), ssc % sumsin_pmp(k), &
                                                             ssc % sumsin_ppm(k), ssc % sumsin_pmm(k)
         end do
         write(hfile, *)
      end do
      call io_close(hfile)
   end subroutine

end module RhoSquared
