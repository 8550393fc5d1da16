! SumSinCosKR.f90 -- drop-in for YAMC's SumSinCosKR module
! (SumSinCosKR.f90:1-689).  The sin/cos tables of a position (SinCosXYZ_fill)
! and of a charge at every k (SinCosKR_fill) are computed by the library; the
! elementwise maintenance of the sums (nulify, add, subtract, copy, scale)
! works on the host arrays the fields are, in place (a device round trip would
! move three times the bytes these passes touch).
module SumSinCosKR
   use iso_c_binding
   use FourierGrid
   implicit none

   type :: TSinCosXYZ
      integer :: nk
      real(8), dimension(:), allocatable :: cos_xkx, cos_yky, cos_zkz
      real(8), dimension(:), allocatable :: sin_xkx, sin_yky, sin_zkz
   end type TSinCosXYZ

   type :: TSinCosKR
      type(TFourierGrid), pointer :: grid
      real(8), dimension(:), allocatable :: cos_ppp
      real(8), dimension(:), allocatable :: sin_ppp
      real(8), dimension(:), allocatable :: cos_ppm
      real(8), dimension(:), allocatable :: sin_ppm
      real(8), dimension(:), allocatable :: cos_pmp
      real(8), dimension(:), allocatable :: sin_pmp
      real(8), dimension(:), allocatable :: cos_pmm
      real(8), dimension(:), allocatable :: sin_pmm
   end type

   type :: TSumSinCosKR
      type(TFourierGrid), pointer :: grid
      real(8), dimension(:), allocatable :: sumcos_ppp
      real(8), dimension(:), allocatable :: sumsin_ppp
      real(8), dimension(:), allocatable :: sumcos_ppm
      real(8), dimension(:), allocatable :: sumsin_ppm
      real(8), dimension(:), allocatable :: sumcos_pmp
      real(8), dimension(:), allocatable :: sumsin_pmp
      real(8), dimension(:), allocatable :: sumcos_pmm
      real(8), dimension(:), allocatable :: sumsin_pmm
   end type TSumSinCosKR

contains

   ! The 8 arrays of a set in the library's order (cos, sin of ppp, ppm, pmp, pmm).
   subroutine sumsincos_ptrs(s, p)
      type(TSumSinCosKR), target :: s
      type(c_ptr) :: p(8)
      p(1) = c_loc(s % sumcos_ppp); p(2) = c_loc(s % sumsin_ppp)
      p(3) = c_loc(s % sumcos_ppm); p(4) = c_loc(s % sumsin_ppm)
      p(5) = c_loc(s % sumcos_pmp); p(6) = c_loc(s % sumsin_pmp)
      p(7) = c_loc(s % sumcos_pmm); p(8) = c_loc(s % sumsin_pmm)
   end subroutine

   subroutine sincoskr_ptrs(s, p)
      type(TSinCosKR), target :: s
      type(c_ptr) :: p(8)
      p(1) = c_loc(s % cos_ppp); p(2) = c_loc(s % sin_ppp)
      p(3) = c_loc(s % cos_ppm); p(4) = c_loc(s % sin_ppm)
      p(5) = c_loc(s % cos_pmp); p(6) = c_loc(s % sin_pmp)
      p(7) = c_loc(s % cos_pmm); p(8) = c_loc(s % sin_pmm)
   end subroutine

   ! ---- SinCosXYZ (SumSinCosKR.f90:150-223) ----
   subroutine SinCosXYZ_alloc(this, kmax)
      type(TSinCosXYZ) :: this
      integer, intent(in) :: kmax
      allocate(this % cos_xkx(0:kmax), this % cos_yky(0:kmax), this % cos_zkz(0:kmax))
      allocate(this % sin_xkx(0:kmax), this % sin_yky(0:kmax), this % sin_zkz(0:kmax))
   end subroutine

   subroutine SinCosXYZ_dealloc(this)
      type(TSinCosXYZ) :: this
      deallocate(this % cos_xkx, this % cos_yky, this % cos_zkz)
      deallocate(this % sin_xkx, this % sin_yky, this % sin_zkz)
   end subroutine

   subroutine SinCosXYZ_fill(this, x, y, z, kmax)
      use yamc_ewald
      use yamc_shim_core
      type(TSinCosXYZ), target :: this
      real(8), intent(in) :: x, y, z
      integer, intent(in) :: kmax
      type(c_ptr) :: t(6)
      t(1) = c_loc(this % cos_xkx); t(2) = c_loc(this % cos_yky); t(3) = c_loc(this % cos_zkz)
      t(4) = c_loc(this % sin_xkx); t(5) = c_loc(this % sin_yky); t(6) = c_loc(this % sin_zkz)
      call shim_check(yamc_op_sincos_xyz(shim_ws(kmax), x, y, z, t), 'SinCosXYZ_fill')
   end subroutine

   ! ---- SinCosKR (SumSinCosKR.f90:229-516) ----
   subroutine SinCosKR_alloc(this, grid)
      type(TSinCosKR) :: this
      type(TFourierGrid), target :: grid
      integer :: nk
      this % grid => grid
      nk = grid % nk
      allocate(this % cos_ppp(nk), this % sin_ppp(nk), this % cos_ppm(nk), this % sin_ppm(nk))
      allocate(this % cos_pmp(nk), this % sin_pmp(nk), this % cos_pmm(nk), this % sin_pmm(nk))
   end subroutine

   subroutine SinCosKR_dealloc(this)
      type(TSinCosKR) :: this
      nullify(this % grid)
      deallocate(this % cos_ppp, this % sin_ppp, this % cos_ppm, this % sin_ppm)
      deallocate(this % cos_pmp, this % sin_pmp, this % cos_pmm, this % sin_pmm)
   end subroutine

   ! charge * sin/cos(sx x kx + sy y ky + sz z kz) from the axis tables.
   subroutine SinCosKR_fill(this, sincos_xyz, charge)
      use yamc_ewald
      use yamc_shim_core
      type(TSinCosKR), target :: this
      type(TSinCosXYZ), target :: sincos_xyz
      real(8), intent(in) :: charge
      type(c_ptr) :: t(6), p(8)
      t(1) = c_loc(sincos_xyz % cos_xkx); t(2) = c_loc(sincos_xyz % cos_yky); t(3) = c_loc(sincos_xyz % cos_zkz)
      t(4) = c_loc(sincos_xyz % sin_xkx); t(5) = c_loc(sincos_xyz % sin_yky); t(6) = c_loc(sincos_xyz % sin_zkz)
      call sincoskr_ptrs(this, p)
      call shim_check(yamc_op_sincos_kr(shim_ws(this % grid % kmax), t, charge, p), 'SinCosKR_fill')
   end subroutine

   ! charge * sin/cos of the atom at (x, y, z) (SinCosXYZ_fill + SinCosKR_fill).
   subroutine SinCosKR_fill_atom(this, x, y, z, charge)
      use yamc_ewald
      use yamc_shim_core
      type(TSinCosKR), target :: this
      real(8), intent(in) :: x, y, z, charge
      type(c_ptr) :: p(8)
      call sincoskr_ptrs(this, p)
      call shim_check(yamc_op_sincos_kr_atom(shim_ws(this % grid % kmax), x, y, z, charge, p), 'SinCosKR_fill')
   end subroutine

   subroutine SinCosKR_nulify(this)
      type(TSinCosKR) :: this
      this % cos_ppp(:) = 0; this % sin_ppp(:) = 0; this % cos_ppm(:) = 0; this % sin_ppm(:) = 0
      this % cos_pmp(:) = 0; this % sin_pmp(:) = 0; this % cos_pmm(:) = 0; this % sin_pmm(:) = 0
   end subroutine

   subroutine SinCosKR_copy(dst, src)
      type(TSinCosKR) :: dst, src
      dst % cos_ppp(:) = src % cos_ppp(:); dst % sin_ppp(:) = src % sin_ppp(:)
      dst % cos_ppm(:) = src % cos_ppm(:); dst % sin_ppm(:) = src % sin_ppm(:)
      dst % cos_pmp(:) = src % cos_pmp(:); dst % sin_pmp(:) = src % sin_pmp(:)
      dst % cos_pmm(:) = src % cos_pmm(:); dst % sin_pmm(:) = src % sin_pmm(:)
   end subroutine

   subroutine SinCosKR_mulScalar(this, multiplier)
      type(TSinCosKR) :: this
      real(8), intent(in) :: multiplier
      this % cos_ppp(:) = this % cos_ppp(:) * multiplier; this % sin_ppp(:) = this % sin_ppp(:) * multiplier
      this % cos_ppm(:) = this % cos_ppm(:) * multiplier; this % sin_ppm(:) = this % sin_ppm(:) * multiplier
      this % cos_pmp(:) = this % cos_pmp(:) * multiplier; this % sin_pmp(:) = this % sin_pmp(:) * multiplier
      this % cos_pmm(:) = this % cos_pmm(:) * multiplier; this % sin_pmm(:) = this % sin_pmm(:) * multiplier
   end subroutine

   ! ---- SumSinCosKR (SumSinCosKR.f90:521-689) ----
   subroutine SumSinCosKR_alloc(this, grid)
      type(TSumSinCosKR) :: this
      type(TFourierGrid), target :: grid
      integer :: nk
      this % grid => grid
      nk = grid % nk
      allocate(this % sumcos_ppp(nk), this % sumsin_ppp(nk), this % sumcos_ppm(nk), this % sumsin_ppm(nk))
      allocate(this % sumcos_pmp(nk), this % sumsin_pmp(nk), this % sumcos_pmm(nk), this % sumsin_pmm(nk))
   end subroutine

   subroutine SumSinCosKR_dealloc(this)
      type(TSumSinCosKR) :: this
      nullify(this % grid)
      deallocate(this % sumcos_ppp, this % sumsin_ppp, this % sumcos_ppm, this % sumsin_ppm)
      deallocate(this % sumcos_pmp, this % sumsin_pmp, this % sumcos_pmm, this % sumsin_pmm)
   end subroutine

   subroutine SumSinCosKR_nulify(this)
      type(TSumSinCosKR) :: this
      this % sumcos_ppp(:) = 0; this % sumsin_ppp(:) = 0; this % sumcos_ppm(:) = 0; this % sumsin_ppm(:) = 0
      this % sumcos_pmp(:) = 0; this % sumsin_pmp(:) = 0; this % sumcos_pmm(:) = 0; this % sumsin_pmm(:) = 0
   end subroutine

   subroutine SumSinCosKR_addAtom(this, sincos_kr)
      type(TSumSinCosKR) :: this
      type(TSinCosKR) :: sincos_kr
      this % sumcos_ppp(:) = this % sumcos_ppp(:) + sincos_kr % cos_ppp(:)
      this % sumsin_ppp(:) = this % sumsin_ppp(:) + sincos_kr % sin_ppp(:)
      this % sumcos_ppm(:) = this % sumcos_ppm(:) + sincos_kr % cos_ppm(:)
      this % sumsin_ppm(:) = this % sumsin_ppm(:) + sincos_kr % sin_ppm(:)
      this % sumcos_pmp(:) = this % sumcos_pmp(:) + sincos_kr % cos_pmp(:)
      this % sumsin_pmp(:) = this % sumsin_pmp(:) + sincos_kr % sin_pmp(:)
      this % sumcos_pmm(:) = this % sumcos_pmm(:) + sincos_kr % cos_pmm(:)
      this % sumsin_pmm(:) = this % sumsin_pmm(:) + sincos_kr % sin_pmm(:)
   end subroutine

   subroutine SumSinCosKR_addSum(this, sumsincos)
      type(TSumSinCosKR) :: this, sumsincos
      this % sumcos_ppp(:) = this % sumcos_ppp(:) + sumsincos % sumcos_ppp(:)
      this % sumsin_ppp(:) = this % sumsin_ppp(:) + sumsincos % sumsin_ppp(:)
      this % sumcos_ppm(:) = this % sumcos_ppm(:) + sumsincos % sumcos_ppm(:)
      this % sumsin_ppm(:) = this % sumsin_ppm(:) + sumsincos % sumsin_ppm(:)
      this % sumcos_pmp(:) = this % sumcos_pmp(:) + sumsincos % sumcos_pmp(:)
      this % sumsin_pmp(:) = this % sumsin_pmp(:) + sumsincos % sumsin_pmp(:)
      this % sumcos_pmm(:) = this % sumcos_pmm(:) + sumsincos % sumcos_pmm(:)
      this % sumsin_pmm(:) = this % sumsin_pmm(:) + sumsincos % sumsin_pmm(:)
   end subroutine

   subroutine SumSinCosKR_subSum(this, sumsincos)
      type(TSumSinCosKR) :: this, sumsincos
      this % sumcos_ppp(:) = this % sumcos_ppp(:) - sumsincos % sumcos_ppp(:)
      this % sumsin_ppp(:) = this % sumsin_ppp(:) - sumsincos % sumsin_ppp(:)
      this % sumcos_ppm(:) = this % sumcos_ppm(:) - sumsincos % sumcos_ppm(:)
      this % sumsin_ppm(:) = this % sumsin_ppm(:) - sumsincos % sumsin_ppm(:)
      this % sumcos_pmp(:) = this % sumcos_pmp(:) - sumsincos % sumcos_pmp(:)
      this % sumsin_pmp(:) = this % sumsin_pmp(:) - sumsincos % sumsin_pmp(:)
      this % sumcos_pmm(:) = this % sumcos_pmm(:) - sumsincos % sumcos_pmm(:)
      this % sumsin_pmm(:) = this % sumsin_pmm(:) - sumsincos % sumsin_pmm(:)
   end subroutine

   subroutine SumSinCosKR_copy(dst, src)
      type(TSumSinCosKR) :: dst, src
      dst % sumcos_ppp(:) = src % sumcos_ppp(:); dst % sumsin_ppp(:) = src % sumsin_ppp(:)
      dst % sumcos_ppm(:) = src % sumcos_ppm(:); dst % sumsin_ppm(:) = src % sumsin_ppm(:)
      dst % sumcos_pmp(:) = src % sumcos_pmp(:); dst % sumsin_pmp(:) = src % sumsin_pmp(:)
      dst % sumcos_pmm(:) = src % sumcos_pmm(:); dst % sumsin_pmm(:) = src % sumsin_pmm(:)
   end subroutine

   subroutine SumSinCosKR_mulScalar(this, multiplier)
      type(TSumSinCosKR) :: this
      real(8), intent(in) :: multiplier
      this % sumcos_ppp(:) = this % sumcos_ppp(:) * multiplier; this % sumsin_ppp(:) = this % sumsin_ppp(:) * multiplier
      this % sumcos_ppm(:) = this % sumcos_ppm(:) * multiplier; this % sumsin_ppm(:) = this % sumsin_ppm(:) * multiplier
      this % sumcos_pmp(:) = this % sumcos_pmp(:) * multiplier; this % sumsin_pmp(:) = this % sumsin_pmp(:) * multiplier
      this % sumcos_pmm(:) = this % sumcos_pmm(:) * multiplier; this % sumsin_pmm(:) = this % sumsin_pmm(:) * multiplier
   end subroutine

end module SumSinCosKR
