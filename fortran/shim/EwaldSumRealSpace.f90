! EwaldSumRealSpace.f90 -- drop-in for YAMC's EwaldSumRealSpace module
! (EwaldSumRealSpace.f90:1-505): the real-space Ewald sum, full or over one
! molecule's atoms.  The fields keep the reference's type and pointer
! aliasing (full sum: uu_back => uu, F*_back => F*, xmol => atomic_data % xx);
! EwaldSumRealSpace_calc runs on the device (yamc_op_rspace) and leaves the
! same fields as the reference's loop, including the early return at the
! first overlapping pair.
module EwaldSumRealSpace
   use iso_c_binding
   use AtomicData
   use LJTypes
   implicit none

   type :: TEwaldSumRealSpace
      type(TAtomicData), pointer :: atomic_data
      type(TLJTypes), pointer :: lj_types
      integer :: first_atom, last_atom
      integer :: natom, nalloc
      real(8) :: uu_ew, uu_lj6, uu_lj12
      real(8) :: uu_ew_intra, uu_lj6_intra, uu_lj12_intra
      real(8), dimension(:), pointer :: uu
      real(8), dimension(:), pointer :: uu_back
      real(8), dimension(:), pointer :: Fx, Fy, Fz
      real(8), dimension(:), pointer :: Fx_back, Fy_back, Fz_back
      real(8), dimension(:), pointer :: xmol, ymol, zmol
      logical :: full_sum
   end type

contains

   ! EwaldSumRealSpace.f90:70-103
   subroutine EwaldSumRealSpace_alloc_full(this, atomic_data, lj_types)
      type(TEwaldSumRealSpace) :: this
      type(TAtomicData), intent(in), target :: atomic_data
      type(TLJTypes), intent(in), target :: lj_types
      integer :: ntotal
      this % atomic_data => atomic_data
      this % lj_types => lj_types
      this % full_sum = .TRUE.
      ntotal = atomic_data % Natom
      this % natom = ntotal
      this % xmol => atomic_data % xx
      this % ymol => atomic_data % yy
      this % zmol => atomic_data % zz
      allocate(this % uu(ntotal))
      this % uu_back => this % uu
      allocate(this % Fx(ntotal), this % Fy(ntotal), this % Fz(ntotal))
      this % Fx_back => this % Fx
      this % Fy_back => this % Fy
      this % Fz_back => this % Fz
   end subroutine

   ! EwaldSumRealSpace.f90:105-131 (nalloc starts at 0 here)
   subroutine EwaldSumRealSpace_alloc_partial(this, atomic_data, lj_types)
      type(TEwaldSumRealSpace) :: this
      type(TAtomicData), intent(in), target :: atomic_data
      type(TLJTypes), intent(in), target :: lj_types
      integer :: ntotal
      this % atomic_data => atomic_data
      this % lj_types => lj_types
      this % full_sum = .FALSE.
      this % natom = 0
      this % nalloc = 0
      ntotal = atomic_data % Natom
      nullify(this % xmol, this % ymol, this % zmol)
      allocate(this % uu_back(ntotal))
      allocate(this % Fx_back(ntotal), this % Fy_back(ntotal), this % Fz_back(ntotal))
   end subroutine

   ! EwaldSumRealSpace.f90:133-185
   subroutine EwaldSumRealSpace_set_molecule(this, xmol, ymol, zmol, first_atom, last_atom)
      use error
      type(TEwaldSumRealSpace) :: this
      real(8), dimension(:), intent(in), target :: xmol, ymol, zmol
      integer, intent(in) :: first_atom, last_atom
      integer :: nmolatom
      if (this % full_sum) then
         write(error_message, *) 'EwaldSimRealSpace_set_molecule: function not valid for the full  sum!'
         call error_throw(ERROR_WRONG_FUNCTION)
         return
      end if
      this % xmol => xmol
      this % ymol => ymol
      this % zmol => zmol
      this % first_atom = first_atom
      this % last_atom = last_atom
      nmolatom = last_atom - first_atom + 1
      if (this % nalloc < nmolatom) then
         if (this % nalloc > 0) deallocate(this % Fx, this % Fy, this % Fz, this % uu)
         allocate(this % uu(nmolatom), this % Fx(nmolatom), this % Fy(nmolatom), this % Fz(nmolatom))
         this % nalloc = nmolatom
      end if
      this % natom = nmolatom
   end subroutine

   ! EwaldSumRealSpace.f90:187-204
   subroutine EwaldSumRealSpace_dealloc_full(this)
      use error
      type(TEwaldSumRealSpace) :: this
      if (.not. this % full_sum) then
         write(error_message, *) 'EwaldSumRealSpace_dealloc_full: cannot deallocate partial sum! use dealloc_partial instead! '
         call error_throw(ERROR_WRONG_FUNCTION)
         return
      end if
      deallocate(this % uu, this % Fx, this % Fy, this % Fz)
   end subroutine

   ! EwaldSumRealSpace.f90:206-231
   subroutine EwaldSumRealSpace_dealloc_partial(this)
      use error
      type(TEwaldSumRealSpace) :: this
      if (this % full_sum) then
         write(error_message, *) 'EwaldSumRealSpace_dealloc_partial: cannot deallocate full_sum! use dealloc_full instead! '
         call error_throw(ERROR_WRONG_FUNCTION)
         return
      end if
      deallocate(this % Fx_back, this % Fy_back, this % Fz_back, this % uu_back)
      if (this % nalloc > 0) deallocate(this % Fx, this % Fy, this % Fz, this % uu)
   end subroutine

   ! EwaldSumRealSpace.f90:233-498: uu and F over the rows (all atoms, or the
   ! molecule's atoms at xmol/ymol/zmol), the back arrays over all atoms,
   ! the energy components; pairs j in first..last when given.  Each
   ! overlapping pair the loop meets prints an 'O'; unless
   ! dont_stop_on_overlap is .TRUE. the sum returns at the first one, with the
   ! fields as the loop left them there.
   subroutine EwaldSumRealSpace_calc(this, overlap, first, last, dont_stop_on_overlap)
      use parameters, only : rmax2, BoxLength, alpha, xclb, kmax
      use yamc_ewald
      use yamc_shim_core
      type(TEwaldSumRealSpace), target :: this
      logical, intent(out) :: overlap
      integer, intent(in), optional :: first, last
      logical, intent(in), optional :: dont_stop_on_overlap
      type(TAtomicData), pointer :: ad
      type(TLJTypes), pointer :: lt
      type(yamc_rs_desc) :: d
      real(c_double) :: comps(6)
      integer(c_int32_t) :: nov
      integer :: k
      ad => this % atomic_data
      lt => this % lj_types
      d % natom = ad % Natom
      d % ntype = lt % NType
      d % full = merge(1, 0, this % full_sum)
      d % nmol_atom = this % natom
      d % first_atom = 0
      if (.not. this % full_sum) d % first_atom = this % first_atom - 1
      d % j0 = -1
      d % j1 = -1
      if (present(first)) d % j0 = first - 1
      if (present(last)) d % j1 = last - 1
      d % stop_on_overlap = 1
      if (present(dont_stop_on_overlap)) then
         if (dont_stop_on_overlap) d % stop_on_overlap = 0
      end if
      d % pad = 0
      d % box_length = BoxLength
      d % alpha = alpha
      d % rmax2 = rmax2
      d % xclb = xclb
      d % xx = c_loc(ad % xx(1)); d % yy = c_loc(ad % yy(1)); d % zz = c_loc(ad % zz(1))
      d % charge = c_loc(ad % charge(1))
      d % hard_core_angstr = c_loc(ad % hard_core_angstr(1))
      d % molnum = c_loc(ad % molnum_by_atomnum(1))
      d % lj_type = c_loc(lt % type_by_index(1))
      d % sigma2 = c_null_ptr; d % four_eps = c_null_ptr; d % sigma6 = c_null_ptr
      if (lt % NType > 0) then
         d % sigma2 = c_loc(lt % sigma2_tab(1))
         d % four_eps = c_loc(lt % four_epsilon_tab(1))
         d % sigma6 = c_loc(lt % sigma6_tab(1))
      end if
      d % xmol = c_null_ptr; d % ymol = c_null_ptr; d % zmol = c_null_ptr
      if (.not. this % full_sum) then
         d % xmol = c_loc(this % xmol(1)); d % ymol = c_loc(this % ymol(1)); d % zmol = c_loc(this % zmol(1))
      end if
      this % uu(:) = 0; this % Fx(:) = 0; this % Fy(:) = 0; this % Fz(:) = 0  ! rows past natom stay 0
      call shim_check(yamc_op_rspace(shim_ws(kmax), d, c_loc(this % uu(1)), c_loc(this % Fx(1)), c_loc(this % Fy(1)), &
                                     c_loc(this % Fz(1)), c_loc(this % uu_back(1)), c_loc(this % Fx_back(1)), &
                                     c_loc(this % Fy_back(1)), c_loc(this % Fz_back(1)), comps, nov), &
                      'EwaldSumRealSpace_calc')
      do k = 1, nov
         write(*, '(A,$)') 'O'
      end do
      overlap = nov > 0
      this % uu_ew = comps(1); this % uu_lj6 = comps(2); this % uu_lj12 = comps(3)
      this % uu_ew_intra = comps(4); this % uu_lj6_intra = comps(5); this % uu_lj12_intra = comps(6)
   end subroutine

end module EwaldSumRealSpace
