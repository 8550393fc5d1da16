! ref_driver.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Golden-vector driver for the compiled YAMC reference.  It uses the
! reference's own modules (compiled where they lie, see build_ref.sh) and
! writes their results for one staged system as labelled text:
!   * staging: MoleculeTable_fillAtomicData, LJTypes_fill, Parameters_recalc
!   * set-up: initmc (mc_calc_total_sums): beta tables, S(k) sums, energy
!     components, per-atom real-space uu and F, mu, tails
!   * per-atom k-space forces: RhoSquared_addAtom + ForceKSpace_coulomb/_LJ
!   * a move trace: movemc with the reference's RND_CONSTANT random mode and a
!     caller-chosen molecule and constant uniform per move;
!   * exchange steps: mc_xchange of given molecule pairs with a constant
!     uniform each, after the volume steps;
!   * volume steps: mcvol with a constant uniform, pressure and maximum
!     volume scaling per step.
!
! Input (free format, file given as the first argument):
!   composition_file
!   BoxLength temp sr sk alphaL eps_ext dr_a d_angle_degree lambda_f lambda_c
!   nmol
!   nmol lines: species(1-based) x y z theta phi psi
!   nmoves
!   nmoves lines: imol(1-based) u
!   nvol (optional)
!   nvol lines: u pressure[Pa] max_volume_scaling   (mcvol steps)
!   nxchg (optional)
!   nxchg lines: imol1 imol2 u   (mc_xchange steps, 1-based molecules)
!   ncycles seed psm pressure[Pa] max_volume_scaling nstream nfreq (optional)
!   nfreq lines: first last move_weight xchange_weight (1-based molecules)
!
! The last section runs the reference's own run_mc (runmc.f90:178-334) with
! its standard random mode (random.f90:25-93, the compiler's random_number):
! after randomize(seed) it writes the first nstream uniforms rand() returns
! (STREAM), re-seeds, and runs ncycles cycles, one run_mc call per cycle,
! with the volume steps (pressure_step_multiplier psm) and, with nfreq > 0,
! a frequency file (read_freq_file) that sets move and exchange weights.
! After each cycle it writes the step counters, the box length and the
! energies (CYCLE), then the next four uniforms (TAIL: they locate how many
! uniforms the run consumed) and the final state.
program ref_driver
   use SystemSettings, only : SYSTEM_STRING_LENGTH
   use parameters
   use composition
   use MoleculeTable
   use AtomicData
   use LJTypes
   use MC
   use MRandom
   use RhoSquared
   use ForceKSpace
   use SumSinCosKR
   use FloatingPoint, only : write_float
   use io, only : io_open, io_close
   use runmc, only : run_mc, ifirst, n_mv_steps, n_mv_accepted, n_vol_steps, n_vol_accepted, n_xchg_steps, &
                     n_xchg_accepted
   implicit none

   character(SYSTEM_STRING_LENGTH) :: case_file, comp_file
   Type(TComposition), target :: comp
   Type(TMoleculeTable) :: mol_table
   Type(TAtomicData), target :: atomic_data
   Type(TLJTypes), target :: lj_types
   Type(TRhoSquared) :: scratch
   Type(TSinCosKR) :: krc, krl
   integer :: nmol, i, j, natom, nk, t1, t2, nt, nmoves, imol, species, f, l, uin, nvol, ios, imol2, nxchg
   integer :: nb, hfile
   integer :: nrun, rseed, nstream, nfreq, ff, fl
   real(8), allocatable :: fkx(:), fky(:), fkz(:), fex(:), fey(:), fez(:)
   real(8) :: psm, wm, wx
   character(SYSTEM_STRING_LENGTH) :: fname
   real(8), dimension(10) :: specials = (/ 0.d0, 1.d0, -1.d0, 0.75d0, -3.5d0, 2.d0, 1.d-15, -1.d15, 123456.789d0, &
                                           -2.5d-7 /)
   real(8) :: bl, u, fxc, fyc, fzc, fxl, fyl, fzl, pvol, mvs
   logical :: accepted
   Type(TMolecule), pointer :: mol_ptr

   call getarg(1, case_file)
   open(newunit=uin, file=case_file, status="old")
   read(uin, '(A)') comp_file
   read(uin, *) bl, temp, sr, sk, alpha, external_permutivity, dr_a, d_angle_degree, xlambda_f, xlambda_c
   pressure = 0.d0
   BoxLength = bl

   comp % nalloc = 0
   call Composition_read_from_file(comp, comp_file)

   read(uin, *) nmol
   call MoleculeTable_alloc(mol_table, nmol)
   natom = 0
   do i = 1, nmol
      read(uin, *) species, mol_table % x(i), mol_table % y(i), mol_table % z(i), &
                  mol_table % theta(i), mol_table % phi(i), mol_table % psi(i)
      mol_table % mol_type(i) = comp % mol_types(species)
      call MoleculeHandler_getMolecule(mol_table % mol_type(i), mol_ptr)
      natom = natom + mol_ptr % Natoms
   end do
   mol_table % nmol = nmol

   Parameters_areRead = .TRUE.
   call Parameters_recalc(bl)

   call AtomicData_alloc(atomic_data, natom, bl, .FALSE.)
   call MoleculeTable_fillAtomicData(mol_table, atomic_data, bl, kT_kcal_mol)
   call LJTypes_fill(lj_types, atomic_data % sigma, atomic_data % epsilon, atomic_data % natom)

   call initmc(comp, mol_table, atomic_data, lj_types)

   nk = KSpace_TotalSum % grid % nk
   nt = lj_types % NType
   write(*, '(A,3I10)') 'SIZES', natom, nk, nt
   write(*, '(A,I10)') 'KMAX', kmax
   write(*, '(A,4ES26.17)') 'PARAMS', alpha, rmax2, xclb, kext
   write(*, '(A,2ES26.17)') 'PARAMS2', dr, d_angle
   call dump_real('XX', atomic_data % xx, natom)
   call dump_real('YY', atomic_data % yy, natom)
   call dump_real('ZZ', atomic_data % zz, natom)
   call dump_real('CHARGE', atomic_data % charge, natom)
   call dump_real('SIGMA', atomic_data % sigma, natom)
   call dump_real('EPSILON', atomic_data % epsilon, natom)
   call dump_real('HARDCORE', atomic_data % hard_core_angstr, natom)
   call dump_int('MOLNUM', atomic_data % molnum_by_atomnum, natom)
   call dump_int('TYPE', lj_types % type_by_index, natom)
   call dump_int('FIRST', mol_table % first_atom, nmol)
   call dump_int('LAST', mol_table % last_atom, nmol)
   call dump_real('LJ6', lj_types % LJ6Tab, nt * nt)
   call dump_real('FOUREPS', lj_types % four_epsilon_tab, nt * nt)
   call dump_real('SIGMA6', lj_types % sigma6_tab, nt * nt)
   ! LJ12Tab is only defined for t2 >= t1 (LJTypes.f90:230); dump that half
   write(*, '(A,I10)') 'LJ12UPPER', nt * (nt + 1) / 2
   do t1 = 1, nt
      do t2 = t1, nt
         write(*, '(ES26.17)') lj_types % LJ12Tab((t1 - 1) * nt + t2)
      end do
   end do
   call dump_int('KX', KSpace_TotalSum % grid % kx, nk)
   call dump_int('KY', KSpace_TotalSum % grid % ky, nk)
   call dump_int('KZ', KSpace_TotalSum % grid % kz, nk)
   call dump_int('IP', KSpace_TotalSum % grid % ip, nk)
   call dump_int('IQ', KSpace_TotalSum % grid % iq, nk)
   call dump_real('BETA', KSpace_TotalSum % beta, nk)
   call dump_real('BETA6', KSpace_TotalSum % beta6, nk)
   call dump_real('BETA12', KSpace_TotalSum % beta12, nk)
   do t1 = 1, nt
      do t2 = 1, nt
         call dump_real('BETALJ', KSpace_TotalSum % beta_LJ((t1 - 1) * nt + t2) % ptr, nk)
      end do
   end do
   call dump_sums(KSpace_TotalSum % rho_squared_total % sumsincos_coulomb, nk)
   do t1 = 1, nt
      call dump_sums(KSpace_TotalSum % rho_squared_total % sumsincos_LJ(t1), nk)
   end do
   call dump_energies('INIT')
   call dump_real('UU', RealSpace_TotalSum % uu, natom)
   call dump_real('FX', RealSpace_TotalSum % Fx, natom)
   call dump_real('FY', RealSpace_TotalSum % Fy, natom)
   call dump_real('FZ', RealSpace_TotalSum % Fz, natom)

   ! per-atom k-space forces against the total sums
   call RhoSquared_alloc(scratch, KSpace_TotalSum % grid, lj_types)
   call RhoSquared_nulify(scratch)
   call SinCosKR_alloc(krc, KSpace_TotalSum % grid)
   call SinCosKR_alloc(krl, KSpace_TotalSum % grid)
   allocate(fkx(natom), fky(natom), fkz(natom), fex(natom), fey(natom), fez(natom))
   write(*, '(A,I10)') 'FK', natom
   do i = 1, natom
      call RhoSquared_addAtom(scratch, atomic_data % xx(i), atomic_data % yy(i), atomic_data % zz(i), &
                              atomic_data % charge(i), lj_types % type_by_index(i), krc, krl)
      call ForceKSpace_coulomb(KSpace_TotalSum % beta, KSpace_TotalSum % rho_squared_total % sumsincos_coulomb, &
                               krc, fxc, fyc, fzc)
      fxl = 0; fyl = 0; fzl = 0
      if (lj_types % type_by_index(i) > 0) then
         call ForceKSpace_LJ(KSpace_TotalSum % beta_LJ, KSpace_TotalSum % rho_squared_total % sumsincos_LJ, nt, &
                             lj_types % type_by_index(i), krl, fxl, fyl, fzl)
      end if
      write(*, '(6ES26.17)') fxc, fyc, fzc, fxl, fyl, fzl
      fkx(i) = fxc + fxl
      fky(i) = fyc + fyl
      fkz(i) = fzc + fzl
   end do
   ! the forces of this configuration in mc_calc_forces' output layout
   ! (store_forces, mc_calc_forces.f90:348-368: per atom F/L x, y, z as
   ! 24+8-bit floats), with the movemc real-space convention (pairs j > i,
   ! RealSpace_TotalSum%F) instead of calc_forces' first = 1 sum, plus the
   ! k-space and external forces as calc_forces adds them
   call EwaldSumExternal_calc_mu(Ext_TotalSum)
   call external_sum_calc_forces(Ext_TotalSum % mu_x, Ext_TotalSum % mu_y, Ext_TotalSum % mu_z, &
                                 atomic_data % charge, natom, fex, fey, fez)
   fname = 'forces.ftraj'
   hfile = io_open(fname, 'b')
   do i = 1, natom
      call write_float(hfile, (RealSpace_TotalSum % Fx(i) + fkx(i) + fex(i)) / BoxLength)
      call write_float(hfile, (RealSpace_TotalSum % Fy(i) + fky(i) + fey(i)) / BoxLength)
      call write_float(hfile, (RealSpace_TotalSum % Fz(i) + fkz(i) + fez(i)) / BoxLength)
   end do
   call io_close(hfile)

   ! move trace with one constant uniform per move
   read(uin, *) nmoves
   random_method = RND_CONSTANT
   write(*, '(A,I10)') 'MOVES', nmoves
   do j = 1, nmoves
      read(uin, *) imol, u
      rnd_constant_const = u
      call movemc(imol, mol_table, accepted)
      f = mol_table % first_atom(imol)
      l = mol_table % last_atom(imol)
      write(*, '(A,2I8,L3)') 'MOVE', j, imol, accepted
      call dump_energies('AFTER')
      write(*, '(I6)') l - f + 1
      do i = f, l
         write(*, '(3ES26.17)') atomic_data % xx(i), atomic_data % yy(i), atomic_data % zz(i)
      end do
   end do
   call dump_real('UUEND', RealSpace_TotalSum % uu, natom)
   call dump_real('FXEND', RealSpace_TotalSum % Fx, natom)
   call dump_real('FYEND', RealSpace_TotalSum % Fy, natom)
   call dump_real('FZEND', RealSpace_TotalSum % Fz, natom)
   call dump_sums(KSpace_TotalSum % rho_squared_total % sumsincos_coulomb, nk)

   ! volume steps (mcvol) with one constant uniform each: lines of
   ! u  pressure[Pa]  max_volume_scaling
   read(uin, *, iostat=ios) nvol
   if (ios /= 0) nvol = 0
   write(*, '(A,I10)') 'VOLS', nvol
   do j = 1, nvol
      read(uin, *) u, pvol, mvs
      pressure = pvol
      max_volume_scaling = mvs
      call Parameters_recalc(BoxLength)
      rnd_constant_const = u
      call mcvol(mol_table, accepted)
      write(*, '(A,I8,L3,ES26.17)') 'VOL', j, accepted, BoxLength
      call dump_energies('AFTER')
      f = mol_table % first_atom(1)
      l = mol_table % last_atom(nmol)
      write(*, '(I6)') l - f + 1
      do i = f, l
         write(*, '(3ES26.17)') atomic_data % xx(i), atomic_data % yy(i), atomic_data % zz(i)
      end do
   end do

   ! exchange steps (mc_xchange) with one constant uniform each: lines of
   ! imol1 imol2 u (1-based molecule numbers)
   read(uin, *, iostat=ios) nxchg
   if (ios /= 0) nxchg = 0
   write(*, '(/,A,I10)') 'XCHGS', nxchg
   do j = 1, nxchg
      read(uin, *) imol, imol2, u
      rnd_constant_const = u
      call mc_xchange(imol, imol2, mol_table, accepted)
      write(*, '(/,A,3I8,L3)') 'XCHG', j, imol, imol2, accepted
      call dump_energies('AFTER')
      f = mol_table % first_atom(1)
      l = mol_table % last_atom(nmol)
      write(*, '(I6)') l - f + 1
      do i = f, l
         write(*, '(3ES26.17)') atomic_data % xx(i), atomic_data % yy(i), atomic_data % zz(i)
      end do
   end do
   if (nxchg > 0) then
      call dump_real('UUX', RealSpace_TotalSum % uu, natom)
      call dump_real('FXX', RealSpace_TotalSum % Fx, natom)
      call dump_real('FYX', RealSpace_TotalSum % Fy, natom)
      call dump_real('FZX', RealSpace_TotalSum % Fz, natom)
      call dump_sums(KSpace_TotalSum % rho_squared_total % sumsincos_coulomb, nk)
   end if
   ! run_mc cycles with the reference's own random stream
   read(uin, *, iostat=ios) nrun, rseed, psm, pvol, mvs, nstream, nfreq
   if (ios /= 0) nrun = 0
   if (nrun > 0) then
      if (nfreq > 0) then
         open(newunit=hfile, file='freq.txt', status='replace')
         write(hfile, '(A)') '# intervals: first-last move_weight xchange_weight'
         write(hfile, '(I6)') nfreq
         do j = 1, nfreq
            read(uin, *) ff, fl, wm, wx
            ! fixed-point weights: read_freq_file turns every '-' of the line into a blank
            write(hfile, '(I0,A,I0,2F18.10)') ff, '-', fl, wm, wx
         end do
         close(hfile)
         freq_file = 'freq.txt'
      else
         freq_file = 'none'
      end if
      pressure = pvol
      max_volume_scaling = mvs
      pressure_step_multiplier = psm
      n_store_traj_interval = 1.d9
      n_store_energy_interval = 1.d9
      call Parameters_recalc(BoxLength)
      random_method = RND_STANDARD
      call randomize(rseed)
      write(*, '(/,A,I10)') 'STREAM', nstream
      do j = 1, nstream
         write(*, '(ES26.17)') rand()
      end do
      call randomize(rseed)
      do j = 1, nrun
         ifirst = j
         call run_mc(mol_table, j)
         write(*, '(/,A,I8,6F16.1,ES26.17)') 'CYCLE', j, n_mv_steps, n_mv_accepted, n_vol_steps, &
                                             n_vol_accepted, n_xchg_steps, n_xchg_accepted, BoxLength
         call dump_energies('AFTER')
      end do
      write(*, '(/,A)') 'TAIL'
      do j = 1, 4
         write(*, '(ES26.17)') rand()
      end do
      call dump_real('RUNXX', atomic_data % xx, natom)
      call dump_real('RUNYY', atomic_data % yy, natom)
      call dump_real('RUNZZ', atomic_data % zz, natom)
      call dump_real('RUNUU', RealSpace_TotalSum % uu, natom)
      call dump_real('RUNFX', RealSpace_TotalSum % Fx, natom)
      call dump_real('RUNFY', RealSpace_TotalSum % Fy, natom)
      call dump_real('RUNFZ', RealSpace_TotalSum % Fz, natom)
      call dump_sums(KSpace_TotalSum % rho_squared_total % sumsincos_coulomb, nk)
   end if

   ! formats: molecule table of the final configuration (positions and
   ! orientations from the atoms), its binary moltab at 1, 2 and 3 bytes per
   ! sample, and 24+8-bit floats of the real-space forces / BoxLength (the
   ! mc_calc_forces output) followed by a few special values
   call MoleculeTable_calcPositionsOrientations(mol_table, atomic_data % xx, atomic_data % yy, &
                                                atomic_data % zz, BoxLength)
   write(*, '(A,I8,ES26.17)') 'MOLTAB', nmol, BoxLength
   do i = 1, nmol
      write(*, '(6ES26.17)') mol_table % x(i), mol_table % y(i), mol_table % z(i), mol_table % theta(i), &
                             mol_table % phi(i), mol_table % psi(i)
   end do
   do nb = 1, 3
      write(fname, '(A,I1,A)') 'moltab', nb, '.bin'
      hfile = io_open(fname, 'b')
      call MoleculeTable_save_binary(mol_table, hfile, nb, nb)
      call io_close(hfile)
   end do
   fname = 'floats.bin'
   hfile = io_open(fname, 'b')
   do i = 1, natom
      call write_float(hfile, RealSpace_TotalSum % Fx(i) / BoxLength)
   end do
   do i = 1, 10
      call write_float(hfile, specials(i))
   end do
   call io_close(hfile)
   write(*, '(A)') 'END'

contains

   subroutine dump_real(tag, a, n)
      character(*), intent(in) :: tag
      real(8), dimension(:), intent(in) :: a
      integer, intent(in) :: n
      integer :: k
      write(*, '(A,1X,A,I10)') 'R', tag, n
      do k = 1, n
         write(*, '(ES26.17)') a(k)
      end do
   end subroutine

   subroutine dump_int(tag, a, n)
      character(*), intent(in) :: tag
      integer, dimension(:), intent(in) :: a
      integer, intent(in) :: n
      integer :: k
      write(*, '(A,1X,A,I10)') 'I', tag, n
      do k = 1, n
         write(*, '(I12)') a(k)
      end do
   end subroutine

   subroutine dump_sums(s, n)
      Type(TSumSinCosKR), intent(in) :: s
      integer, intent(in) :: n
      call dump_real('SUMS', s % sumcos_ppp, n)
      call dump_real('SUMS', s % sumsin_ppp, n)
      call dump_real('SUMS', s % sumcos_ppm, n)
      call dump_real('SUMS', s % sumsin_ppm, n)
      call dump_real('SUMS', s % sumcos_pmp, n)
      call dump_real('SUMS', s % sumsin_pmp, n)
      call dump_real('SUMS', s % sumcos_pmm, n)
      call dump_real('SUMS', s % sumsin_pmm, n)
   end subroutine

   subroutine dump_energies(tag)
      character(*), intent(in) :: tag
      real(8) :: ut
      ut = mc_total_energy()
      write(*, '(A,1X,A)') 'E', tag
      write(*, '(5ES26.17)') ut, U_total_ew, U_total_lj, RealSpace_TotalSum % uu_ew, RealSpace_TotalSum % uu_lj6
      write(*, '(5ES26.17)') RealSpace_TotalSum % uu_lj12, KSpace_TotalSum % energy_coulomb, &
                             KSpace_TotalSum % energy_LJ, U_tail, U_ext
      write(*, '(6ES26.17)') U_tail_coulomb, U_tail_lj6, U_tail_lj12, Ext_TotalSum % mu_x, Ext_TotalSum % mu_y, &
                             Ext_TotalSum % mu_z
   end subroutine

end program ref_driver
