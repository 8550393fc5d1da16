! mc_host.f90 -- a Fortran host driving the MI355X library through the
! iso_c_binding module yamc_ewald, in the order the YAMC run driver calls
! its Ewald modules: stage the system (TAtomicData/TLJTypes/parameters
! values), initmc (mc_calc_total_sums), run_mc cycles with the device move
! loop, and mc_calc_forces-style forces.
!
! Usage: mc_host case_file
! The case file (text, free format):
!   natom nmol ntype kmax precision
!   box_length alpha rmax2 xclb kext dr d_angle lambda_f lambda_c
!   natom lines:  charge hard_core_angstr mass molnum lj_type x y z
!   nmol lines:   mol_first mol_last            (0-based)
!   ntype*ntype lines: lj6 lj12 four_eps sigma6
!   ntype lines:  tail_count
!   seed stream ncycles
! Output: 'E_INIT' + 10 energies, 'ACCEPTED n', 'E_RUN' + 10 energies,
! 'FORCE_SUM' + the sum of the reference-convention forces.
program mc_host
   use iso_c_binding
   use yamc_ewald
   implicit none
   character(len=1024) :: case_file
   integer :: uin, i, natom, nmol, ntype, kmax, prec, ncycles
   integer(c_int64_t) :: seed, stream, nacc
   type(yamc_system_desc), target :: d
   real(c_double), allocatable, target :: q(:), hc(:), mass(:), x(:), y(:), z(:), lj6(:), lj12(:), feps(:), sig6(:)
   real(c_double), allocatable, target :: rx(:), ry(:), rz(:)
   integer(c_int32_t), allocatable, target :: molnum(:), ljt(:), first(:), last(:), counts(:)
   real(c_double) :: e(10)
   integer(c_int32_t) :: overlap
   type(c_ptr) :: ctx

   call get_command_argument(1, case_file)
   open(newunit=uin, file=trim(case_file), status='old', action='read')
   read(uin, *) natom, nmol, ntype, kmax, prec
   d%natom = natom; d%nmol = nmol; d%ntype = ntype; d%kmax = kmax; d%precision = prec; d%pad = 0
   read(uin, *) d%box_length, d%alpha, d%rmax2, d%xclb, d%kext, d%dr, d%d_angle, d%lambda_f, d%lambda_c
   allocate(q(natom), hc(natom), mass(natom), x(natom), y(natom), z(natom), molnum(natom), ljt(natom))
   do i = 1, natom
      read(uin, *) q(i), hc(i), mass(i), molnum(i), ljt(i), x(i), y(i), z(i)
   end do
   allocate(first(nmol), last(nmol))
   do i = 1, nmol
      read(uin, *) first(i), last(i)
   end do
   allocate(lj6(max(1, ntype * ntype)), lj12(max(1, ntype * ntype)), feps(max(1, ntype * ntype)), &
            sig6(max(1, ntype * ntype)), counts(max(1, ntype)))
   do i = 1, ntype * ntype
      read(uin, *) lj6(i), lj12(i), feps(i), sig6(i)
   end do
   do i = 1, ntype
      read(uin, *) counts(i)
   end do
   read(uin, *) seed, stream, ncycles
   close(uin)

   d%charge = c_loc(q); d%hard_core_angstr = c_loc(hc); d%mass = c_loc(mass)
   d%molnum = c_loc(molnum); d%lj_type = c_loc(ljt); d%mol_first = c_loc(first); d%mol_last = c_loc(last)
   d%lj6 = c_loc(lj6); d%lj12 = c_loc(lj12); d%four_eps = c_loc(feps); d%sigma6 = c_loc(sig6)
   d%tail_counts = c_loc(counts)

   call yamc_check(yamc_ctx_create(d, ctx), 'yamc_ctx_create')
   call yamc_check(yamc_set_positions(ctx, x, y, z), 'yamc_set_positions')
   call yamc_check(yamc_mc_init(ctx, e, overlap), 'yamc_mc_init')
   write(*, '(A,10ES25.16)') 'E_INIT ', e
   call yamc_check(yamc_mc_set_rng(ctx, seed, stream, 0_c_int64_t), 'yamc_mc_set_rng')
   call yamc_check(yamc_mc_run(ctx, int(ncycles, c_int32_t), 0_c_int32_t, c_null_ptr, nacc), 'yamc_mc_run')
   write(*, '(A,I0)') 'ACCEPTED ', nacc
   call yamc_check(yamc_mc_energies(ctx, e), 'yamc_mc_energies')
   write(*, '(A,10ES25.16)') 'E_RUN ', e
   allocate(rx(natom), ry(natom), rz(natom))
   call yamc_check(yamc_forces(ctx, c_null_ptr, c_null_ptr, c_null_ptr, c_loc(rx), c_loc(ry), c_loc(rz)), &
                   'yamc_forces')
   write(*, '(A,3ES25.16)') 'FORCE_SUM ', sum(rx), sum(ry), sum(rz)
   call yamc_check(yamc_ctx_destroy(ctx), 'yamc_ctx_destroy')
end program mc_host
