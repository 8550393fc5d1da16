! yamc_ewald.f90 -- Fortran binding of libyamc_ewald.so (include/yamc_ewald.h).
!
! A Fortran host (the YAMC run loop or any caller of its Ewald modules) uses
! this module to drive the MI355X library through iso_c_binding: the derived
! type yamc_system_desc is the C descriptor, every function returns the C
! status (YAMC_OK = 0, otherwise the error.f90-style code), and a context is
! an opaque type(c_ptr).  Arrays that the C ABI allows to be NULL are passed
! as type(c_ptr) values (c_loc(array) or c_null_ptr).
module yamc_ewald
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: YAMC_OK = 0, YAMC_ERROR_IO = 1, YAMC_ERROR_LIMITS = 2, &
                                YAMC_ERROR_PARAMETER = 3, YAMC_ERROR_INITIALIZATION = 4, &
                                YAMC_ERROR_WRONG_FUNCTION = 5, YAMC_ERROR_DEVICE = 6
   integer(c_int32_t), parameter :: YAMC_FP64 = 0, YAMC_FP32C = 1

   type, bind(C) :: yamc_system_desc
      integer(c_int32_t) :: natom, nmol, ntype, kmax, precision, pad
      real(c_double) :: box_length, alpha, rmax2, xclb, kext, dr, d_angle, lambda_f, lambda_c
      type(c_ptr) :: charge, hard_core_angstr, mass
      type(c_ptr) :: molnum, lj_type, mol_first, mol_last
      type(c_ptr) :: lj6, lj12, four_eps, sigma6
      type(c_ptr) :: tail_counts
   end type yamc_system_desc

   type, bind(C) :: yamc_move_rec
      integer(c_int32_t) :: imol, accepted, overlap, imol2
      real(c_double) :: du_real, du_kc, du_klj, du_ext, du_total
      real(c_double) :: f_old(3), t_old(3), f_new(3), t_new(3), delta(3)
      real(c_double) :: angle
   end type yamc_move_rec

   type, bind(C) :: yamc_rs_desc
      integer(c_int32_t) :: natom, ntype, full, nmol_atom, first_atom, j0, j1, stop_on_overlap, pad
      real(c_double) :: box_length, alpha, rmax2, xclb
      type(c_ptr) :: xx, yy, zz, charge, hard_core_angstr
      type(c_ptr) :: molnum, lj_type
      type(c_ptr) :: sigma2, four_eps, sigma6
      type(c_ptr) :: xmol, ymol, zmol
   end type yamc_rs_desc

   interface
      integer(c_int) function yamc_ctx_create(desc, ctx) bind(C, name="yamc_ctx_create")
         import :: c_int, c_ptr, yamc_system_desc
         type(yamc_system_desc), intent(in) :: desc
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function yamc_ctx_destroy(ctx) bind(C, name="yamc_ctx_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function yamc_ctx_info(ctx, device, nk) bind(C, name="yamc_ctx_info")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), intent(out) :: device, nk
      end function
      integer(c_int) function yamc_set_positions(ctx, xx, yy, zz) bind(C, name="yamc_set_positions")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: xx(*), yy(*), zz(*)
      end function
      integer(c_int) function yamc_get_positions(ctx, xx, yy, zz) bind(C, name="yamc_get_positions")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: xx(*), yy(*), zz(*)
      end function
      integer(c_int) function yamc_grid(ctx, kx, ky, kz, ip, iq) bind(C, name="yamc_grid")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), intent(out) :: kx(*), ky(*), kz(*), ip(*), iq(*)
      end function
      integer(c_int) function yamc_kspace_init(ctx, scale_beta, e_coulomb, e_lj) bind(C, name="yamc_kspace_init")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: scale_beta
         real(c_double), intent(out) :: e_coulomb, e_lj
      end function
      integer(c_int) function yamc_kspace_get_beta(ctx, beta, beta6, beta12, beta_lj) &
            bind(C, name="yamc_kspace_get_beta")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, beta, beta6, beta12, beta_lj
      end function
      integer(c_int) function yamc_kspace_get_sums(ctx, sums) bind(C, name="yamc_kspace_get_sums")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: sums(*)
      end function
      integer(c_int) function yamc_kspace_delta(ctx, imol, x, y, z, du_coulomb, du_lj) &
            bind(C, name="yamc_kspace_delta")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: imol
         real(c_double), intent(in) :: x(*), y(*), z(*)
         real(c_double), intent(out) :: du_coulomb, du_lj
      end function
      integer(c_int) function yamc_kspace_forces(ctx, first, n, x, y, z, fx, fy, fz) &
            bind(C, name="yamc_kspace_forces")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: first, n
         real(c_double), intent(in) :: x(*), y(*), z(*)
         real(c_double), intent(out) :: fx(*), fy(*), fz(*)
      end function
      integer(c_int) function yamc_rspace_full(ctx, comps, uu, fx, fy, fz, overlap) bind(C, name="yamc_rspace_full")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: comps(6)
         type(c_ptr), value :: uu, fx, fy, fz
         integer(c_int32_t), intent(out) :: overlap
      end function
      integer(c_int) function yamc_rspace_full_noalloc(ctx, comps, uu, fx, fy, fz, overlap) &
            bind(C, name="yamc_rspace_full_noalloc")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: comps(6)
         type(c_ptr), value :: uu, fx, fy, fz
         integer(c_int32_t), intent(out) :: overlap
      end function
      integer(c_int) function yamc_rspace_partial(ctx, imol, x, y, z, comps, uu, fx, fy, fz, uu_back, fx_back, &
                                                  fy_back, fz_back, overlap) bind(C, name="yamc_rspace_partial")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: imol
         real(c_double), intent(in) :: x(*), y(*), z(*)
         real(c_double), intent(out) :: comps(6)
         type(c_ptr), value :: uu, fx, fy, fz, uu_back, fx_back, fy_back, fz_back
         integer(c_int32_t), intent(out) :: overlap
      end function
      integer(c_int) function yamc_tails(ctx, self_coulomb, lj6_tail, lj12_tail) bind(C, name="yamc_tails")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: self_coulomb, lj6_tail, lj12_tail
      end function
      integer(c_int) function yamc_external_mu(ctx, mu3) bind(C, name="yamc_external_mu")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: mu3(3)
      end function
      integer(c_int) function yamc_mc_init(ctx, e, overlap) bind(C, name="yamc_mc_init")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: e(10)
         integer(c_int32_t), intent(out) :: overlap
      end function
      integer(c_int) function yamc_mc_energies(ctx, e) bind(C, name="yamc_mc_energies")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: e(10)
      end function
      integer(c_int) function yamc_mc_real_arrays(ctx, uu, fx, fy, fz) bind(C, name="yamc_mc_real_arrays")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, uu, fx, fy, fz
      end function
      integer(c_int) function yamc_mc_set_rng(ctx, seed, stream, counter) bind(C, name="yamc_mc_set_rng")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: seed, stream, counter
      end function
      integer(c_int) function yamc_mc_set_rng_list(ctx, u, n) bind(C, name="yamc_mc_set_rng_list")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         integer(c_int64_t), value :: n
      end function
      integer(c_int) function yamc_mc_get_rng(ctx, counter) bind(C, name="yamc_mc_get_rng")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), intent(out) :: counter
      end function
      integer(c_int) function yamc_mc_run(ctx, ncycles, resync_every, recs, n_accepted) bind(C, name="yamc_mc_run")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: ncycles, resync_every
         type(c_ptr), value :: recs
         integer(c_int64_t), intent(out) :: n_accepted
      end function
      integer(c_int) function yamc_mc_run_const(ctx, nmoves, imol, u, recs, n_accepted) &
            bind(C, name="yamc_mc_run_const")
         import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nmoves
         integer(c_int32_t), intent(in) :: imol(*)
         real(c_double), intent(in) :: u(*)
         type(c_ptr), value :: recs
         integer(c_int64_t), intent(out) :: n_accepted
      end function
      integer(c_int) function yamc_resync(ctx) bind(C, name="yamc_resync")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function yamc_mc_volume_move(ctx, pressure_kT_A3, log_max_volume_scaling, u, accepted, &
            box_length) bind(C, name="yamc_mc_volume_move")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: pressure_kT_A3, log_max_volume_scaling
         type(c_ptr), value :: u
         integer(c_int32_t), intent(out) :: accepted
         real(c_double), intent(out) :: box_length
      end function
      integer(c_int) function yamc_mc_run_npt(ctx, ncycles, pressure_step_multiplier, pressure_kT_A3, &
            log_max_volume_scaling, n_accepted, n_vol_steps, n_vol_accepted) bind(C, name="yamc_mc_run_npt")
         import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: ncycles
         real(c_double), value :: pressure_step_multiplier, pressure_kT_A3, log_max_volume_scaling
         integer(c_int64_t), intent(out) :: n_accepted, n_vol_steps, n_vol_accepted
      end function
      integer(c_int) function yamc_mc_set_frequencies(ctx, move_weights, xchange_weights) &
            bind(C, name="yamc_mc_set_frequencies")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, move_weights, xchange_weights
      end function
      integer(c_int) function yamc_mc_xchange(ctx, imol1, imol2, u, accepted, rec) bind(C, name="yamc_mc_xchange")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: imol1, imol2
         type(c_ptr), value :: u
         integer(c_int32_t), intent(out) :: accepted
         type(c_ptr), value :: rec
      end function
      integer(c_int) function yamc_set_box(ctx, box_length, scale_positions, overlap) bind(C, name="yamc_set_box")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: box_length
         integer(c_int32_t), value :: scale_positions
         integer(c_int32_t), intent(out) :: overlap
      end function
      integer(c_int) function yamc_mc_commit_move(ctx, imol, x, y, z, du_c, du_lj, overlap) &
            bind(C, name="yamc_mc_commit_move")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: imol
         real(c_double), intent(in) :: x(*), y(*), z(*)
         real(c_double), value :: du_c, du_lj
         integer(c_int32_t), intent(out) :: overlap
      end function
      integer(c_int) function yamc_box_length(ctx, box_length) bind(C, name="yamc_box_length")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: box_length
      end function
      integer(c_int) function yamc_batch_run(ctxs, n, ncycles, resync_every, n_accepted) &
            bind(C, name="yamc_batch_run")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), intent(in) :: ctxs(*)
         integer(c_int32_t), value :: n, ncycles, resync_every
         integer(c_int64_t), intent(out) :: n_accepted(*)
      end function
      integer(c_int) function yamc_batch_create(ctxs, n, ncycles, resync_every, batch) &
            bind(C, name="yamc_batch_create")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), intent(in) :: ctxs(*)
         integer(c_int32_t), value :: n, ncycles, resync_every
         type(c_ptr), intent(out) :: batch
      end function
      integer(c_int) function yamc_batch_launch(batch, n_accepted) bind(C, name="yamc_batch_launch")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: batch
         integer(c_int64_t), intent(out) :: n_accepted(*)
      end function
      integer(c_int) function yamc_batch_destroy(batch) bind(C, name="yamc_batch_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
      end function
      integer(c_int) function yamc_forces(ctx, gx, gy, gz, rx, ry, rz) bind(C, name="yamc_forces")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, gx, gy, gz, rx, ry, rz
      end function
      integer(c_int) function yamc_rdf(ctx, mol_species, nspecies, dr_angstr, rmax_angstr, nbins, counts) &
            bind(C, name="yamc_rdf")
         import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), intent(in) :: mol_species(*)
         integer(c_int32_t), value :: nspecies, nbins
         real(c_double), value :: dr_angstr, rmax_angstr
         integer(c_int64_t), intent(inout) :: counts(*)
      end function
      ! ---- per-procedure surface (fortran/shim/ forwards the reference's
      ! module procedures to these); 8-pointer arguments are the 8 sin/cos
      ! arrays of a set, beta_LJ the ntype*ntype row pointers
      integer(c_int) function yamc_ws_create(kmax, ws) bind(C, name="yamc_ws_create")
         import :: c_int, c_ptr, c_int32_t
         integer(c_int32_t), value :: kmax
         type(c_ptr), intent(out) :: ws
      end function
      integer(c_int) function yamc_ws_destroy(ws) bind(C, name="yamc_ws_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ws
      end function
      integer(c_int) function yamc_ws_grid(ws, nk, kx, ky, kz, ip, iq, xk, xk2) bind(C, name="yamc_ws_grid")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ws
         integer(c_int32_t), intent(out) :: nk
         type(c_ptr), value :: kx, ky, kz, ip, iq, xk, xk2
      end function
      integer(c_int) function yamc_op_beta(ws, alpha, xclb, beta, beta6, beta12) bind(C, name="yamc_op_beta")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ws
         real(c_double), value :: alpha, xclb
         real(c_double), intent(out) :: beta(*), beta6(*), beta12(*)
      end function
      integer(c_int) function yamc_op_beta_lj(ws, lj6, lj12, beta6, beta12, out) bind(C, name="yamc_op_beta_lj")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ws
         real(c_double), value :: lj6, lj12
         real(c_double), intent(in) :: beta6(*), beta12(*)
         type(c_ptr), value :: out
      end function
      integer(c_int) function yamc_op_sincos_xyz(ws, x, y, z, tab6) bind(C, name="yamc_op_sincos_xyz")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ws
         real(c_double), value :: x, y, z
         type(c_ptr), intent(in) :: tab6(6)
      end function
      integer(c_int) function yamc_op_sincos_kr(ws, tab6, q, kr8) bind(C, name="yamc_op_sincos_kr")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ws
         type(c_ptr), intent(in) :: tab6(6)
         real(c_double), value :: q
         type(c_ptr), intent(in) :: kr8(8)
      end function
      integer(c_int) function yamc_op_sincos_kr_atom(ws, x, y, z, q, kr8) bind(C, name="yamc_op_sincos_kr_atom")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ws
         real(c_double), value :: x, y, z, q
         type(c_ptr), intent(in) :: kr8(8)
      end function
      integer(c_int) function yamc_op_sums(ws, n, x, y, z, q, lj_type, ntype, sums, present) &
            bind(C, name="yamc_op_sums")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ws
         integer(c_int32_t), value :: n, ntype
         real(c_double), intent(in) :: x(*), y(*), z(*), q(*)
         integer(c_int32_t), intent(in) :: lj_type(*)
         type(c_ptr), intent(in) :: sums(*)
         integer(c_int32_t), intent(out) :: present(*)
      end function
      integer(c_int) function yamc_op_rho_squared(ws, s8, rho2) bind(C, name="yamc_op_rho_squared")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ws
         type(c_ptr), intent(in) :: s8(8)
         real(c_double), intent(out) :: rho2(*)
      end function
      integer(c_int) function yamc_op_energy_coulomb(ws, beta, rho2, e) bind(C, name="yamc_op_energy_coulomb")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ws
         real(c_double), intent(in) :: beta(*), rho2(*)
         real(c_double), intent(out) :: e
      end function
      integer(c_int) function yamc_op_energy_lj(ws, ntype, beta_lj, sums_lj, e) bind(C, name="yamc_op_energy_lj")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ws
         integer(c_int32_t), value :: ntype
         type(c_ptr), intent(in) :: beta_lj(*), sums_lj(*)
         real(c_double), intent(out) :: e
      end function
      integer(c_int) function yamc_op_du_coulomb(ws, beta, s8, d8, du) bind(C, name="yamc_op_du_coulomb")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ws
         real(c_double), intent(in) :: beta(*)
         type(c_ptr), intent(in) :: s8(8), d8(8)
         real(c_double), intent(out) :: du
      end function
      integer(c_int) function yamc_op_du_lj(ws, ntype, beta_lj, sums_lj, delta_lj, present, du) &
            bind(C, name="yamc_op_du_lj")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ws
         integer(c_int32_t), value :: ntype
         type(c_ptr), intent(in) :: beta_lj(*), sums_lj(*), delta_lj(*)
         integer(c_int32_t), intent(in) :: present(*)
         real(c_double), intent(out) :: du
      end function
      integer(c_int) function yamc_op_force_coulomb(ws, beta, s8, kr8, f3) bind(C, name="yamc_op_force_coulomb")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ws
         real(c_double), intent(in) :: beta(*)
         type(c_ptr), intent(in) :: s8(8), kr8(8)
         real(c_double), intent(out) :: f3(3)
      end function
      integer(c_int) function yamc_op_force_lj(ws, ntype, ityp, beta_lj, sums_lj, kr8, f3) &
            bind(C, name="yamc_op_force_lj")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ws
         integer(c_int32_t), value :: ntype, ityp
         type(c_ptr), intent(in) :: beta_lj(*), sums_lj(*), kr8(8)
         real(c_double), intent(out) :: f3(3)
      end function
      integer(c_int) function yamc_op_rspace(ws, d, uu, fx, fy, fz, uu_back, fx_back, fy_back, fz_back, comps, &
                                             overlap) bind(C, name="yamc_op_rspace")
         import :: c_int, c_ptr, c_int32_t, c_double, yamc_rs_desc
         type(c_ptr), value :: ws
         type(yamc_rs_desc), intent(in) :: d
         type(c_ptr), value :: uu, fx, fy, fz, uu_back, fx_back, fy_back, fz_back
         real(c_double), intent(out) :: comps(6)
         integer(c_int32_t), intent(out) :: overlap
      end function
      integer(c_int) function yamc_op_pair(ws, alpha, xclb, r2, same_molecule, qi, qj, lj, sigma6, four_eps, out) &
            bind(C, name="yamc_op_pair")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ws
         real(c_double), value :: alpha, xclb, r2, qi, qj, sigma6, four_eps
         integer(c_int32_t), value :: same_molecule, lj
         real(c_double), intent(out) :: out(14)
      end function
      integer(c_int) function yamc_op_tails(ws, natom, charge, ntype, lj6, lj12, counts, alpha, xclb, self_coulomb, &
                                            lj6_tail, lj12_tail) bind(C, name="yamc_op_tails")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ws
         integer(c_int32_t), value :: natom, ntype
         real(c_double), intent(in) :: charge(*), lj6(*), lj12(*)
         integer(c_int32_t), intent(in) :: counts(*)
         real(c_double), value :: alpha, xclb
         real(c_double), intent(out) :: self_coulomb, lj6_tail, lj12_tail
      end function
      integer(c_int) function yamc_set_block_size(ctx, block) bind(C, name="yamc_set_block_size")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: block
      end function
      integer(c_int) function yamc_set_chain_groups(ctx, groups) bind(C, name="yamc_set_chain_groups")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: groups
      end function
      integer(c_int) function yamc_set_cell_div(ctx, div) bind(C, name="yamc_set_cell_div")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: div
      end function
      integer(c_int) function yamc_last_kernel_ms(ctx, ms) bind(C, name="yamc_last_kernel_ms")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: ms
      end function
      integer(c_int) function yamc_phase_times(ctx, us) bind(C, name="yamc_phase_times")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: us(9)
      end function
      type(c_ptr) function yamc_last_error_c() bind(C, name="yamc_last_error")
         import :: c_ptr
      end function
   end interface

contains

   ! The library's last error message as a Fortran string.
   function yamc_last_error() result(msg)
      character(len=:), allocatable :: msg
      character(kind=c_char), pointer :: p(:)
      type(c_ptr) :: cp
      integer :: n
      interface
         integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
         end function
      end interface
      cp = yamc_last_error_c()
      if (.not. c_associated(cp)) then
         msg = ""
         return
      end if
      n = int(c_strlen(cp))
      call c_f_pointer(cp, p, [n])
      allocate(character(len=n) :: msg)
      msg = transfer(p(1:n), msg)
   end function

   ! Stops with the library's message unless rc is YAMC_OK, as the
   ! reference's error_throw does for an uncaught code (error.f90:112-130).
   subroutine yamc_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: where
      if (rc /= YAMC_OK) then
         write(*, '(A,A,A,I0,A,A)') 'yamc: ', where, ' failed with status ', rc, ': ', yamc_last_error()
         stop 1
      end if
   end subroutine

end module yamc_ewald
