! iso_c_binding interface of libsis_gpu.so (include/sis_gpu.h) for the
! Fortran host.  A host that keeps the reference's module state calls
! sis_gpu_set_state / sis_gpu_get_state around the calls that replace
! neighbor_list (src/neighbor_list.F90:1), force (src/force.F90:1),
! energy_sumup (src/energy_sumup.F90:1) and the job_md loop body
! (src/job_md.F90:360-460).  Struct layouts match include/sis_gpu.h.

module sis_gpu_iface

   use, intrinsic :: iso_c_binding
   implicit none

   integer(c_int), parameter :: SIS_OK = 0

   type, bind(C) :: sis_bp_para
      real(c_double) :: cutoff_ddist, U0, bond_k, bond_r
      real(c_double) :: angl_k1, angl_k2, angl_k3, angl_k4
      real(c_double) :: angl_theta1, angl_theta2, angl_theta3, angl_theta4
      real(c_double) :: dihd_k1, dihd_k2, dihd_phi1, dihd_phi2
   end type sis_bp_para

   type, bind(C) :: sis_params
      real(c_double) :: bond_k, bond_r0
      real(c_double) :: angl_k, angl_t0
      real(c_double) :: dih_k, dih_w, dih_p0
      real(c_double) :: wca_sigma, wca_eps
      real(c_double) :: ele_cutoff, length_per_charge, ionic_strength
      real(c_double) :: bp_cutoff_energy
      real(c_double) :: bp_dH0, bp_dS0, bp_dHend0, bp_dSend0, bp_coef_dG
      real(c_double) :: nn_dH(21), nn_dS(21)
      real(c_double) :: nnend_dH(6), nnend_dS(6)
      type(sis_bp_para) :: bp(3)
      real(c_double) :: tempK
      real(c_double) :: dt, viscosity_pas, nl_margin
      real(c_double) :: box(3)
      integer(c_int32_t) :: angl_reb, dih_type
      integer(c_int32_t) :: ele_on, ele_cutoff_type, ele_excl_bond, ele_excl_angl, dummy_has_charge
      integer(c_int32_t) :: bp_model, bp_min_loop, max_bp_per_nt, nstep_bp_mc, nnend
      integer(c_int32_t) :: pbc, integrator
   end type sis_params

   type, bind(C) :: sis_topo
      integer(c_int32_t) :: nchains, nmp
      type(c_ptr) :: nmp_chain, moltype, seq, no_charge
      integer(c_int32_t) :: n_no_charge, pad
   end type sis_topo

   type, bind(C) :: sis_options
      integer(c_int32_t) :: nrep, precision, parity, mode, block, max_nbr
      integer(c_int32_t) :: rebuild_every, exchange_every
      integer(c_int64_t) :: seed
      integer(c_int32_t) :: grid, flags
   end type sis_options

   ! external terms (src/force.F90:62-75): c_loc of the reference's arrays
   ! (twz_DCF_pairs, twz_DCF_forces, ..., rest_sigb_para), c_null_ptr when unused
   type, bind(C) :: sis_external
      integer(c_int32_t) :: stage, n_dcf, n_fr, rg, n_rest, pad
      real(c_double) :: stage_sigma, stage_eps, rg_k
      type(c_ptr) :: rg_0, dcf_pairs, dcf_forces, fr_pairs, fr_k, fr_init, fr_velo
      type(c_ptr) :: rest_id, rest_rcut, rest_para
   end type sis_external

   interface
      integer(c_int) function sis_gpu_create(topo, prm, opt, ctx) bind(C, name="sis_gpu_create")
         import :: c_int, c_ptr, sis_topo, sis_params, sis_options
         type(sis_topo), intent(in) :: topo
         type(sis_params), intent(in) :: prm
         type(sis_options), intent(in) :: opt
         type(c_ptr), intent(out) :: ctx
      end function

      integer(c_int) function sis_gpu_destroy(ctx) bind(C, name="sis_gpu_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      type(c_ptr) function sis_gpu_last_error() bind(C, name="sis_gpu_last_error")
         import :: c_ptr
      end function

      ! xyz, velos, accels: real(8) (3, nmp, nrep) as var_state (src/var_state.F90:26-28)
      integer(c_int) function sis_gpu_set_state(ctx, xyz, velos, accels) bind(C, name="sis_gpu_set_state")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, xyz, velos, accels
      end function

      integer(c_int) function sis_gpu_get_state(ctx, xyz, velos, accels) bind(C, name="sis_gpu_get_state")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, xyz, velos, accels
      end function

      ! coordinates of one replica, lists kept (the caller's own skin test
      ! decides rebuilds, src/job_md.F90:372-395)
      integer(c_int) function sis_gpu_set_positions(ctx, irep, xyz) bind(C, name="sis_gpu_set_positions")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: irep
         type(c_ptr), value :: xyz
      end function

      integer(c_int) function sis_gpu_set_rep_conditions(ctx, tempK, ionic) &
            bind(C, name="sis_gpu_set_rep_conditions")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, tempK, ionic
      end function

      ! replaces: call neighbor_list(irep)
      integer(c_int) function sis_gpu_rebuild(ctx, irep) bind(C, name="sis_gpu_rebuild")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: irep
      end function

      ! replaces: call list_bp(); call list_exv(); if (flg_ele) call list_ele()
      !           (src/main.F90:144-154, jobs other than MD)
      integer(c_int) function sis_gpu_analysis_lists(ctx) bind(C, name="sis_gpu_analysis_lists")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      ! replaces: call force(irep, forces)
      integer(c_int) function sis_gpu_force(ctx, irep, bp_mc, forces) bind(C, name="sis_gpu_force")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: irep, bp_mc
         type(c_ptr), value :: forces
      end function

      ! replaces: call energy_sumup(irep, tempK, energies)
      integer(c_int) function sis_gpu_energy(ctx, irep, tempK, flags, energies) bind(C, name="sis_gpu_energy")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: irep, flags
         real(c_double), value :: tempK
         real(c_double), intent(out) :: energies(0:10)
      end function

      ! energy_sumup without device allocation (no reference counterpart;
      ! for callers that cannot allocate): flags bit1 only
      integer(c_int) function sis_gpu_energy_noalloc(ctx, irep, tempK, flags, energies) &
            bind(C, name="sis_gpu_energy_noalloc")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: irep, flags
         real(c_double), value :: tempK
         real(c_double), intent(out) :: energies(0:10)
      end function

      integer(c_int) function sis_gpu_get_pair_count(ctx, irep, kind, n) bind(C, name="sis_gpu_get_pair_count")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: irep, kind
         integer(c_int64_t), intent(out) :: n
      end function

      integer(c_int) function sis_gpu_get_pairs(ctx, irep, kind, out, cap, n) bind(C, name="sis_gpu_get_pairs")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: irep, kind
         type(c_ptr), value :: out
         integer(c_int64_t), value :: cap
         integer(c_int64_t), intent(out) :: n
      end function

      integer(c_int) function sis_gpu_get_bp(ctx, irep, bp_mp, bp_coef, status, ene, cap, n) &
            bind(C, name="sis_gpu_get_bp")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: irep
         type(c_ptr), value :: bp_mp, bp_coef, status, ene
         integer(c_int64_t), value :: cap
         integer(c_int64_t), intent(out) :: n
      end function

      integer(c_int) function sis_gpu_set_rng(ctx, seed) bind(C, name="sis_gpu_set_rng")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: seed
      end function

      integer(c_int) function sis_gpu_set_rng_tape(ctx, irep, words, n) bind(C, name="sis_gpu_set_rng_tape")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: irep
         type(c_ptr), value :: words
         integer(c_int64_t), value :: n
      end function

      ! words of the tape used so far: the caller advances mts(irep) by them
      integer(c_int) function sis_gpu_get_rng_tape_pos(ctx, irep, pos) bind(C, name="sis_gpu_get_rng_tape_pos")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: irep
         integer(c_int64_t), intent(out) :: pos
      end function

      integer(c_int) function sis_gpu_init_velocities(ctx) bind(C, name="sis_gpu_init_velocities")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      ! replaces the body of the job_md time loop for nsteps steps
      integer(c_int) function sis_gpu_step(ctx, nsteps, rebuilds) bind(C, name="sis_gpu_step")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: nsteps
         integer(c_int64_t), intent(out) :: rebuilds
      end function

      integer(c_int) function sis_gpu_exchange(ctx, replica_energies, accepted) bind(C, name="sis_gpu_exchange")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, replica_energies, accepted
      end function

      ! replaces: make_replica_exchange_pair_tb (src/var_replica.F90:162-259)
      ! for ndim_replica > 1; nvalues = nrep(ivar) of the variables with ladders
      integer(c_int) function sis_gpu_set_rep_dims(ctx, ndim, nvalues) bind(C, name="sis_gpu_set_rep_dims")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: ndim
         integer(c_int32_t), intent(in) :: nvalues(*)
      end function

      integer(c_int) function sis_gpu_get_labels(ctx, rep2lab) bind(C, name="sis_gpu_get_labels")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, rep2lab
      end function

      ! trajectory output (DCD, src/dcd.F90 layout); paths is an array of
      ! nrep c_ptr to NUL-terminated file names
      integer(c_int) function sis_gpu_dcd_open(ctx, paths, fix_com_origin, nstep_save) &
            bind(C, name="sis_gpu_dcd_open")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx, paths
         integer(c_int32_t), value :: fix_com_origin
         integer(c_int64_t), value :: nstep_save
      end function

      integer(c_int) function sis_gpu_dcd_write(ctx) bind(C, name="sis_gpu_dcd_write")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function sis_gpu_dcd_close(ctx) bind(C, name="sis_gpu_dcd_close")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      ! base-pair output (src/write_bp.F90 layout); kind 0 = .bp, 1 = .bpall
      integer(c_int) function sis_gpu_bp_open(ctx, paths, kind, nstep_save) bind(C, name="sis_gpu_bp_open")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx, paths
         integer(c_int32_t), value :: kind
         integer(c_int64_t), value :: nstep_save
      end function

      integer(c_int) function sis_gpu_bp_write(ctx) bind(C, name="sis_gpu_bp_write")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function sis_gpu_bp_close(ctx) bind(C, name="sis_gpu_bp_close")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      ! energy lines (.out, src/write_out.F90 format)
      integer(c_int) function sis_gpu_out_open(ctx, paths, nstep_save) bind(C, name="sis_gpu_out_open")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx, paths
         integer(c_int64_t), value :: nstep_save
      end function

      integer(c_int) function sis_gpu_out_write(ctx) bind(C, name="sis_gpu_out_write")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function sis_gpu_out_close(ctx) bind(C, name="sis_gpu_out_close")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      ! tweezers' force-ramp lines (.twz)
      integer(c_int) function sis_gpu_twz_open(ctx, path, nstep_save) bind(C, name="sis_gpu_twz_open")
         import :: c_int, c_ptr, c_int64_t, c_char
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int64_t), value :: nstep_save
      end function

      integer(c_int) function sis_gpu_twz_close(ctx) bind(C, name="sis_gpu_twz_close")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      ! replica exchange over the tweezers' constant force
      integer(c_int) function sis_gpu_set_rep_dcf(ctx, k, direction) bind(C, name="sis_gpu_set_rep_dcf")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, k, direction
      end function

      ! replica exchange over the Rg-bias target
      integer(c_int) function sis_gpu_set_rep_rg(ctx, rg0) bind(C, name="sis_gpu_set_rep_rg")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, rg0
      end function

      ! timed Rg bias (read_timed_bias_rg)
      integer(c_int) function sis_gpu_set_rg_schedule(ctx, n, steps, k, r0) bind(C, name="sis_gpu_set_rg_schedule")
         import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: n
         integer(c_int64_t), intent(in) :: steps(*)
         real(c_double), intent(in) :: k(*), r0(*)
      end function

      ! box resizing ([PBC_box] resize_step / resize_change / resize_target)
      integer(c_int) function sis_gpu_set_box_resize(ctx, resize_step, change, target) &
            bind(C, name="sis_gpu_set_box_resize")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx, change, target
         integer(c_int64_t), value :: resize_step
      end function

      ! annealing schedule (read_anneal)
      integer(c_int) function sis_gpu_set_anneal(ctx, n, steps, temps) bind(C, name="sis_gpu_set_anneal")
         import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: n
         integer(c_int64_t), intent(in) :: steps(*)
         real(c_double), intent(in) :: temps(*)
      end function

      ! secondary-structure restriction (read_ss pairs)
      integer(c_int) function sis_gpu_set_ss(ctx, npair, pairs) bind(C, name="sis_gpu_set_ss")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: npair
         integer(c_int32_t), intent(in) :: pairs(2, *)
      end function

      ! stage wall, tweezers, Rg bias, restraints
      integer(c_int) function sis_gpu_set_external(ctx, ext) bind(C, name="sis_gpu_set_external")
         import :: c_int, c_ptr, sis_external
         type(c_ptr), value :: ctx
         type(sis_external), intent(in) :: ext
      end function

      ! restart files (.rst, src/write_rst.F90 blocks) and read_rst
      integer(c_int) function sis_gpu_rst_open(ctx, paths, nstep_save_rst) bind(C, name="sis_gpu_rst_open")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx, paths
         integer(c_int64_t), value :: nstep_save_rst
      end function

      integer(c_int) function sis_gpu_rst_write(ctx) bind(C, name="sis_gpu_rst_write")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function sis_gpu_rst_close(ctx) bind(C, name="sis_gpu_rst_close")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function sis_gpu_rst_read(ctx, paths) bind(C, name="sis_gpu_rst_read")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, paths
      end function

      ! job_check_force: forces against central differences of the energy
      integer(c_int) function sis_gpu_check_force(ctx, irep, h, out_path, f_energy, f_force) &
            bind(C, name="sis_gpu_check_force")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx, out_path, f_energy, f_force
         integer(c_int32_t), value :: irep
         real(c_double), value :: h
      end function

      ! job_dcd: energies of the frames of a DCD file
      integer(c_int) function sis_gpu_dcd_energies(ctx, path, out_path, max_frames, energies, nframes) &
            bind(C, name="sis_gpu_dcd_energies")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx, path, out_path, energies, nframes
         integer(c_int64_t), value :: max_frames
      end function

      integer(c_int) function sis_gpu_kinetic(ctx, irep, ekin) bind(C, name="sis_gpu_kinetic")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx, ekin
         integer(c_int32_t), value :: irep
      end function

      integer(c_int) function sis_gpu_set_step(ctx, istep) bind(C, name="sis_gpu_set_step")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: istep
      end function

      integer(c_int) function sis_gpu_set_labels(ctx, rep2lab) bind(C, name="sis_gpu_set_labels")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, rep2lab
      end function

      integer(c_int) function sis_gpu_get_info(ctx, out, n) bind(C, name="sis_gpu_get_info")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx, out
         integer(c_int32_t), value :: n
      end function

      real(c_double) function sis_gpu_last_kernel_ms(ctx, which) bind(C, name="sis_gpu_last_kernel_ms")
         import :: c_double, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: which
      end function
   end interface

contains

   ! Message of the last failed call as a Fortran string (for sis_abort).
   function sis_gpu_error_message() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: s(:)
      integer :: n
      p = sis_gpu_last_error()
      n = 0
      call c_f_pointer(p, s, [4096])
      do while (n < 4096)
         if (s(n+1) == c_null_char) exit
         n = n + 1
      enddo
      allocate(character(len=n) :: msg)
      msg = transfer(s(1:n), msg)
   end function sis_gpu_error_message

end module sis_gpu_iface
