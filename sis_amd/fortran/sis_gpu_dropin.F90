! Drop-in replacements of the reference's force, neighbor_list and
! energy_sumup (src/force.F90:1-77, src/neighbor_list.F90:1-233,
! src/energy_sumup.F90:1-56): the same names and arguments, the work done
! by libsis_gpu through the C ABI.  Compiled and linked in place of those
! three files, the reference's own job_md (time loop, integrator, replica
! exchange, output; src/job_md.F90) runs unchanged on top of the device.
!
! The context is created at the first call from the module state that
! read_input / read_force_field / read_fasta leave (var_top, var_state,
! var_potential, var_replica, pbc), in FP64 with the reference's pair and
! list order (parity mode).  job_md integrates on the host, so each call
! sends the replica's coordinates.  The limit MC draws the words of the
! replica's own stream mts(irep): a tape is cut from a copy of the stream
! before the call and the stream advanced by the words the device used,
! so the Langevin noise drawn by job_md afterwards is the reference's.
!
! Covered: temperature ladders (labels permuted by the reference's own
! replica_exchange; each call hands the device the replicas' current
! temperatures), nstep_bp_MC of 0 or 1.  Refused with a message: ladders
! over the ionic strength, force or Rg target, external terms, the
! secondary-structure restriction, annealing and temperature-independent
! MC, nstep_bp_MC > 1 (bp_status_MC would have to follow job_md's copies).

module sis_dropin

   use, intrinsic :: iso_c_binding
   use, intrinsic :: iso_fortran_env, only : error_unit
   use sis_gpu_iface
   implicit none

   type(c_ptr), save :: gctx = c_null_ptr
   real(c_double), allocatable, target, save :: cur_T(:), cur_I(:)
   integer(c_int32_t), allocatable, target, save :: t_nmp_chain(:), t_moltype(:), t_seq(:), t_noq(:)
   real(c_double), allocatable, target, save :: xbuf(:,:), fbuf(:,:)
   integer(c_int32_t), allocatable, target, save :: tape(:)

contains

   subroutine dropin_check(rc, what)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: what
      if (rc /= SIS_OK) then
         write(error_unit, '(a,a,a,a)') 'sis_gpu (', what, '): ', sis_gpu_error_message()
         error stop 1
      endif
   end subroutine dropin_check

   subroutine dropin_refuse(what)
      character(*), intent(in) :: what
      write(error_unit, '(a,a)') 'sis_gpu drop-in: not supported: ', what
      error stop 1
   end subroutine dropin_refuse

   ! sis_params / sis_topo from the reference's module state
   subroutine dropin_create()
      use const_idx, only : INTGRT, REPT
      use var_top, only : nchains, nmp, nmp_chain, moltypes, seq, dummy_has_charge, inp_no_charge
      use var_state, only : tempK, dt, viscosity_Pas, nl_margin, ionic_strength, length_per_charge, &
                            nstep_bp_MC, integrator, rng_seed, opt_anneal, temp_independent
      use var_potential, only : bond_k, bond_r0, flg_angl_ReB, angl_k, angl_t0, flg_dih_cos, flg_dih_exp, &
                                dih_k, dih_w, dih_p0, wca_sigma, wca_eps, flg_ele, ele_cutoff_type, &
                                ele_cutoff_inp, ele_exclude_covalent_bond_pairs, &
                                ele_exclude_covalent_angle_pairs, bp_model, bp_min_loop, max_bp_per_nt, &
                                bp_cutoff_energy, coef_dG, dH0, dS0, dHend0, dSend0, flg_NNend, NN_dH, NN_dS, &
                                NNend_dH, NNend_dS, bp_paras, flg_stage, flg_twz, flg_bias_rg, &
                                flg_restraint, flg_bias_ss
      use var_replica, only : nrep_proc, flg_repvar
      use pbc, only : flg_pbc, pbc_box
      type(sis_params) :: p
      type(sis_topo) :: t
      type(sis_options) :: o
      integer :: ich, i, k, imp

      if (flg_stage .or. flg_twz .or. flg_bias_rg .or. flg_restraint) call dropin_refuse('external terms')
      if (flg_bias_ss) call dropin_refuse('secondary-structure restriction')
      if (flg_repvar(REPT%ION) .or. flg_repvar(REPT%TWZDCF) .or. flg_repvar(REPT%RG)) &
         call dropin_refuse('replica ladders other than the temperature')
      if (opt_anneal /= 0 .or. temp_independent /= 0) call dropin_refuse('annealing / temperature-independent MC')
      if (nstep_bp_MC > 1) call dropin_refuse('nstep_bp_MC > 1')

      p%bond_k = bond_k
      p%bond_r0 = bond_r0
      p%angl_k = angl_k
      p%angl_t0 = angl_t0
      p%dih_k = dih_k
      p%dih_w = dih_w
      p%dih_p0 = dih_p0
      p%wca_sigma = wca_sigma
      p%wca_eps = wca_eps
      p%ele_cutoff = ele_cutoff_inp
      p%length_per_charge = length_per_charge
      p%ionic_strength = ionic_strength
      p%bp_cutoff_energy = bp_cutoff_energy
      p%bp_dH0 = dH0
      p%bp_dS0 = dS0
      p%bp_dHend0 = dHend0
      p%bp_dSend0 = dSend0
      p%bp_coef_dG = coef_dG
      p%nn_dH(:) = 0.0_c_double
      p%nn_dS(:) = 0.0_c_double
      p%nnend_dH(:) = 0.0_c_double
      p%nnend_dS(:) = 0.0_c_double
      if (allocated(NN_dH)) p%nn_dH(:) = NN_dH(1:21)
      if (allocated(NN_dS)) p%nn_dS(:) = NN_dS(1:21)
      if (allocated(NNend_dH)) p%nnend_dH(:) = NNend_dH(1:6)
      if (allocated(NNend_dS)) p%nnend_dS(:) = NNend_dS(1:6)
      do k = 1, 3   ! GC, AU, GU; the engine derives CG, UA, UG itself
         p%bp(k)%cutoff_ddist = bp_paras(k)%cutoff_ddist
         p%bp(k)%U0 = bp_paras(k)%U0
         p%bp(k)%bond_k = bp_paras(k)%bond_k
         p%bp(k)%bond_r = bp_paras(k)%bond_r
         p%bp(k)%angl_k1 = bp_paras(k)%angl_k1
         p%bp(k)%angl_k2 = bp_paras(k)%angl_k2
         p%bp(k)%angl_k3 = bp_paras(k)%angl_k3
         p%bp(k)%angl_k4 = bp_paras(k)%angl_k4
         p%bp(k)%angl_theta1 = bp_paras(k)%angl_theta1
         p%bp(k)%angl_theta2 = bp_paras(k)%angl_theta2
         p%bp(k)%angl_theta3 = bp_paras(k)%angl_theta3
         p%bp(k)%angl_theta4 = bp_paras(k)%angl_theta4
         p%bp(k)%dihd_k1 = bp_paras(k)%dihd_k1
         p%bp(k)%dihd_k2 = bp_paras(k)%dihd_k2
         p%bp(k)%dihd_phi1 = bp_paras(k)%dihd_phi1
         p%bp(k)%dihd_phi2 = bp_paras(k)%dihd_phi2
      enddo
      p%tempK = tempK
      p%dt = dt
      p%viscosity_pas = viscosity_Pas
      p%nl_margin = nl_margin
      p%box(:) = 0.0_c_double
      if (flg_pbc) p%box(:) = pbc_box(:)
      p%angl_reb = merge(1, 0, flg_angl_ReB)
      p%dih_type = 0
      if (flg_dih_cos) p%dih_type = 1
      if (flg_dih_exp) p%dih_type = 2
      p%ele_on = merge(1, 0, flg_ele)
      p%ele_cutoff_type = ele_cutoff_type
      p%ele_excl_bond = merge(1, 0, ele_exclude_covalent_bond_pairs)
      p%ele_excl_angl = merge(1, 0, ele_exclude_covalent_angle_pairs)
      p%dummy_has_charge = merge(1, 0, dummy_has_charge)
      p%bp_model = bp_model
      p%bp_min_loop = bp_min_loop
      p%max_bp_per_nt = max_bp_per_nt
      p%nstep_bp_mc = int(nstep_bp_MC, c_int32_t)
      p%nnend = merge(1, 0, flg_NNend)
      p%pbc = merge(1, 0, flg_pbc)
      p%integrator = 1
      if (integrator == INTGRT%BD_EM) p%integrator = 2

      allocate(t_nmp_chain(nchains), t_moltype(nchains), t_seq(nmp))
      imp = 0
      do ich = 1, nchains
         t_nmp_chain(ich) = nmp_chain(ich)
         t_moltype(ich) = moltypes(ich)
         do i = 1, nmp_chain(ich)
            imp = imp + 1
            t_seq(imp) = seq(i, ich)
         enddo
      enddo
      t%nchains = nchains
      t%nmp = nmp
      t%nmp_chain = c_loc(t_nmp_chain)
      t%moltype = c_loc(t_moltype)
      t%seq = c_loc(t_seq)
      t%no_charge = c_null_ptr
      t%n_no_charge = 0
      t%pad = 0
      if (allocated(inp_no_charge)) then
         if (size(inp_no_charge) > 0) then
            allocate(t_noq(size(inp_no_charge)))
            t_noq(:) = inp_no_charge(:)
            t%no_charge = c_loc(t_noq)
            t%n_no_charge = size(inp_no_charge)
         endif
      endif

      o%nrep = nrep_proc
      o%precision = 0
      o%parity = 1
      o%mode = 1
      o%block = 0
      o%max_nbr = 0
      o%rebuild_every = 0
      o%exchange_every = 0   ! job_md's own replica_exchange permutes the labels
      o%seed = int(rng_seed, c_int64_t)
      o%grid = 0
      o%flags = 2            ! SIS_OPT_NO_GRAPH: no steps are taken on the device
      call dropin_check(sis_gpu_create(t, p, o, gctx), 'create')

      allocate(xbuf(3, nmp), fbuf(3, nmp))
      allocate(tape(4096 + 64 * nmp))
   end subroutine dropin_create

   ! The reference's arrays that job_md and energy_replica touch around the
   ! calls (src/job_md.F90:345-347,399-401, src/energy_replica.F90:64-70,
   ! src/write_bp.F90:36): the device holds the real lists and base-pair
   ! state, so these stay one-entry placeholders with nbp = 0.
   subroutine dropin_placeholders()
      use var_replica, only : nrep_proc
      use var_state, only : bp_status, bp_status_MC, ene_bp
      use var_potential, only : nbp, nbp_max, bp_mp
      if (allocated(nbp)) return
      nbp_max = 1
      allocate(nbp(nrep_proc), bp_mp(7, 1, nrep_proc))
      allocate(bp_status(1, nrep_proc), bp_status_MC(1, nrep_proc), ene_bp(1, nrep_proc))
      nbp(:) = 0
      bp_mp(:,:,:) = 0
      bp_status(:,:) = .False.
      bp_status_MC(:,:) = .False.
      ene_bp(:,:) = 0.0_c_double
   end subroutine dropin_placeholders

   subroutine dropin_ready()
      if (.not. c_associated(gctx)) then
         call dropin_create()
         call dropin_placeholders()
      endif
      call dropin_conditions()
   end subroutine dropin_ready

   ! Each replica's temperature (its label's, src/force_bp_limit_triplet.F90:61-67)
   ! and ionic strength; a change (after an exchange) goes to the device,
   ! which then rebuilds its lists at the positions it is given.
   subroutine dropin_conditions()
      use const_idx, only : REPT
      use var_state, only : tempK, ionic_strength
      use var_replica, only : nrep_proc, flg_repvar, rep2val, irep2grep
      real(c_double) :: T(nrep_proc)
      integer :: r
      do r = 1, nrep_proc
         if (flg_repvar(REPT%TEMP)) then
            T(r) = rep2val(irep2grep(r), REPT%TEMP)
         else
            T(r) = tempK
         endif
      enddo
      if (allocated(cur_T)) then
         if (all(cur_T(:) == T(:))) return
      else
         allocate(cur_T(nrep_proc), cur_I(nrep_proc))
         cur_I(:) = ionic_strength
      endif
      cur_T(:) = T(:)
      call dropin_check(sis_gpu_set_rep_conditions(gctx, c_loc(cur_T), c_loc(cur_I)), 'set_rep_conditions')
      call dropin_send_all()
      call dropin_check(sis_gpu_rebuild(gctx, 0_c_int32_t), 'rebuild')
   end subroutine dropin_conditions

   subroutine dropin_send(irep)
      use var_state, only : xyz
      integer, intent(in) :: irep
      xbuf(:,:) = xyz(:,:,irep)
      call dropin_check(sis_gpu_set_positions(gctx, int(irep, c_int32_t), c_loc(xbuf)), 'set_positions')
   end subroutine dropin_send

   subroutine dropin_send_all()
      use var_replica, only : nrep_proc
      integer :: r
      do r = 1, nrep_proc
         call dropin_send(r)
      enddo
   end subroutine dropin_send_all

   ! A tape of the words mts(irep) would give next, for the limit MC
   subroutine dropin_tape_begin(irep)
      use mt_stream, only : mt_state, genrand_int32
      use var_state, only : mts
      integer, intent(in) :: irep
      type(mt_state) :: tmp
      integer :: k
      tmp = mts(irep)
      allocate(tmp%state(0:tmp%nn-1))
      tmp%state(:) = mts(irep)%state(:)
      do k = 1, size(tape)
         tape(k) = genrand_int32(tmp)
      enddo
      deallocate(tmp%state)
      call dropin_check(sis_gpu_set_rng_tape(gctx, int(irep, c_int32_t), c_loc(tape), &
                                             int(size(tape), c_int64_t)), 'set_rng_tape')
   end subroutine dropin_tape_begin

   ! mts(irep) advanced past the words the device drew
   subroutine dropin_tape_end(irep)
      use mt_stream, only : genrand_int32
      use var_state, only : mts
      integer, intent(in) :: irep
      integer(c_int64_t) :: used, k
      integer(c_int32_t) :: w
      call dropin_check(sis_gpu_get_rng_tape_pos(gctx, int(irep, c_int32_t), used), 'get_rng_tape_pos')
      do k = 1, used
         w = genrand_int32(mts(irep))
      enddo
   end subroutine dropin_tape_end

end module sis_dropin


! neighbor_list(irep) (src/neighbor_list.F90:1-233): pbc_wrap and the lists
! of replica irep on the device; the wrapped coordinates come back to xyz.
subroutine neighbor_list(irep)
   use, intrinsic :: iso_c_binding
   use sis_gpu_iface
   use sis_dropin
   use var_state, only : xyz, lambdaD
   use var_top, only : nmp
   use var_replica, only : nrep_proc
   use var_potential, only : flg_ele, ele_cutoff_type, ele_cutoff_inp, ele_cutoff
   use pbc, only : flg_pbc
   implicit none
   integer, intent(in) :: irep
   real(c_double), allocatable, target :: all_xyz(:,:,:)

   call dropin_ready()
   call dropin_send_all()
   call dropin_check(sis_gpu_rebuild(gctx, int(irep, c_int32_t)), 'rebuild')
   if (flg_pbc) then
      allocate(all_xyz(3, nmp, nrep_proc))
      call dropin_check(sis_gpu_get_state(gctx, c_loc(all_xyz), c_null_ptr, c_null_ptr), 'get_state')
      xyz(:,:,irep) = all_xyz(:,:,irep)
   endif
   ! the cutoff follows the Debye length at each rebuild (src/neighbor_list.F90:80-82)
   if (flg_ele .and. ele_cutoff_type == 2) ele_cutoff(irep) = ele_cutoff_inp * lambdaD(irep)
end subroutine neighbor_list


! force(irep, forces) (src/force.F90:1-77); flg_bp_MC selects the limit MC.
! As the reference's base-pair force routines do (src/force_bp.F90:208,
! src/force_bp_triplet.F90:244, src/force_bp_limit_triplet.F90:375), it
! leaves flg_bp_energy set: a following energy_sumup takes the base-pair
! state of this call.
subroutine force(irep, forces)
   use, intrinsic :: iso_c_binding
   use sis_gpu_iface
   use sis_dropin
   use const, only : PREC
   use var_state, only : flg_bp_MC, flg_bp_energy
   use var_top, only : nmp
   implicit none
   integer, intent(in) :: irep
   real(PREC), intent(out) :: forces(3, nmp)
   integer(c_int32_t) :: mc

   call dropin_ready()
   call dropin_send(irep)
   mc = merge(1, 0, flg_bp_MC)
   if (flg_bp_MC) call dropin_tape_begin(irep)
   call dropin_check(sis_gpu_force(gctx, int(irep, c_int32_t), mc, c_loc(fbuf)), 'force')
   if (flg_bp_MC) call dropin_tape_end(irep)
   forces(:,:) = fbuf(:,:)
   flg_bp_energy = .True.
end subroutine force


! energy_sumup(irep, tempK_in, energies) (src/energy_sumup.F90:1-56);
! flg_bp_MC / flg_bp_energy as the reference reads them.
subroutine energy_sumup(irep, tempK_in, energies)
   use, intrinsic :: iso_c_binding
   use sis_gpu_iface
   use sis_dropin
   use const, only : PREC
   use const_idx, only : ENE
   use var_state, only : flg_bp_MC, flg_bp_energy
   implicit none
   integer, intent(in) :: irep
   real(PREC), intent(in) :: tempK_in
   real(PREC), intent(out) :: energies(0:ENE%MAX)
   real(c_double) :: e(0:10)
   integer(c_int32_t) :: flags
   logical :: mc

   call dropin_ready()
   call dropin_send(irep)
   flags = 0
   if (flg_bp_MC) flags = flags + 1
   if (flg_bp_energy) flags = flags + 2
   mc = flg_bp_MC .and. .not. flg_bp_energy
   if (mc) call dropin_tape_begin(irep)
   call dropin_check(sis_gpu_energy(gctx, int(irep, c_int32_t), real(tempK_in, c_double), flags, e), 'energy')
   if (mc) call dropin_tape_end(irep)
   energies(:) = 0.0_PREC
   energies(0:min(10, ENE%MAX)) = e(0:min(10, ENE%MAX))
end subroutine energy_sumup
