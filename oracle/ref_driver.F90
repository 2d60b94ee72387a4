! Test-only driver around the reference SIS Fortran (compiled from its own
! sources by oracle/Makefile into oracle/_ref/).  It is part of the test
! oracle: only tests and the CPU-baseline leg of bench.py load it.
!
! The reference reads its configuration with read_input/read_force_field,
! which need the un-vendored toml-f library; those two files are not
! compiled.  This driver instead fills the same module variables that
! read_input.F90 and read_force_field.F90 would fill, then runs the
! reference's own initialisation chain in the order of src/main.F90:28-140
! (init_replica, init_mts, init_pbc, init_out_files, init_bp, init_ele,
! list_local) and exposes the hot-path routines through bind(C) entry
! points:
!   neighbor_list  (src/neighbor_list.F90)
!   force          (src/force.F90)
!   energy_sumup   (src/energy_sumup.F90)
!   rnd_boxmuller  (src/rnd_boxmuller.F90)
!   job_md         (src/job_md.F90)
! Nothing here re-implements physics.

module sisref_driver

   use, intrinsic :: iso_c_binding
   implicit none

   ! One base-pair parameter block, field order of basepair_parameters
   ! (src/var_potential.F90:63-80).
   type, bind(C) :: c_bp_para
      real(c_double) :: cutoff_ddist, U0, bond_k, bond_r
      real(c_double) :: angl_k1, angl_k2, angl_k3, angl_k4
      real(c_double) :: angl_theta1, angl_theta2, angl_theta3, angl_theta4
      real(c_double) :: dihd_k1, dihd_k2, dihd_phi1, dihd_phi2
   end type c_bp_para

   ! Same layout as struct sis_params in include/sis_gpu.h.
   type, bind(C) :: c_params
      real(c_double) :: bond_k, bond_r0
      real(c_double) :: angl_k, angl_t0
      real(c_double) :: dih_k, dih_w, dih_p0
      real(c_double) :: wca_sigma, wca_eps
      real(c_double) :: ele_cutoff, length_per_charge, ionic_strength
      real(c_double) :: bp_cutoff_energy
      real(c_double) :: bp_dH0, bp_dS0, bp_dHend0, bp_dSend0, bp_coef_dG
      real(c_double) :: nn_dH(21), nn_dS(21)
      real(c_double) :: nnend_dH(6), nnend_dS(6)
      type(c_bp_para) :: bp(3)
      real(c_double) :: tempK
      real(c_double) :: dt, viscosity_pas, nl_margin
      real(c_double) :: box(3)
      integer(c_int32_t) :: angl_reb, dih_type
      integer(c_int32_t) :: ele_on, ele_cutoff_type, ele_excl_bond, ele_excl_angl, dummy_has_charge
      integer(c_int32_t) :: bp_model, bp_min_loop, max_bp_per_nt, nstep_bp_mc, nnend
      integer(c_int32_t) :: pbc, integrator
   end type c_params

   ! REMD over another variable than the temperature (REPT id: 3 ionic
   ! strength, 4 Rg target): set by sisref_remd_var before sisref_setup
   integer, save :: remd_kind = 0
   ! a temperature dimension beside remd_kind's (a two-dimensional grid):
   ! set by sisref_remd_temps before sisref_setup
   real(c_double), allocatable, save :: remd_temps(:)
   integer(c_int64_t), save :: twz_nsave = 0   ! nstep_save of a job_md run that writes .twz lines
   real(c_double), allocatable, save :: remd_values(:)

contains

   subroutine sisref_remd_var(kind, n, values) bind(C, name="sisref_remd_var")
      integer(c_int), value :: kind, n
      real(c_double), intent(in) :: values(n)
      remd_kind = kind
      if (allocated(remd_values)) deallocate(remd_values)
      allocate(remd_values(n))
      remd_values(:) = values(:)
   end subroutine sisref_remd_var

   subroutine sisref_remd_temps(n, values) bind(C, name="sisref_remd_temps")
      integer(c_int), value :: n
      real(c_double), intent(in) :: values(n)
      if (allocated(remd_temps)) deallocate(remd_temps)
      allocate(remd_temps(n))
      remd_temps(:) = values(:)
   end subroutine sisref_remd_temps

   subroutine copy_bp(src, dst)
      use var_potential, only : basepair_parameters
      type(c_bp_para), intent(in) :: src
      type(basepair_parameters), intent(out) :: dst
      dst%cutoff_ddist = src%cutoff_ddist
      dst%U0 = src%U0
      dst%bond_k = src%bond_k
      dst%bond_r = src%bond_r
      dst%angl_k1 = src%angl_k1
      dst%angl_k2 = src%angl_k2
      dst%angl_k3 = src%angl_k3
      dst%angl_k4 = src%angl_k4
      dst%angl_theta1 = src%angl_theta1
      dst%angl_theta2 = src%angl_theta2
      dst%angl_theta3 = src%angl_theta3
      dst%angl_theta4 = src%angl_theta4
      dst%dihd_k1 = src%dihd_k1
      dst%dihd_k2 = src%dihd_k2
      dst%dihd_phi1 = src%dihd_phi1
      dst%dihd_phi2 = src%dihd_phi2
   end subroutine copy_bp

   ! Mirror of the CG/UA/UG derivation done by read_force_field
   ! (src/read_force_field.F90:203-207 and the AU/GU analogues): the
   ! mirrored type swaps angl_k1<->angl_k2 and angl_k3<->angl_k4.
   subroutine mirror_bp(src, dst)
      use var_potential, only : basepair_parameters
      type(basepair_parameters), intent(in) :: src
      type(basepair_parameters), intent(out) :: dst
      dst = src
      dst%angl_k1 = src%angl_k2
      dst%angl_k2 = src%angl_k1
      dst%angl_k3 = src%angl_k4
      dst%angl_k4 = src%angl_k3
   end subroutine mirror_bp

   ! --------------------------------------------------------------------
   ! Set up a system.  seq_flat holds SEQT codes chain after chain;
   ! xyz is (3, nmp, nrep) column-major.  rep_temps(nrep) gives the
   ! temperature ladder when nrep > 1 (temperature REMD), else ignored.
   ! prefix is a NUL-terminated output prefix for the .out/.dcd files
   ! that the reference's job_md writes.
   ! --------------------------------------------------------------------
   subroutine sisref_setup(p, nchains_in, nmp_chain_in, moltype_in, seq_flat, nrep_in, &
                           rep_temps, rng_seed_in, xyz_in, prefix, nstep_rep_exchange_in) &
                           bind(C, name="sisref_setup")
      use const
      use const_phys, only : BOLTZ_KCAL_MOL
      use const_idx, only : JOBT, INTGRT, REPT, MOLT
      use var_parallel, only : init_parallel
      use var_state
      use var_top
      use var_potential
      use var_io
      use var_replica
      use pbc, only : flg_pbc, pbc_box_input, flg_pbc_resize, flg_pbc_ignore_rst, init_pbc

      type(c_params), intent(in) :: p
      integer(c_int), value :: nchains_in, nrep_in, rng_seed_in
      integer(c_int), intent(in) :: nmp_chain_in(nchains_in), moltype_in(nchains_in)
      integer(c_int), intent(in) :: seq_flat(*)
      real(c_double), intent(in) :: rep_temps(*)
      real(c_double), intent(in) :: xyz_in(*)
      character(kind=c_char), intent(in) :: prefix(*)
      integer(c_int64_t), value :: nstep_rep_exchange_in

      integer :: ichain, i, imp, maxnmp, n, irep
      character(len=:), allocatable :: cpre

      call init_const()
      call init_parallel()

      ! ---- what read_input.F90 would set ([Job], [Condition], [MD]) ----
      job = JOBT%MD
      restarted = .False.
      reset_step = .False.
      tempK = p%tempK
      kT = BOLTZ_KCAL_MOL * tempK
      rng_seed = rng_seed_in
      integrator = INTGRT%LD_GJF2GJ
      if (p%integrator == INTGRT%BD_EM) integrator = INTGRT%BD_EM
      viscosity_Pas = p%viscosity_pas
      dt = p%dt
      nstep = 0
      nstep_save = huge(nstep_save)
      nstep_save_rst = huge(nstep_save_rst)
      nstep_check_stop = huge(nstep_check_stop)
      stop_wall_time_sec = 0
      fix_com_origin = 0
      nl_margin = p%nl_margin
      opt_anneal = 0
      temp_independent = 0
      ionic_strength = p%ionic_strength
      length_per_charge = p%length_per_charge
      nstep_bp_MC = p%nstep_bp_mc
      flg_progress = .False.
      flg_freeze = .False.
      dummy_has_charge = (p%dummy_has_charge /= 0)

      flg_in_ct = .False.
      flg_in_bpseq = .False.
      flg_in_bpl = .False.
      flg_in_fasta = .True.
      flg_out_bp = .False.
      flg_out_bpall = .False.
      flg_out_bpe = .False.
      flg_out_bpcoef = .False.

      n = 0
      do while (prefix(n+1) /= c_null_char)
         n = n + 1
      enddo
      allocate(character(len=n) :: cpre)
      do i = 1, n
         cpre(i:i) = prefix(i)
      enddo
      cfile_prefix = cpre

      ! ---- [Replica] ----
      flg_repvar(:) = .False.
      nrep(:) = 1
      if (nrep_in > 1) then
         flg_replica = .True.
         flg_exchange = .True.
         replica_values(:,:) = 0.0_PREC
         if (allocated(remd_temps)) then
            ! [Replica] nrep_temp x the other variable's ladder
            flg_repvar(REPT%TEMP) = .True.
            nrep(REPT%TEMP) = size(remd_temps)
            replica_values(1:size(remd_temps), REPT%TEMP) = remd_temps(:)
            flg_repvar(remd_kind) = .True.
            nrep(remd_kind) = size(remd_values)
            replica_values(1:size(remd_values), remd_kind) = remd_values(:)
            if (nrep(REPT%TEMP) * nrep(remd_kind) /= nrep_in) error stop "sisref: grid size /= nrep"
         else if (remd_kind == REPT%ION .or. remd_kind == REPT%RG .or. remd_kind == REPT%TWZDCF) then
            ! [Replica] ion_strength / rg / force ladder
            flg_repvar(remd_kind) = .True.
            nrep(remd_kind) = nrep_in
            do irep = 1, nrep_in
               replica_values(irep, remd_kind) = remd_values(irep)
            enddo
         else
            flg_repvar(REPT%TEMP) = .True.
            nrep(REPT%TEMP) = nrep_in
            do irep = 1, nrep_in
               replica_values(irep, REPT%TEMP) = rep_temps(irep)
            enddo
         endif
         nstep_rep_exchange = nstep_rep_exchange_in
         nstep_rep_save = nstep_rep_exchange_in
      else
         flg_replica = .False.
         flg_exchange = .False.
      endif

      ! ---- [PBC_box] ----
      flg_pbc = (p%pbc /= 0)
      pbc_box_input(:) = p%box(:)
      flg_pbc_resize = .False.
      flg_pbc_ignore_rst = .True.

      ! ---- [Basepair], [Electrostatic] sections of read_input ----
      bp_model = p%bp_model
      bp_min_loop = p%bp_min_loop
      max_bp_per_nt = p%max_bp_per_nt
      flg_ele = (p%ele_on /= 0)
      ele_cutoff_type = p%ele_cutoff_type
      ele_cutoff_inp = p%ele_cutoff
      ele_exclude_covalent_bond_pairs = (p%ele_excl_bond /= 0)
      ele_exclude_covalent_angle_pairs = (p%ele_excl_angl /= 0)
      flg_stage = .False.
      flg_twz = .False.
      flg_bias_ss = .False.
      flg_bias_rg = .False.
      flg_timed_bias_rg = .False.
      flg_restraint = .False.
      ntwz_DCF = 0
      ntwz_FR = 0

      call init_replica()
      call init_mts()

      ! ---- what read_force_field.F90 would set ----
      bond_k = p%bond_k
      bond_r0 = p%bond_r0
      flg_angl_ReB = (p%angl_reb /= 0)
      angl_k = p%angl_k
      angl_t0 = p%angl_t0
      flg_dih_cos = (p%dih_type == 1)
      flg_dih_exp = (p%dih_type == 2)
      dih_k = p%dih_k
      dih_w = p%dih_w
      dih_p0 = p%dih_p0
      wca_sigma = p%wca_sigma
      wca_eps = p%wca_eps
      bp_cutoff_energy = p%bp_cutoff_energy
      coef_dG = p%bp_coef_dG
      dH0 = p%bp_dH0
      dS0 = p%bp_dS0
      dHend0 = p%bp_dHend0
      dSend0 = p%bp_dSend0
      flg_NNend = (p%nnend /= 0)
      allocate(NN_dH(21), NN_dS(21), NNend_dH(6), NNend_dS(6))
      NN_dH(:) = p%nn_dH(:)
      NN_dS(:) = p%nn_dS(:)
      NNend_dH(:) = p%nnend_dH(:)
      NNend_dS(:) = p%nnend_dS(:)
      call copy_bp(p%bp(1), bp_paras(1))   ! GC
      call copy_bp(p%bp(2), bp_paras(2))   ! AU
      call copy_bp(p%bp(3), bp_paras(3))   ! GU
      call mirror_bp(bp_paras(1), bp_paras(4))   ! CG
      call mirror_bp(bp_paras(2), bp_paras(5))   ! UA
      call mirror_bp(bp_paras(3), bp_paras(6))   ! UG

      if (flg_pbc) call init_pbc(restarted)

      call init_out_files()

      ! ---- what read_fasta.F90 would set (init_sequence) ----
      nchains = nchains_in
      allocate(nmp_chain(nchains), moltypes(nchains))
      nmp_chain(:) = nmp_chain_in(:)
      moltypes(:) = moltype_in(:)
      maxnmp = maxval(nmp_chain)
      nmp = sum(nmp_chain)
      allocate(seq(maxnmp, nchains), imp_chain(maxnmp, nchains), ichain_mp(nmp), lmp_mp(nmp))
      seq(:,:) = -1
      imp_chain(:,:) = 0
      imp = 0
      do ichain = 1, nchains
         do i = 1, nmp_chain(ichain)
            imp = imp + 1
            seq(i, ichain) = seq_flat(imp)
            imp_chain(i, ichain) = imp
            ichain_mp(imp) = ichain
            lmp_mp(imp) = i
         enddo
      enddo

      allocate(xyz(3, nmp, nrep_proc))
      do irep = 1, nrep_proc
         do imp = 1, nmp
            do i = 1, 3
               xyz(i, imp, irep) = xyz_in(i + 3*(imp-1) + 3*nmp*(irep-1))
            enddo
         enddo
      enddo

      call init_bp()
      if (flg_ele) call init_ele()
      call list_local()

   end subroutine sisref_setup

   subroutine sisref_set_xyz(xyz_in) bind(C, name="sisref_set_xyz")
      use var_state, only : xyz
      use var_top, only : nmp
      use var_replica, only : nrep_proc
      real(c_double), intent(in) :: xyz_in(3, nmp, nrep_proc)
      xyz(:,:,:) = xyz_in(:,:,:)
   end subroutine sisref_set_xyz

   subroutine sisref_get_xyz(xyz_out) bind(C, name="sisref_get_xyz")
      use var_state, only : xyz
      use var_top, only : nmp
      use var_replica, only : nrep_proc
      real(c_double), intent(out) :: xyz_out(3, nmp, nrep_proc)
      xyz_out(:,:,:) = xyz(:,:,:)
   end subroutine sisref_get_xyz

   subroutine sisref_get_state(velos_out, accels_out) bind(C, name="sisref_get_state")
      use var_state, only : velos, accels
      use var_top, only : nmp
      use var_replica, only : nrep_proc
      real(c_double), intent(out) :: velos_out(3, nmp, nrep_proc), accels_out(3, nmp, nrep_proc)
      velos_out(:,:,:) = velos(:,:,:)
      accels_out(:,:,:) = accels(:,:,:)
   end subroutine sisref_get_state

   subroutine sisref_set_velos(velos_in) bind(C, name="sisref_set_velos")
      use var_state, only : velos
      use var_top, only : nmp
      use var_replica, only : nrep_proc
      real(c_double), intent(in) :: velos_in(3, nmp, nrep_proc)
      if (.not. allocated(velos)) allocate(velos(3, nmp, nrep_proc))   ! job_md allocates it otherwise
      velos(:,:,:) = velos_in(:,:,:)
   end subroutine sisref_set_velos

   ! energy_kinetic needs var_top's mass, which job_md fills
   ! (src/job_md.F90:116-133); set it the same way here
   subroutine sisref_kinetic(irep, ekin) bind(C, name="sisref_kinetic")
      use const, only : PREC
      use const_idx, only : SEQT
      use var_top, only : nmp, mass, seq, lmp_mp, ichain_mp
      integer(c_int), value :: irep
      real(c_double), intent(out) :: ekin
      integer :: imp
      if (.not. allocated(mass)) allocate(mass(nmp))
      if (.true.) then
         do imp = 1, nmp
            select case (seq(lmp_mp(imp), ichain_mp(imp)))
               case (SEQT%A)
                  mass(imp) = 328.212_PREC
               case (SEQT%G)
                  mass(imp) = 344.212_PREC
               case (SEQT%C)
                  mass(imp) = 304.182_PREC
               case (SEQT%U)
                  mass(imp) = 305.164_PREC
               case default
                  mass(imp) = 320.443_PREC
            endselect
         enddo
      endif
      call energy_kinetic(irep, ekin)
   end subroutine sisref_kinetic

   subroutine sisref_neighbor_list(irep) bind(C, name="sisref_neighbor_list")
      integer(c_int), value :: irep
      call neighbor_list(irep)
   end subroutine sisref_neighbor_list

   ! The cutoff-free lists of the DCD and CHECK_FORCE jobs
   ! (src/main.F90:144-154): list_bp, list_exv, list_ele in place of
   ! neighbor_list.  list_bp sizes nbp by count(bp3_map /= 0)
   ! (src/list_bp.F90:18,28), which counts both triangles of the symmetric
   ! map of models 1/3 and the cross-chain pairs its jmp >= imp +
   ! bp_min_loop + 1 loop skips; the entries it leaves zero would be read at
   ! index 0 by the energy and force routines, so nbp is cut to the entries
   ! it filled.
   subroutine sisref_analysis_lists() bind(C, name="sisref_analysis_lists")
      use var_potential, only : flg_ele, nbp, bp_mp
      call list_bp()
      nbp(:) = count(bp_mp(1, :, 1) > 0)
      call list_exv()
      if (flg_ele) call list_ele()
   end subroutine sisref_analysis_lists

   subroutine sisref_counts(irep, counts) bind(C, name="sisref_counts")
      use var_potential, only : nwca, nele, nbp, flg_ele
      integer(c_int), value :: irep
      integer(c_int), intent(out) :: counts(3)
      counts(1) = nwca(irep)
      counts(2) = 0
      if (flg_ele) counts(2) = nele(irep)
      counts(3) = nbp(irep)
   end subroutine sisref_counts

   ! wca(2,nwca), ele(2,nele), bp(7,nbp), bpcoef(2,nbp), 1-based bead ids
   subroutine sisref_lists(irep, wca, ele, bp, bpcoef) bind(C, name="sisref_lists")
      use var_potential, only : nwca, nele, nbp, wca_mp, ele_mp, bp_mp, bp_coef, flg_ele
      integer(c_int), value :: irep
      integer(c_int), intent(out) :: wca(2,*), ele(2,*), bp(7,*)
      real(c_double), intent(out) :: bpcoef(2,*)
      integer :: k
      do k = 1, nwca(irep)
         wca(:,k) = wca_mp(:,k,irep)
      enddo
      if (flg_ele) then
         do k = 1, nele(irep)
            ele(:,k) = ele_mp(:,k,irep)
         enddo
      endif
      do k = 1, nbp(irep)
         bp(:,k) = bp_mp(:,k,irep)
         ! list_bp allocates bp_coef for model 5 only (src/list_bp.F90:36-39)
         if (allocated(bp_coef)) then
            bpcoef(:,k) = bp_coef(:,k,irep)
         else
            bpcoef(:,k) = 0.0_c_double
         endif
      enddo
   end subroutine sisref_lists

   ! The BP MC flag is a module variable set per step by job_md
   ! (src/job_md.F90:364-367); here the caller chooses it.
   subroutine sisref_force(irep, bp_mc, forces) bind(C, name="sisref_force")
      use var_state, only : flg_bp_MC, bp_status_MC, nstep_bp_MC
      use var_top, only : nmp
      integer(c_int), value :: irep, bp_mc
      real(c_double), intent(out) :: forces(3, nmp)
      flg_bp_MC = (bp_mc /= 0)
      call force(irep, forces)
   end subroutine sisref_force

   subroutine sisref_set_bp_status_mc(irep, st) bind(C, name="sisref_set_bp_status_mc")
      use var_state, only : bp_status_MC
      use var_potential, only : nbp
      integer(c_int), value :: irep
      integer(c_int8_t), intent(in) :: st(*)
      integer :: k
      do k = 1, nbp(irep)
         bp_status_MC(k, irep) = (st(k) /= 0)
      enddo
   end subroutine sisref_set_bp_status_mc

   subroutine sisref_bp_state(irep, st, ene) bind(C, name="sisref_bp_state")
      use var_state, only : bp_status, ene_bp
      use var_potential, only : nbp
      integer(c_int), value :: irep
      integer(c_int8_t), intent(out) :: st(*)
      real(c_double), intent(out) :: ene(*)
      integer :: k
      do k = 1, nbp(irep)
         st(k) = 0
         if (bp_status(k, irep)) st(k) = 1
         ene(k) = ene_bp(k, irep)
      enddo
   end subroutine sisref_bp_state

   ! energies(0:10) as ENE indices (src/const_idx.F90:21)
   subroutine sisref_energy(irep, tempk_in, bp_mc, bp_energy_valid, energies) bind(C, name="sisref_energy")
      use var_state, only : flg_bp_MC, flg_bp_energy, rg
      use var_replica, only : nrep_proc
      integer(c_int), value :: irep, bp_mc, bp_energy_valid
      real(c_double), value :: tempk_in
      real(c_double), intent(out) :: energies(0:10)
      interface
         subroutine energy_sumup(irep, tempK_in, energies)
            use const, only : PREC
            use const_idx, only : ENE
            integer,    intent(in)  :: irep
            real(PREC), intent(in)  :: tempK_in
            real(PREC), intent(out) :: energies(0:ENE%MAX)
         end subroutine energy_sumup
      end interface
      logical :: own_rg
      flg_bp_MC = (bp_mc /= 0)
      flg_bp_energy = (bp_energy_valid /= 0)
      ! energy_rg stores Rg in var_state's rg, which job_md allocates
      own_rg = .not. allocated(rg)
      if (own_rg) allocate(rg(nrep_proc))
      call energy_sumup(irep, tempk_in, energies)
      if (own_rg) deallocate(rg)
   end subroutine sisref_energy

   ! The external terms, set as read_input / init_tweezers / init_restraint
   ! leave them (src/var_potential.F90:115-163); istep keys the force-ramp
   ! anchors of force_tweezers.
   subroutine sisref_set_external(stage, sigma, eps, ndcf, dcf_pairs, dcf_forces, nfr, fr_pairs, fr_k, &
                                  fr_init, fr_velo, rgpot, rgk, rg0, nrest, rest_id, rest_rcut, rest_para, &
                                  istep_in) bind(C, name="sisref_set_external")
      use var_potential, only : flg_stage, stage_sigma, stage_eps, flg_twz, ntwz_DCF, twz_DCF_pairs, &
                                twz_DCF_forces, ntwz_FR, twz_FR_pairs, twz_FR_k, twz_FR_init, twz_FR_velo, &
                                flg_bias_rg, bias_rg_pott, bias_rg_k, bias_rg_0, flg_restraint, flg_rest_sigb, &
                                nrest_sigb, rest_sigb_id, rest_sigb_rcut, rest_sigb_para
      use var_io, only : flg_out_twz
      use var_state, only : istep
      use var_replica, only : nrep_proc
      integer(c_int), value :: stage, ndcf, nfr, rgpot, nrest
      real(c_double), value :: sigma, eps, rgk
      integer(c_int64_t), value :: istep_in
      integer(c_int), intent(in) :: dcf_pairs(2, *), fr_pairs(2, *), rest_id(3, *)
      real(c_double), intent(in) :: dcf_forces(3, max(ndcf, 1), *), fr_k(2, *), fr_init(3, 2, *), &
                                    fr_velo(3, 2, *), rg0(*), rest_rcut(*), rest_para(3, *)
      flg_stage = (stage /= 0)
      stage_sigma = sigma
      stage_eps = eps
      flg_out_twz = .False.
      ntwz_DCF = ndcf
      ntwz_FR = nfr
      flg_twz = (ndcf > 0 .or. nfr > 0)
      if (allocated(twz_DCF_pairs)) deallocate(twz_DCF_pairs, twz_DCF_forces)
      allocate(twz_DCF_pairs(2, max(ndcf, 1)), twz_DCF_forces(3, max(ndcf, 1), nrep_proc))
      if (ndcf > 0) then
         twz_DCF_pairs(:, 1:ndcf) = dcf_pairs(:, 1:ndcf)
         twz_DCF_forces(:, 1:ndcf, 1:nrep_proc) = dcf_forces(:, 1:ndcf, 1:nrep_proc)
      endif
      if (allocated(twz_FR_pairs)) deallocate(twz_FR_pairs, twz_FR_k, twz_FR_init, twz_FR_velo)
      allocate(twz_FR_pairs(2, max(nfr, 1)), twz_FR_k(2, max(nfr, 1)), twz_FR_init(3, 2, max(nfr, 1)), &
               twz_FR_velo(3, 2, max(nfr, 1)))
      if (nfr > 0) then
         twz_FR_pairs(:, 1:nfr) = fr_pairs(:, 1:nfr)
         twz_FR_k(:, 1:nfr) = fr_k(:, 1:nfr)
         twz_FR_init(:, :, 1:nfr) = fr_init(:, :, 1:nfr)
         twz_FR_velo(:, :, 1:nfr) = fr_velo(:, :, 1:nfr)
      endif
      flg_bias_rg = (rgpot /= 0)
      bias_rg_pott = rgpot
      bias_rg_k = rgk
      if (allocated(bias_rg_0)) deallocate(bias_rg_0)
      allocate(bias_rg_0(nrep_proc))
      if (rgpot /= 0) bias_rg_0(1:nrep_proc) = rg0(1:nrep_proc)
      nrest_sigb = nrest
      flg_rest_sigb = (nrest > 0)
      flg_restraint = (nrest > 0)
      if (allocated(rest_sigb_id)) deallocate(rest_sigb_id, rest_sigb_rcut, rest_sigb_para)
      allocate(rest_sigb_id(3, max(nrest, 1)), rest_sigb_rcut(max(nrest, 1)), rest_sigb_para(3, max(nrest, 1)))
      if (nrest > 0) then
         rest_sigb_id(:, 1:nrest) = rest_id(:, 1:nrest)
         rest_sigb_rcut(1:nrest) = rest_rcut(1:nrest)
         rest_sigb_para(:, 1:nrest) = rest_para(:, 1:nrest)
      endif
      istep = istep_in
   end subroutine sisref_set_external

   ! Re-run init_bp with a base-pair list (BPL) file, as a run with
   ! [Files.In] bpl does (src/init_bp.F90:177-181 -> src/read_ss.F90).
   subroutine sisref_set_ss(path, plen) bind(C, name="sisref_set_ss")
      use var_io, only : flg_in_bpl, cfile_bpl_in
      use var_potential, only : bp_map, bp3_map
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: plen
      integer :: k
      character(len=:), allocatable :: name
      interface
         subroutine init_bp()
         end subroutine init_bp
      end interface
      allocate(character(len=plen) :: name)
      do k = 1, plen
         name(k:k) = path(k)
      enddo
      cfile_bpl_in = name
      flg_in_bpl = .True.
      if (allocated(bp_map)) deallocate(bp_map)
      if (allocated(bp3_map)) deallocate(bp3_map)
      call init_bp()
   end subroutine sisref_set_ss

   ! An annealing schedule as read_anneal leaves it, with what main does
   ! next (src/main.F90:90-97): tempK = the first entry, and init_ele's
   ! set_ele at that temperature; kT keeps the input temperature's value.
   subroutine sisref_set_anneal(n, steps, temps) bind(C, name="sisref_set_anneal")
      use var_state, only : opt_anneal, nanneal, anneal_step, anneal_tempK, tempK, ionic_strength
      use var_replica, only : nrep_proc
      integer(c_int), value :: n
      integer(c_int64_t), intent(in) :: steps(n)
      real(c_double), intent(in) :: temps(n)
      integer :: irep
      interface
         subroutine set_ele(irep, tempk, ionstr, out_lb, out_Zp)
            use const, only : PREC
            integer, intent(in) :: irep
            real(PREC), intent(in) :: tempk
            real(PREC), intent(in) :: ionstr
            real(PREC), intent(out), optional :: out_lb
            real(PREC), intent(out), optional :: out_Zp
         end subroutine set_ele
      end interface
      opt_anneal = 1
      nanneal = n
      if (allocated(anneal_step)) deallocate(anneal_step)
      if (allocated(anneal_tempK)) deallocate(anneal_tempK)
      allocate(anneal_step(n), anneal_tempK(n))
      anneal_step(:) = steps(:)
      anneal_tempK(:) = temps(:)
      tempK = anneal_tempK(1)
      do irep = 1, nrep_proc
         call set_ele(irep, tempK, ionic_strength)
      enddo
   end subroutine sisref_set_anneal

   ! [PBC_box] resize settings as read_input and init_pbc leave them
   ! (src/read_input.F90:821-870, src/pbc.F90:51-52)
   subroutine sisref_set_box_resize(step, change, has_target, target) bind(C, name="sisref_set_box_resize")
      use pbc, only : flg_pbc_resize, pbc_resize_step, pbc_resize_change, flg_pbc_resize_target, &
                      pbc_resize_target, pbc_resize_count, pbc_box_original, pbc_box
      integer(c_int64_t), value :: step
      integer(c_int), value :: has_target
      real(c_double), intent(in) :: change(3), target(3)
      flg_pbc_resize = .True.
      pbc_resize_step = step
      pbc_resize_change(:) = change(:)
      flg_pbc_resize_target = (has_target /= 0)
      pbc_resize_target(:) = target(:)
      pbc_resize_count(:) = 0
      pbc_box_original(:) = pbc_box(:)
   end subroutine sisref_set_box_resize

   subroutine sisref_get_box(box) bind(C, name="sisref_get_box")
      use pbc, only : pbc_box
      real(c_double), intent(out) :: box(3)
      box(:) = pbc_box(:)
   end subroutine sisref_get_box

   ! A timed Rg-bias schedule file for job_md's read_timed_bias_rg
   ! (src/job_md.F90:231-265): flg_timed_bias_rg and the file name.
   subroutine sisref_set_rg_schedule(path, plen) bind(C, name="sisref_set_rg_schedule")
      use var_io, only : cfile_bias_rg_in
      use var_potential, only : flg_timed_bias_rg
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: plen
      character(len=:), allocatable :: name
      integer :: k
      allocate(character(len=plen) :: name)
      do k = 1, plen
         name(k:k) = path(k)
      enddo
      cfile_bias_rg_in = name
      flg_timed_bias_rg = .True.
   end subroutine sisref_set_rg_schedule

   ! f-REMD: the normalised direction of the one constant-force pair and
   ! its per-replica forces, as init_tweezers sets them (src/init_tweezers.F90:24-40)
   ! Let force_tweezers write its .twz lines (src/force_tweezers.F90:50-59)
   ! into path, with the header of src/init_out_files.F90:133-139, every
   ! nsave steps of the job_md run that follows.
   subroutine sisref_open_twz(path, plen, nsave) bind(C, name="sisref_open_twz")
      use var_io, only : flg_out_twz, hdl_twz
      use var_state, only : nstep_save
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: plen
      integer(c_int64_t), value :: nsave
      character(len=:), allocatable :: name
      integer :: k
      allocate(character(len=plen) :: name)
      do k = 1, plen
         name(k:k) = path(k)
      enddo
      hdl_twz = 95
      open(hdl_twz, file=name, status='replace', action='write', form='formatted')
      write(hdl_twz, '(a)') '#  (1)step  (2)id (3)d[traps]   (4)d[beads]   (5)d[t1-b1]   (6)d[t2-b2]   (7)f[b1]      (8)f[b2]'
      flg_out_twz = .True.
      twz_nsave = nsave
   end subroutine sisref_open_twz

   subroutine sisref_set_dcf_remd(dir) bind(C, name="sisref_set_dcf_remd")
      use var_potential, only : twz_DCF_direction, twz_DCF_forces
      use var_replica, only : nrep_proc, rep2val, irep2grep
      use const_idx, only : REPT
      real(c_double), intent(in) :: dir(3)
      integer :: irep
      if (allocated(twz_DCF_direction)) deallocate(twz_DCF_direction)
      allocate(twz_DCF_direction(3, 1))
      twz_DCF_direction(:, 1) = dir(:) / norm2(dir(:))
      do irep = 1, nrep_proc
         twz_DCF_forces(:, 1, irep) = rep2val(irep2grep(irep), REPT%TWZDCF) * twz_DCF_direction(:, 1)
      enddo
   end subroutine sisref_set_dcf_remd

   subroutine sisref_ele_params(irep, out) bind(C, name="sisref_ele_params")
      use var_state, only : lambdaD, diele
      use var_potential, only : ele_coef, ele_cutoff
      integer(c_int), value :: irep
      real(c_double), intent(out) :: out(4)
      out(1) = lambdaD(irep)
      out(2) = ele_coef(irep)
      out(3) = ele_cutoff(irep)
      out(4) = diele(irep)
   end subroutine sisref_ele_params

   ! bp3 triplet table (src/init_bp.F90:39-172) and the BP cutoffs
   subroutine sisref_bp_tables(dh, ds, cut) bind(C, name="sisref_bp_tables")
      use var_potential, only : bp3_dH, bp3_dS, bp_cutoff_dist, bp_paras
      real(c_double), intent(out) :: dh(468), ds(468), cut(7)
      integer :: k
      dh(:) = bp3_dH(:)
      ds(:) = bp3_dS(:)
      cut(1) = bp_cutoff_dist
      do k = 1, 6
         cut(k+1) = bp_paras(k)%cutoff_ddist
      enddo
   end subroutine sisref_bp_tables

   ! dense bp_map / bp3_map (src/init_bp.F90:28-29), (nmp, nmp) column-major
   subroutine sisref_bp_maps(bpm, bp3m) bind(C, name="sisref_bp_maps")
      use var_potential, only : bp_map, bp3_map
      use var_top, only : nmp
      integer(c_int), intent(out) :: bpm(nmp, nmp), bp3m(nmp, nmp)
      bpm(:,:) = bp_map(:,:)
      bp3m(:,:) = bp3_map(:,:)
   end subroutine sisref_bp_maps

   ! n raw tempered MT19937 words from a deep copy of mts(irep)
   ! (irep = -1 selects mts_rep); the stream itself does not advance.
   subroutine sisref_mt_tape(irep, n, words) bind(C, name="sisref_mt_tape")
      use mt_stream, only : mt_state, genrand_int32
      use var_state, only : mts, mts_rep
      integer(c_int), value :: irep
      integer(c_int64_t), value :: n
      integer(c_int32_t), intent(out) :: words(n)
      type(mt_state) :: tmp
      integer(c_int64_t) :: k
      if (irep < 0) then
         tmp = mts_rep
         allocate(tmp%state(0:tmp%nn-1))
         tmp%state(:) = mts_rep%state(:)
      else
         tmp = mts(irep)
         allocate(tmp%state(0:tmp%nn-1))
         tmp%state(:) = mts(irep)%state(:)
      endif
      do k = 1, n
         words(k) = genrand_int32(tmp)
      enddo
      deallocate(tmp%state)
   end subroutine sisref_mt_tape

   subroutine sisref_genrand(irep, kind, n, out) bind(C, name="sisref_genrand")
      use mt_stream, only : genrand_double1, genrand_double3
      use var_state, only : mts
      integer(c_int), value :: irep, kind
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: out(n)
      integer(c_int64_t) :: k
      do k = 1, n
         if (kind == 1) then
            out(k) = genrand_double1(mts(irep))
         else
            out(k) = genrand_double3(mts(irep))
         endif
      enddo
   end subroutine sisref_genrand

   subroutine sisref_boxmuller(irep, n, out) bind(C, name="sisref_boxmuller")
      integer(c_int), value :: irep
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: out(n)
      real(c_double) :: rnd_boxmuller
      integer(c_int64_t) :: k
      do k = 1, n
         out(k) = rnd_boxmuller(irep)
      enddo
   end subroutine sisref_boxmuller

   ! Run the reference MD loop for nstep steps (single use per process:
   ! job_md allocates its arrays and keeps first-call state).
   subroutine sisref_job_md(nstep_in, energies_out) bind(C, name="sisref_job_md")
      use var_state, only : nstep, nstep_save, energies
      use var_replica, only : nrep_proc
      integer(c_int64_t), value :: nstep_in
      real(c_double), intent(out) :: energies_out(0:10, *)
      nstep = nstep_in
      nstep_save = max(nstep_in, 1_c_int64_t)
      if (twz_nsave > 0) nstep_save = twz_nsave
      call job_md()
      energies_out(:, 1:nrep_proc) = energies(:, 1:nrep_proc)
   end subroutine sisref_job_md

   ! Write every replica's restart file with the reference's own writer
   ! (src/write_rst.F90): replica irep goes to <prefix><irep, i4.4>.rst.
   subroutine sisref_write_rst(prefix, plen) bind(C, name="sisref_write_rst")
      use var_io, only : hdl_rst, cfile_rst
      use var_replica, only : nrep_proc
      character(kind=c_char), intent(in) :: prefix(*)
      integer(c_int), value :: plen
      character(len=:), allocatable :: pre
      character(len=4) :: tag
      integer :: irep, k
      interface
         subroutine write_rst()
         end subroutine write_rst
      end interface
      allocate(character(len=plen) :: pre)
      do k = 1, plen
         pre(k:k) = prefix(k)
      enddo
      if (allocated(hdl_rst)) deallocate(hdl_rst)
      if (allocated(cfile_rst)) deallocate(cfile_rst)
      allocate(hdl_rst(nrep_proc), cfile_rst(nrep_proc))
      do irep = 1, nrep_proc
         write(tag, '(i4.4)') irep
         hdl_rst(irep) = 700 + irep
         cfile_rst(irep) = pre // tag // '.rst'
      enddo
      call write_rst()
   end subroutine sisref_write_rst

   subroutine sisref_replica_labels(rep2lab_out) bind(C, name="sisref_replica_labels")
      use var_replica, only : rep2lab, nrep_all
      integer(c_int), intent(out) :: rep2lab_out(*)
      rep2lab_out(1:nrep_all) = rep2lab(1:nrep_all)
   end subroutine sisref_replica_labels

   ! Write nframe frames of xyz(3, nmp, nframe) to a DCD file with the
   ! reference's own writer (file_dcd of src/dcd.F90: write_header,
   ! write_onestep), for byte-level comparison with other writers.
   subroutine sisref_write_dcd(path, plen, nmp_in, nframe, xyz, unitcell, box, fix_com) &
                               bind(C, name="sisref_write_dcd")
      use dcd, only : file_dcd, DCD_OPEN_MODE
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: plen, nmp_in, nframe, unitcell, fix_com
      real(c_double), intent(in) :: xyz(3, nmp_in, *)
      real(c_double), intent(in) :: box(3)
      type(file_dcd) :: f
      character(len=:), allocatable :: fname
      integer :: k
      allocate(character(len=plen) :: fname)
      do k = 1, plen
         fname(k:k) = path(k)
      enddo
      f = file_dcd(93, fname, DCD_OPEN_MODE%WRITE)
      if (unitcell /= 0) then
         f%flg_unitcell = .True.
         f%box(:) = box(:)
      endif
      call f%write_header(nmp_in)
      do k = 1, nframe
         call f%write_onestep(nmp_in, xyz(:, :, k), fix_com)
      enddo
      call f%close()
   end subroutine sisref_write_dcd

   ! One .out file through the reference's write_out_header and write_out
   ! (src/write_out.F90), for nline lines of (istep, label, tK, Ekin,
   ! energies(0:10)); used to compare other writers byte by byte.
   subroutine sisref_write_out(path, plen, nline, isteps, labels, temps, ekin, evals, replica, ele) &
                               bind(C, name="sisref_write_out")
      use, intrinsic :: iso_fortran_env, only : INT64
      use const_idx, only : ENE
      use var_io, only : hdl_out
      use var_state, only : energies, Ekinetic
      use var_replica, only : flg_replica
      use var_potential, only : flg_ele, flg_stage, flg_twz, flg_bias_rg, flg_restraint
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: plen, nline, replica, ele
      integer(c_int64_t), intent(in) :: isteps(nline)
      integer(c_int), intent(in) :: labels(nline)
      real(c_double), intent(in) :: temps(nline), ekin(nline), evals(0:10, nline)
      character(len=:), allocatable :: fname
      integer :: k
      integer(INT64) :: istep_out
      allocate(character(len=plen) :: fname)
      do k = 1, plen
         fname(k:k) = path(k)
      enddo
      if (.not. allocated(hdl_out)) allocate(hdl_out(1))
      if (.not. allocated(energies)) allocate(energies(0:ENE%MAX, 1))
      if (.not. allocated(Ekinetic)) allocate(Ekinetic(1))
      flg_replica = replica /= 0
      flg_ele = ele /= 0
      flg_stage = .False.
      flg_twz = .False.
      flg_bias_rg = .False.
      flg_restraint = .False.
      hdl_out(1) = 94
      open(hdl_out(1), file=fname, status='replace', action='write', form='formatted')
      call write_out_header(1)
      do k = 1, nline
         energies(0:10, 1) = evals(0:10, k)
         Ekinetic(1) = ekin(k)
         istep_out = isteps(k)
         call write_out(1, istep_out, labels(k), temps(k))
      enddo
      close(hdl_out(1))
   end subroutine sisref_write_out

end module sisref_driver
