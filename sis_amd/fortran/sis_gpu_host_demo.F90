! Fortran host that drives libsis_gpu.so through module sis_gpu_iface in
! the order a patched job_md uses it (src/job_md.F90:196-198 neighbor_list,
! :290 energy_sumup, :397 force, :360-460 the time loop).
!
! Usage: sis_gpu_host_demo <in.bin> <out.bin>
!   in.bin  (stream, native byte order): one sis_params record; int32
!           nchains, nmp, nrep, precision; int32 nmp_chain(nchains),
!           moltype(nchains), seq(nmp); real64 xyz(3, nmp, nrep);
!           int64 nsteps, seed
!   out.bin : real64 forces(3, nmp) of replica 1 after the first rebuild,
!           energies(0:10) of replica 1 (ENE order, src/const_idx.F90:21),
!           xyz(3, nmp, nrep) after nsteps; int64 rebuilds
! A failing call prints sis_gpu_last_error() and stops with status 1, as
! sis_abort would (src/sis_abort.F90:1-20).
program sis_gpu_host_demo
   use, intrinsic :: iso_c_binding
   use sis_gpu_iface
   implicit none

   type(sis_params) :: prm
   type(sis_topo) :: topo
   type(sis_options) :: opt
   type(c_ptr) :: ctx
   integer(c_int32_t), allocatable, target :: nmp_chain(:), moltype(:), seq(:)
   real(c_double), allocatable, target :: xyz(:,:,:), forces(:,:)
   real(c_double) :: energies(0:10)
   integer(c_int32_t) :: nchains, nmp, nrep, precision
   integer(c_int64_t) :: nsteps, seed, nreb
   character(len=1024) :: fin, fout
   integer :: u

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
   read(u) prm
   read(u) nchains, nmp, nrep, precision
   allocate(nmp_chain(nchains), moltype(nchains), seq(nmp), xyz(3, nmp, nrep), forces(3, nmp))
   read(u) nmp_chain, moltype, seq, xyz, nsteps, seed
   close(u)

   topo%nchains = nchains
   topo%nmp = nmp
   topo%nmp_chain = c_loc(nmp_chain)
   topo%moltype = c_loc(moltype)
   topo%seq = c_loc(seq)
   topo%no_charge = c_null_ptr
   topo%n_no_charge = 0
   topo%pad = 0

   opt%nrep = nrep
   opt%precision = precision
   opt%parity = 0
   opt%mode = 0
   opt%block = 0
   opt%max_nbr = 0
   opt%rebuild_every = 0
   opt%exchange_every = 0
   opt%seed = seed

   call check(sis_gpu_create(topo, prm, opt, ctx), 'sis_gpu_create')
   call check(sis_gpu_set_state(ctx, c_loc(xyz), c_null_ptr, c_null_ptr), 'sis_gpu_set_state')
   call check(sis_gpu_rebuild(ctx, 0_c_int32_t), 'neighbor_list')
   call check(sis_gpu_force(ctx, 1_c_int32_t, 1_c_int32_t, c_loc(forces)), 'force')
   ! bit 1: reuse the base-pair state of the force call just made
   call check(sis_gpu_energy(ctx, 1_c_int32_t, prm%tempK, 3_c_int32_t, energies), 'energy_sumup')
   call check(sis_gpu_init_velocities(ctx), 'init_velocities')
   nreb = 0
   call check(sis_gpu_step(ctx, nsteps, nreb), 'job_md steps')
   call check(sis_gpu_get_state(ctx, c_loc(xyz), c_null_ptr, c_null_ptr), 'sis_gpu_get_state')
   call check(sis_gpu_destroy(ctx), 'sis_gpu_destroy')

   open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(u) forces, energies, xyz, nreb
   close(u)

contains

   subroutine check(rc, what)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: what
      if (rc /= SIS_OK) then
         write(*, '(a, a, a, i0, a, a)') 'sis_gpu: ', what, ' failed (', rc, '): ', sis_gpu_error_message()
         error stop 1
      endif
   end subroutine check

end program sis_gpu_host_demo
