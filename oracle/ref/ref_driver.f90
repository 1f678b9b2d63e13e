! Harness driver for the reference GP-MOOD Fortran (test infrastructure only).
!
! This program is written for this repository.  It links against the reference
! modules compiled from their own sources (constants, parameters, global_variables,
! GP_init, mesh_init, physics, BC, mod_FE, mod_SSP_RK3, time_stepper, ...) and
! replaces only the reference's `main.f90` driver, whose HDF5 I/O cannot be built
! with this image's Fortran compiler.  It performs the reference step loop of
! main.f90:145-162 (dtfinal, Setdt, time_stepping, Boundary_C, t=t+dt, U=Ur) on a
! conserved initial state read from a raw binary file, and dumps:
!   tables.bin   GP tables from GP_presim_sphere (GP_init.f90:26-357)
!   pol.bin      Pol_zT_o3 from pol_presim (GP_init.f90:359-411)
!   steps.bin    per step: dt (f64), t (f64), count_detected_cell (i32)
!   masks.bin    CellGPO(1:lf,1:nf) as int8 after every RK stage (mode 1) or
!                after every step (mode 0)
!   U_final.bin  U(4,lb:le,nb:ne) after the last step
!   U_stage.bin  (mode 1) U after each RK stage of the first step
!   nn.bin       (methods NN_GP_MOOD, NN_GP_MOOD_CC) per step:
!                count_detected_cell_a_priori, count_detected_cell_a_posteriori,
!                count_steps_NN_produced_NAN (i32 each)
!   priori.bin   (the same methods) CellGPO_priori(1:lf,1:nf) as int8 per step
!   dataset.bin  (write_NN_dataset) index, n_overwrite, NR0, NR1, dataset_size, L
!                (i32 each), inputs(dataset_size, L), labels(dataset_size, 2)
!                (f32, Fortran order), steps_NN_sample(1:nsteps) (i32)
! The NN methods load their network with load_NN (mod_NN.f90:60-117) first, as
! main.f90:116 does.
!
! Usage: gpm_ref.exe <ic.bin> <nsteps> <outdir> <mode> [restart.bin]
!   restart.bin (optional): the restart of main.f90:124-136 -- after the
!           initial condition and Boundary_C(U), the interior of U is
!           overwritten from this file (conserved, as read() leaves it) with
!           no refill of the ghost cells, before the step loop
!   mode 0: call time_stepping (time_stepper.f90:12-23) exactly as main.f90:155
!   mode 1: the three Forward_Euler stages of SSP_RK3_ (mod_SSP_RK3.f90:12-39)
!           written out here so that the order mask can be dumped per stage
program gpm_ref_driver
   use constants
   use parameters
   use global_variables
   use mesh_init
   use GP_init
   use physics
   use BC
   use mod_FE
   use time_stepper
   use mod_NN
   implicit none

   character(len=1024) :: icfile, outdir, arg
   integer :: nsteps, mode, s, uo, um, us, ust, unn, upr
   logical :: nnm
   real(PR), allocatable :: Uint(:,:,:)
   real(PR), allocatable :: U1(:,:,:), U2s(:,:,:)
   real(PR) :: tic, tac
   integer(1), allocatable :: m8(:,:)

   call get_command_argument(1, icfile)
   call get_command_argument(2, arg); read(arg,*) nsteps
   call get_command_argument(3, outdir)
   call get_command_argument(4, arg); read(arg,*) mode

   call init_mesh()
   call GP_presim_sphere()
   call pol_presim()
   nnm = (method == NN_GP_MOOD) .or. (method == NN_GP_MOOD_CC)
   if (nnm) call load_NN()

   open(newunit=uo, file=trim(outdir)//'/tables.bin', access='stream', form='unformatted', status='replace')
   write(uo) int(Mord,4), int(ngp,4), int(sz_sphere,4), int(lf,4), int(nf,4), int(BC_type,4)
   write(uo) zT_sp
   write(uo) GP_d2x_sp
   write(uo) GP_d2y_sp
   write(uo) gauss_weight
   write(uo) int(ixiy_sp,4)
   write(uo) dx, dy
   close(uo)
   open(newunit=uo, file=trim(outdir)//'/pol.bin', access='stream', form='unformatted', status='replace')
   write(uo) Pol_zT_o3
   close(uo)

   allocate(Uint(4,lf,nf))
   open(newunit=uo, file=trim(icfile), access='stream', form='unformatted', status='old')
   read(uo) Uint
   close(uo)
   U = 0.
   U(:,1:lf,1:nf) = Uint
   call Boundary_C(U)
   if (command_argument_count() >= 5) then
      call get_command_argument(5, arg)
      open(newunit=uo, file=trim(arg), access='stream', form='unformatted', status='old')
      read(uo) Uint
      close(uo)
      U(:,1:lf,1:nf) = Uint
   end if
   deallocate(Uint)

   allocate(m8(lf,nf))
   if (mode == 1) then
      allocate(U1(4,lb:le,nb:ne), U2s(4,lb:le,nb:ne))
   end if

   open(newunit=uo, file=trim(outdir)//'/steps.bin', access='stream', form='unformatted', status='replace')
   open(newunit=um, file=trim(outdir)//'/masks.bin', access='stream', form='unformatted', status='replace')
   if (mode == 1) open(newunit=ust, file=trim(outdir)//'/U_stage.bin', access='stream', form='unformatted', status='replace')
   if (nnm) then
      open(newunit=unn, file=trim(outdir)//'/nn.bin', access='stream', form='unformatted', status='replace')
      open(newunit=upr, file=trim(outdir)//'/priori.bin', access='stream', form='unformatted', status='replace')
   end if

   niter = 0
   t = 0.
   call cpu_time(tic)
   do s = 1, nsteps
      niter = niter + 1
      dtfinal = tmax - t
      call Setdt(U)
      if (mode == 0) then
         call time_stepping(U, Ur)
         m8 = int(CellGPO(1:lf,1:nf), 1)
         write(um) m8
      else
         ! Same statements as SSP_RK3_ (mod_SSP_RK3.f90:23-33), with mask dumps.
         t_rk = t
         call Forward_Euler(U, U1, .true.)
         m8 = int(CellGPO(1:lf,1:nf), 1); write(um) m8
         if (s == 1) write(ust) U1
         t_rk = t + dt
         call Forward_Euler(U1, U2s, .false.)
         m8 = int(CellGPO(1:lf,1:nf), 1); write(um) m8
         U2s = (0.75)*U + 0.25*U2s
         if (s == 1) write(ust) U2s
         t_rk = t + 0.5*dt
         call Forward_Euler(U2s, Ur, .false.)
         m8 = int(CellGPO(1:lf,1:nf), 1); write(um) m8
         Ur = (1./3)*U + (2./3)*Ur
         if (s == 1) write(ust) Ur
      end if
      call Boundary_C(U)
      t = t + dt
      U = Ur
      write(uo) dt, t, int(count_detected_cell,4)
      if (nnm) then
         write(unn) int(count_detected_cell_a_priori,4), int(count_detected_cell_a_posteriori,4), &
            int(count_steps_NN_produced_NAN,4)
         m8 = int(CellGPO_priori(1:lf,1:nf), 1)
         write(upr) m8
      end if
   end do
   call cpu_time(tac)
   close(uo)
   close(um)
   if (mode == 1) close(ust)
   if (nnm) then
      close(unn)
      close(upr)
   end if

   open(newunit=us, file=trim(outdir)//'/U_final.bin', access='stream', form='unformatted', status='replace')
   write(us) U
   close(us)

   if (write_NN_dataset) then
      ! the training-sample arrays of append_to_NN_datatset and their counters
      open(newunit=us, file=trim(outdir)//'/dataset.bin', access='stream', form='unformatted', status='replace')
      write(us) int(index,4), int(n_overwrite,4), int(NR0,4), int(NR1,4), int(dataset_size,4), int(L,4)
      write(us) inputs
      write(us) labels
      write(us) int(steps_NN_sample(1:nsteps),4)
      close(us)
   end if

   write(*,'(A,I0,A,I0,A,I0,A,ES12.5,A,ES12.5)') 'N=', lf, ' steps=', nsteps, ' mode=', mode, &
      ' cpu_s=', tac-tic, ' cell_stage_per_s=', 3.0d0*real(lf,8)*real(nf,8)*real(nsteps,8)/max(tac-tic,1.0d-9)
end program gpm_ref_driver
