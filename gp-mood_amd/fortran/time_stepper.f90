! Drop-in replacement of the reference's time_stepper module
! (gp-mood-code/time_stepper.f90:1-25): same module and subroutine names and
! the same signature time_stepping(Uin, Uout), so main.f90, mod_*.f90 and the
! parameters.f90 / problems interface compile against it unchanged.  The step of
! the configured integrator (FE, SSP_RK2, SSP_RK3 or SSP_RK4, time_stepper.f90:17-21)
! runs on the GPU through libgpmood; the GP tables
! are the ones GP_presim_sphere left in global_variables, so both codes use the
! same weights.  With method NN_GP_MOOD or NN_GP_MOOD_CC the network is the one
! load_NN read into global_variables (main.f90:116), and the step sets the
! counters mod_FE.f90 keeps for it.  With write_NN_dataset the first stage's
! samples go into the dataset arrays inputs / labels with index, n_overwrite,
! NR0, NR1 and freq_R0 as append_to_NN_datatset leaves them
! (mod_FE.f90:176-181); which order-3 samples are kept is drawn by the library.
! Link with -lgpmood.
module time_stepper
   use iso_c_binding
   use parameters
   use constants
   use global_variables
   use gpm_capi
   implicit none

   type(c_ptr), save :: gpm_handle = c_null_ptr
   integer(c_int32_t), allocatable, save :: gpm_order(:,:)

contains

   subroutine gpm_open()
      type(gpm_config) :: cfg
      integer(c_int) :: rc
      integer(c_int32_t), allocatable :: ixiy32(:,:,:)
      cfg%lf = lf; cfg%nf = nf; cfg%mord = Mord; cfg%ngp = ngp
      cfg%method = method; cfg%integrator = integrator; cfg%num_flux = numFlux; cfg%bc_type = BC_type
      cfg%dmp = merge(1, 0, DMP); cfg%u2 = merge(1, 0, U2); cfg%u2_tol = merge(1, 0, U2_tol)
      cfg%cascade = GPM_CASCADE_REFERENCE
      cfg%problem = problem
      cfg%cfl = CFL; cfg%lx = Lx; cfg%ly = Ly; cfg%ell_o_dx = real(ell_o_dx, c_double)
      rc = gpm_create(cfg, gpm_handle)
      if (rc /= 0) then
         print *, 'gpm_create failed: ', rc
         stop 1
      end if
      allocate(ixiy32(7, 25, 2))
      ixiy32 = int(ixiy_sp, c_int32_t)
      rc = gpm_set_gp_tables(gpm_handle, zT_sp, GP_d2x_sp, GP_d2y_sp, gauss_weight, ixiy32)
      if (rc /= 0) then
         print *, 'gpm_set_gp_tables failed: ', trim(gpm_error_message(gpm_handle))
         stop 1
      end if
      allocate(gpm_order(lb:le, nb:ne))
      if (write_NN_dataset) then
         rc = gpm_set_reference_dataset(gpm_handle, int(L, c_int32_t), int(NR0, c_int32_t), int(NR1, c_int32_t))
         if (rc /= 0) then
            print *, 'gpm_set_reference_dataset failed: ', trim(gpm_error_message(gpm_handle))
            stop 1
         end if
      end if
      if ((method == NN_GP_MOOD) .or. (method == NN_GP_MOOD_CC)) then
         rc = gpm_set_reference_nn(gpm_handle, int(L, c_int32_t), int(length, c_int32_t), weight0, bias0, &
            weight1, bias1, weight2, bias2)
         if (rc /= 0) then
            print *, 'gpm_set_reference_nn failed: ', trim(gpm_error_message(gpm_handle))
            stop 1
         end if
      end if
   end subroutine gpm_open

   subroutine time_stepping(Uin, Uout)
      real(PR), intent(inout), dimension(4,lb:le, nb:ne) :: Uin
      real(PR), intent(out  ), dimension(4,lb:le, nb:ne) :: Uout
      type(gpm_stats) :: st
      integer(c_int) :: rc

      if (.not. c_associated(gpm_handle)) call gpm_open()
      t_rk = t
      ! main.f90:147 has advanced niter for this step; the library advances it itself
      rc = gpm_set_niter(gpm_handle, int(niter - 1, c_int64_t))
      rc = gpm_time_stepping(gpm_handle, Uin, Uout, t, dt, gpm_order, st)
      if (rc /= 0) then
         print *, 'gpm_time_stepping failed: ', trim(gpm_error_message(gpm_handle))
         stop 1
      end if
      if (write_NN_dataset) then
         ! append_to_NN_datatset after the first stage of this step (mod_FE.f90:178-181)
         steps_NN_sample(niter) = 1
         rc = gpm_append_reference_dataset(gpm_handle, inputs, labels, int(dataset_size, c_int32_t), index, &
            n_overwrite, NR0, NR1)
         if (rc /= 0) then
            print *, 'gpm_append_reference_dataset failed: ', trim(gpm_error_message(gpm_handle))
            stop 1
         end if
         freq_R0 = NR0*one/(NR0+NR1)
      end if
      ! t_rk as the integrator leaves it (mod_SSP_RK3.f90:31 sets the last stage's)
      if (integrator == SSP_RK3) t_rk = t + 0.5*dt
      CellGPO = gpm_order
      count_detected_cell = int(st%count_detected_cell)
      count_detected_cell_a_priori = int(st%count_detected_cell_a_priori)
      count_detected_cell_a_posteriori = count_detected_cell - count_detected_cell_a_priori
      if ((method == NN_GP_MOOD) .or. (method == NN_GP_MOOD_CC)) then
         ! CellGPO_priori, steps_NN_produced_NAN and count_steps_NN_produced_NAN (mod_FE.f90:49-56,154-157)
         rc = gpm_get_cellgpo_priori(gpm_handle, gpm_order)
         CellGPO_priori = gpm_order
         if (st%nn_nan_passes > 0) then
            steps_NN_produced_NAN(niter) = 1
            count_steps_NN_produced_NAN = count_steps_NN_produced_NAN + int(st%nn_nan_passes)
         end if
      end if
   end subroutine time_stepping

end module time_stepper
