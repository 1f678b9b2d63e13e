! iso_c_binding interface of libgpmood (include/gpmood.h) for Fortran callers.
module gpm_capi
   use iso_c_binding
   implicit none

   integer(c_int32_t), parameter :: GPM_CASCADE_REFERENCE = 0, GPM_CASCADE_FULL = 1

   type, bind(C) :: gpm_config
      integer(c_int32_t) :: lf, nf, mord, ngp, method, integrator, num_flux, bc_type
      integer(c_int32_t) :: dmp, u2, u2_tol, cascade, problem
      real(c_double) :: cfl, lx, ly, ell_o_dx
   end type gpm_config

   type, bind(C) :: gpm_stats
      integer(c_int64_t) :: steps, count_detected_cell, detected_total
      integer(c_int32_t) :: passes(3), max_passes
      integer(c_int64_t) :: second_pass_cells, a_priori_lowered
      real(c_double) :: t, dt
      integer(c_int64_t) :: count_detected_cell_a_priori, nn_nan_passes, nn_nan_passes_total, niter
   end type gpm_stats

   interface
      integer(c_int) function gpm_create(cfg, ctx) bind(C, name='gpm_create')
         import :: c_int, c_ptr, gpm_config
         type(gpm_config), intent(in) :: cfg
         type(c_ptr), intent(out) :: ctx
      end function gpm_create

      integer(c_int) function gpm_destroy(ctx) bind(C, name='gpm_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function gpm_destroy

      type(c_ptr) function gpm_last_error(ctx) bind(C, name='gpm_last_error')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function gpm_last_error

      integer(c_int) function gpm_set_gp_tables(ctx, zT_sp, gp_d2x_sp, gp_d2y_sp, gauss_weight, ixiy_sp) &
         bind(C, name='gpm_set_gp_tables')
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: zT_sp(*), gp_d2x_sp(*), gp_d2y_sp(*), gauss_weight(*)
         integer(c_int32_t), intent(in) :: ixiy_sp(*)
      end function gpm_set_gp_tables

      integer(c_int) function gpm_set_state(ctx, U) bind(C, name='gpm_set_state')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: U(*)
      end function gpm_set_state

      integer(c_int) function gpm_get_state(ctx, U, cellgpo) bind(C, name='gpm_get_state')
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: U(*)
         integer(c_int32_t), intent(out) :: cellgpo(*)
      end function gpm_get_state

      integer(c_int) function gpm_compute_dt(ctx, dtfinal, dt) bind(C, name='gpm_compute_dt')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: dtfinal
         real(c_double), intent(out) :: dt
      end function gpm_compute_dt

      integer(c_int) function gpm_time_stepping(ctx, Uin, Uout, t, dt, cellgpo, stats) &
         bind(C, name='gpm_time_stepping')
         import :: c_int, c_ptr, c_double, c_int32_t, gpm_stats
         type(c_ptr), value :: ctx
         real(c_double), intent(inout) :: Uin(*)
         real(c_double), intent(out) :: Uout(*)
         real(c_double), value :: t, dt
         integer(c_int32_t), intent(out) :: cellgpo(*)
         type(gpm_stats), intent(out) :: stats
      end function gpm_time_stepping

      integer(c_int) function gpm_run(ctx, nsteps, tmax, dts) bind(C, name='gpm_run')
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nsteps
         real(c_double), value :: tmax
         real(c_double), intent(out) :: dts(*)
      end function gpm_run

      ! the reference's own network: the arrays load_NN filled (global_variables.f90:74-80)
      integer(c_int) function gpm_set_reference_nn(ctx, L, length, weight0, bias0, weight1, bias1, weight2, bias2) &
         bind(C, name='gpm_set_reference_nn')
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: L, length
         real(c_float), intent(in) :: weight0(*), bias0(*), weight1(*), bias1(*), weight2(*), bias2(*)
      end function gpm_set_reference_nn

      integer(c_int) function gpm_set_niter(ctx, niter) bind(C, name='gpm_set_niter')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: niter
      end function gpm_set_niter

      integer(c_int) function gpm_get_cellgpo_priori(ctx, priori) bind(C, name='gpm_get_cellgpo_priori')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), intent(out) :: priori(*)
      end function gpm_get_cellgpo_priori

      integer(c_int) function gpm_set_reference_dataset(ctx, L, nr0, nr1) bind(C, name='gpm_set_reference_dataset')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: L, nr0, nr1
      end function gpm_set_reference_dataset

      integer(c_int) function gpm_append_reference_dataset(ctx, inputs, labels, dataset_size, index, n_overwrite, &
         nr0, nr1) bind(C, name='gpm_append_reference_dataset')
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: ctx
         real(c_float), intent(inout) :: inputs(*), labels(*)
         integer(c_int32_t), value :: dataset_size
         integer(c_int32_t), intent(inout) :: index, n_overwrite, nr0, nr1
      end function gpm_append_reference_dataset
   end interface

contains

   ! Message of the last failed call, as a Fortran string.
   function gpm_error_message(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=256) :: msg
      character(kind=c_char), pointer :: p(:)
      integer :: i
      msg = ''
      if (.not. c_associated(gpm_last_error(ctx))) return
      call c_f_pointer(gpm_last_error(ctx), p, [256])
      do i = 1, 256
         if (p(i) == c_null_char) exit
         msg(i:i) = p(i)
      end do
   end function gpm_error_message
end module gpm_capi
