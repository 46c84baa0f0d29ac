! lbcuda_amd_mod.f90 -- ISO_C_BINDING interface to liblbcuda_amd (include/lbcuda_amd.h).
!
! The reference host program (main.CUF) launches CUDA-Fortran kernels with
! chevrons and keeps device arrays in module globals (kernels_fluid.CUF).  A
! Fortran host calls the MI355X library through these bind(C) interfaces
! instead; arguments are plain values and C pointers, arrays are passed as
! contiguous Fortran arrays (x fastest), so no transposes are needed.
module lbcuda_amd
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: LBM_STORE_F32 = 0, LBM_STORE_F16 = 1, LBM_STORE_BF16 = 2
  integer(c_int32_t), parameter :: LBM_WET_NONE = 0, LBM_WET_JANUS = 1, LBM_WET_NEUTRAL = 2
  integer(c_int32_t), parameter :: LBM_INIT_DROPLET = 1, LBM_INIT_ZPLANES = 2, LBM_INIT_RHO = 3, &
                                   LBM_INIT_POPS = 4, LBM_INIT_CYLINDER = 5, LBM_INIT_RANDOM = 6, &
                                   LBM_INIT_VELFIELD = 7
  integer(c_int32_t), parameter :: LBM_KERNEL_FUSED = 0, LBM_KERNEL_TWOPASS = 1

  ! struct lbm_config, field for field
  type, bind(C) :: lbm_config
    integer(c_int32_t) :: nx, ny, nz
    real(c_float)      :: tau_r, tau_b
    real(c_float)      :: sigma, beta
    integer(c_int32_t) :: storage
    real(c_float)      :: store_rho0
    real(c_float)      :: min_rho, max_rho
    integer(c_int32_t) :: with_particles, n_particles, lrotate
    real(c_float)      :: radius, rmass, max_vel
    real(c_float)      :: ext_force(3), ext_torque(3)
    integer(c_int32_t) :: wetting, hertz
    real(c_float)      :: hz_k, hz_rmin, hz_rcap
    integer(c_int32_t) :: lubrication
    real(c_float)      :: lub_k, lub_rmin, lub_rcap
    integer(c_int32_t) :: closure
    integer(c_int32_t) :: diffdens
    real(c_float)      :: dens_r, dens_b
    integer(c_int32_t) :: reserved(5)
  end type lbm_config

  interface
    subroutine lbm_config_default(cfg, n) bind(C, name="lbm_config_default")
      import :: lbm_config, c_int32_t
      type(lbm_config), intent(out) :: cfg
      integer(c_int32_t), value :: n
    end subroutine

    integer(c_int) function lbm_create(cfg, ctx) bind(C, name="lbm_create")
      import :: lbm_config, c_int, c_ptr
      type(lbm_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function lbm_destroy(ctx) bind(C, name="lbm_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function lbm_init_fields(ctx, init_type, u0, rho_r, rho_b, vel) &
        bind(C, name="lbm_init_fields")
      import :: c_int, c_int32_t, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: init_type
      real(c_float), intent(in) :: u0(3)
      type(c_ptr), value :: rho_r, rho_b, vel
    end function

    integer(c_int) function lbm_set_pops(ctx, pops_r, pops_b) bind(C, name="lbm_set_pops")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, pops_r, pops_b
    end function

    integer(c_int) function lbm_set_particles(ctx, n, pos) bind(C, name="lbm_set_particles")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n
      real(c_double), intent(in) :: pos(13, *)
    end function

    integer(c_int) function lbm_step(ctx, nsteps) bind(C, name="lbm_step")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function

    integer(c_int) function lbm_get_moments(ctx, rho_r, rho_b, vel, phase) bind(C, name="lbm_get_moments")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rho_r, rho_b, vel, phase
    end function

    integer(c_int) function lbm_get_pops(ctx, color, pops) bind(C, name="lbm_get_pops")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: color
      type(c_ptr), value :: pops
    end function

    integer(c_int) function lbm_get_rho(ctx, rho_r, rho_b) bind(C, name="lbm_get_rho")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rho_r, rho_b
    end function

    integer(c_int) function lbm_get_particles(ctx, pos, force) bind(C, name="lbm_get_particles")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, pos, force
    end function

    integer(c_int) function lbm_get_sums(ctx, sums) bind(C, name="lbm_get_sums")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: sums(5)
    end function

    integer(c_int) function lbm_checkpoint_size(ctx, nbytes) bind(C, name="lbm_checkpoint_size")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: nbytes
    end function

    integer(c_int) function lbm_checkpoint_save(ctx, buf, nbytes) bind(C, name="lbm_checkpoint_save")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx, buf
      integer(c_int64_t), value :: nbytes
    end function

    integer(c_int) function lbm_checkpoint_load(ctx, buf, nbytes) bind(C, name="lbm_checkpoint_load")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx, buf
      integer(c_int64_t), value :: nbytes
    end function

    integer(c_int) function lbm_status(ctx, code, step, msg, msg_len) bind(C, name="lbm_status")
      import :: c_int, c_int32_t, c_ptr, c_char
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: code, step
      character(kind=c_char), intent(out) :: msg(*)
      integer(c_int32_t), value :: msg_len
    end function

    integer(c_int) function lbm_fill_random_rho(seed, mean, amp, nx, ny, nz, rho_r, rho_b) &
        bind(C, name="lbm_fill_random_rho")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      integer(c_int64_t), value :: seed
      real(c_double), value :: mean, amp
      integer(c_int32_t), value :: nx, ny, nz
      type(c_ptr), value :: rho_r, rho_b
    end function

    integer(c_int) function lbm_set_kernel(ctx, variant, tile, zchunk) bind(C, name="lbm_set_kernel")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: variant, tile, zchunk
    end function

    integer(c_int) function lbm_autotune(ctx, reps, tile, zchunk, best_ms) bind(C, name="lbm_autotune")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: reps
      integer(c_int32_t), intent(out) :: tile, zchunk
      real(c_double), intent(out) :: best_ms
    end function
  end interface
end module lbcuda_amd
