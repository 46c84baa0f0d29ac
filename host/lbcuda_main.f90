! lbcuda_main.f90 -- Fortran host program of the MI355X LBcuda step.
!
! Keeps the reference host's input semantics: the positional input.dat read
! by readParameters (main.CUF:409-496) and the particle file input.xyz read by
! read_input_atom (main.CUF:1230-1350, Euler angles -> quaternion with
! eul2q(e1,e3,e2), :1284, :1352-1370).  Parameters the reference fixes at
! compile time (box size, sigma, beta, tau, R, tile, precision) come from an
! optional namelist file lbcuda.nml:
!   &lbcuda box=256, taur=1.0, taub=1.0, sigma=0.1, beta=0.999, storage='f32',
!           wetting='janus', radius=5.5, rmass=6.6667e-4, lubrication=.false.,
!           seed=1234, autotune=.false., restart_every=0,
!           diffdens=.false., densr=1.0, densb=1.0 /
! (diffdens selects the density-ratio model of the reference's DIFFDENS build,
! whose densR/densB the reference leaves to be set in SetConstants.)
! Output follows OutputVTK (main.CUF:1812-1929): at step 0 and every nIterVTK
! steps, rhoB ('rho2'), phase, rho ('rho1'), vel and the spheres go to
! output/ in the reference's formats.  restart_every > 0 writes the initType 4
! files (plus the sphere state) to restart_<step>/ every that many steps.
! The time loop replaces the chevron launches of main.CUF:125-308 with
! lbm_step calls of nIterOut steps and prints the reference's progress line
! (main.CUF:282-289; Mlup/s counts (n+2)^3 nodes there) plus the
! interior-node rate.
program lbcuda_host
  use iso_c_binding
  use lbcuda_amd
  use lbcuda_io
  implicit none

  integer :: nIter, nIterOut, nIterVTK, initType, numAtoms, found, itest, i
  integer :: box, seed, restart_every
  real :: vx, vy, vz, ext_f(3), ext_t(3), max_vel
  logical :: withParticles, lrotate, lubrication, lexist
  character(len=120) :: string
  character(len=8) :: storage, wetting
  real :: taur, taub, sigma, beta, radius, rmass, densr, densb
  logical :: diffdens
  type(lbm_config) :: cfg
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  integer(c_int32_t) :: code, estep
  character(kind=c_char) :: msg(256)
  real(c_float), allocatable, target :: rhoR(:,:,:), rhoB(:,:,:), popR(:,:,:,:), popB(:,:,:,:)
  real(c_double), allocatable :: pos(:,:)
  real(c_double) :: sums(5)
  real(c_float) :: u0(3)
  integer(8) :: c0, c1, crate
  real(8) :: secs
  integer :: done, chunk, nacc
  real(8) :: tacc
  logical :: autotune
  integer(c_int32_t) :: tile, zchunk
  real(c_double) :: best_ms
  namelist /lbcuda/ box, taur, taub, sigma, beta, storage, wetting, radius, rmass, lubrication, seed, autotune, &
                     restart_every, diffdens, densr, densb

  ! defaults: dimensions_m.CUF:7 (MYDIMESION 64), :50, :59-60, :103-104
  box = 64; taur = 1.0; taub = 1.0; sigma = 0.1; beta = 0.999
  storage = 'f32'; wetting = 'janus'; radius = 5.5; rmass = 1.0/1500.0
  lubrication = .false.; seed = 1234; autotune = .false.; restart_every = 0
  diffdens = .false.; densr = 1.0; densb = 1.0

  inquire(file='lbcuda.nml', exist=lexist)
  if (lexist) then
    open(unit=12, file='lbcuda.nml', status='old', action='read')
    read(12, nml=lbcuda)
    close(12)
  end if

  ! ---- readParameters, main.CUF:409-496 ----
  nIter = 5; nIterOut = 1; nIterVTK = 100000000; initType = -1
  ext_f = 0.0; ext_t = 0.0; max_vel = 0.0
  withParticles = .false.; lrotate = .false.; numAtoms = 1
  open(unit=11, file='input.dat', status='old', form='formatted', action='read', iostat=itest)
  if (itest /= 0) then
    write(*,*) 'input.dat] Error opening file. STOP'
    stop 1
  end if
  found = 0
  read(11, *, end=120) nIter
  found = 1
  read(11, *, end=120) nIterOut
  found = 2
  read(11, *, end=120) nIterVTK
  found = 3
  read(11, *, end=120) initType
  found = 4
  read(11, *, end=120) vx, vy, vz
  found = 5
  read(11, *, end=120) string, numAtoms
  found = 6
  if (string(1:3) == 'YES') then
    withParticles = .true.
  else if (string(1:3) == 'NO') then
    withParticles = .false.
  else
    found = 5
  end if
  if (withParticles) then
    read(11, *, end=120) ext_f(1), ext_f(2), ext_f(3)
    found = 7
    read(11, *, end=120) string, max_vel
    found = 8
    read(11, *, end=120) string
    found = 9
    if (string(1:3) == 'YES') then
      lrotate = .true.
    else if (string(1:3) == 'NO') then
      lrotate = .false.
    else
      found = 8
    end if
    if (lrotate) then
      read(11, *, end=120) ext_t(1), ext_t(2), ext_t(3)
      found = 10
    end if
  end if
120 continue
  close(11)
  if ((lrotate .and. found /= 10) .or. (.not. lrotate .and. withParticles .and. found /= 9) .or. &
      (.not. withParticles .and. found /= 6)) then
    if (found < 1) write(*,*) 'Number of iterations not found'
    if (found < 2) write(*,*) 'nIterOut not found'
    if (found < 3) write(*,*) 'nIterVTK not found'
    if (found < 4) write(*,*) 'initType not found'
    if (found < 5) write(*,*) 'vx,vy,vz,densR,densB not found'
    if (found < 6) write(*,*) 'withParticles not found'
    if (withParticles .and. found < 7) write(*,*) 'ext_fxx,ext_fyy,ext_fzz not found'
    if (withParticles .and. found < 9) write(*,*) 'lrotate not found'
    if (lrotate .and. found < 10) write(*,*) 'ext_tqx,ext_tqy,ext_tqz not found'
    stop 1
  end if

  ! ---- configuration ----
  call lbm_config_default(cfg, int(box, c_int32_t))
  cfg%tau_r = taur; cfg%tau_b = taub; cfg%sigma = sigma; cfg%beta = beta
  if (diffdens) then
    cfg%diffdens = 1; cfg%dens_r = densr; cfg%dens_b = densb
  end if
  select case (trim(storage))
  case ('f16');  cfg%storage = LBM_STORE_F16
  case ('bf16'); cfg%storage = LBM_STORE_BF16
  case default;  cfg%storage = LBM_STORE_F32
  end select
  select case (trim(wetting))
  case ('neutral'); cfg%wetting = LBM_WET_NEUTRAL
  case ('none');    cfg%wetting = LBM_WET_NONE
  case default;     cfg%wetting = LBM_WET_JANUS
  end select
  if (withParticles) then
    cfg%with_particles = 1
    cfg%n_particles = numAtoms
    cfg%lrotate = merge(1, 0, lrotate)
    cfg%radius = radius
    cfg%rmass = rmass
    cfg%max_vel = max_vel
    cfg%ext_force = ext_f
    cfg%ext_torque = ext_t
    cfg%lubrication = merge(1, 0, lubrication)
  end if

  write(*,'(a,3i6,a,a,a,2f8.4,a,2f6.3)') ' box:', cfg%nx, cfg%ny, cfg%nz, '  storage: ', trim(storage), &
      '  sigma,beta:', cfg%sigma, cfg%beta, '  tau:', cfg%tau_r, cfg%tau_b
  rc = lbm_create(cfg, ctx)
  call check(rc, 'lbm_create')

  ! ---- particles: read_input_atom, main.CUF:1230-1350 ----
  if (withParticles) then
    allocate(pos(13, numAtoms))
    lexist = .false.
    if (initType == 4) lexist = read_particles_restart('particles.init.bin', pos, numAtoms)
    if (lexist) then
      write(*,*) 'particles.init.bin] Using the saved state of', numAtoms, ' particles'
    else
      call read_input_atom()
    end if
    rc = lbm_set_particles(ctx, int(numAtoms, c_int32_t), pos)
    call check(rc, 'lbm_set_particles')
  end if

  ! ---- setupRho_Pops, main.CUF:2531-2608 ----
  u0 = (/ vx, vy, vz /)
  select case (initType)
  case (1, 2, 5)
    rc = lbm_init_fields(ctx, int(initType, c_int32_t), u0, c_null_ptr, c_null_ptr, c_null_ptr)
  case (3)
    call read_rho_files()
    rc = lbm_init_fields(ctx, LBM_INIT_RHO, u0, c_loc(rhoR), c_loc(rhoB), c_null_ptr)
  case (4)
    call read_rho_files()
    rc = lbm_init_fields(ctx, LBM_INIT_POPS, u0, c_loc(rhoR), c_loc(rhoB), c_null_ptr)
    call check(rc, 'lbm_init_fields')
    call read_pops_files()
    rc = lbm_set_pops(ctx, c_loc(popR), c_loc(popB))
  case (6)
    allocate(rhoR(cfg%nx, cfg%ny, cfg%nz), rhoB(cfg%nx, cfg%ny, cfg%nz))
    rc = lbm_fill_random_rho(int(seed, c_int64_t), 0.5d0, 1.0d-4, cfg%nx, cfg%ny, cfg%nz, &
                             c_loc(rhoR), c_loc(rhoB))
    call check(rc, 'lbm_fill_random_rho')
    rc = lbm_init_fields(ctx, LBM_INIT_RANDOM, u0, c_loc(rhoR), c_loc(rhoB), c_null_ptr)
  case default
    write(*,*) 'Wront init type=', initType
    stop 1
  end select
  call check(rc, 'lbm_init_fields')
  if (initType == 4) then
    inquire(file='state.ckpt', exist=lexist)
    if (lexist) call load_checkpoint('state.ckpt')
  end if

  ! tile of the collide-stream kernel: the reference fixes TILE1..3 at compile
  ! time (defines.h:14-24); here the library times its tile table on this box
  if (autotune) then
    rc = lbm_autotune(ctx, 3_c_int32_t, tile, zchunk, best_ms)
    call check(rc, 'lbm_autotune')
    write(*,'(a,i3,a,i4,a,f9.3,a)') ' autotune: tile', tile, ' z-chunk', zchunk, ' (', best_ms, ' ms)'
  end if

  rc = lbm_get_sums(ctx, sums)
  write(*,'(a,2es16.8)') ' initial mass R,B:', sums(1), sums(2)
  write(*,*) 'Run...'

  call system_clock(count_rate=crate)
  call execute_command_line('mkdir -p output')
  if (nIterVTK > 0) call output_vtk(0)      ! main.CUF:114

  ! ---- step loop, main.CUF:125-308 ----
  ! lbm_step runs up to the next progress line, VTK output or restart dump;
  ! the progress line times the steps only (output is not in the interval)
  done = 0; nacc = 0; tacc = 0.0d0
  do while (done < nIter)
    chunk = min(nIterOut - mod(done, nIterOut), nIter - done)
    if (nIterVTK > 0) chunk = min(chunk, nIterVTK - mod(done, nIterVTK))
    if (restart_every > 0) chunk = min(chunk, restart_every - mod(done, restart_every))
    call system_clock(c0)
    rc = lbm_step(ctx, int(chunk, c_int32_t))
    call system_clock(c1)
    done = done + chunk
    if (rc /= 0) then
      rc = lbm_status(ctx, code, estep, msg, 256_c_int32_t)
      write(*,'(a,i4,a,i10,a,256a)') ' lbm_step error ', code, ' at step', estep, ': ', msg(1:cstrlen(msg))
      stop 10
    end if
    nacc = nacc + chunk
    tacc = tacc + real(c1 - c0, 8) / real(crate, 8)
    if (mod(done, nIterOut) == 0 .or. done == nIter) then
      secs = tacc
      write(*,100) done, secs, secs*1.0d3/nacc, &
          real(cfg%nx + 2, 8) * real(cfg%ny + 2, 8) * real(cfg%nz + 2, 8) * nacc / secs / 1.0d6, &
          real(cfg%nx, 8) * real(cfg%ny, 8) * real(cfg%nz, 8) * nacc / secs / 1.0d6
100   format('Iter:',I10,'   Time(s):', G10.4, '   ms/it:',G9.3, '   Mlup/s:',F8.0, '   interior Mlup/s:', F8.0)
      nacc = 0; tacc = 0.0d0
    end if
    if (nIterVTK > 0) then
      if (mod(done, nIterVTK) == 0) call output_vtk(done)   ! main.CUF:303-306
    end if
    if (restart_every > 0) then
      if (mod(done, restart_every) == 0) call write_restart(done)
    end if
  end do

  rc = lbm_get_sums(ctx, sums)
  write(*,'(a,2es16.8)') ' final mass R,B:', sums(1), sums(2)
  if (withParticles) then
    block
      real(c_double), allocatable, target :: p(:,:), f(:,:)
      allocate(p(13, numAtoms), f(12, numAtoms))
      rc = lbm_get_particles(ctx, c_loc(p), c_loc(f))
      write(*,'(a,3f12.5)') ' particle 1 position:', p(1:3, 1)
    end block
  end if
  rc = lbm_destroy(ctx)

contains

  subroutine check(r, what)
    integer(c_int), intent(in) :: r
    character(len=*), intent(in) :: what
    if (r /= 0) then
      write(*,*) trim(what), ' failed with code ', r
      stop 2
    end if
  end subroutine

  integer function cstrlen(s)
    character(kind=c_char), intent(in) :: s(:)
    integer :: k
    cstrlen = 0
    do k = 1, size(s)
      if (s(k) == c_null_char) exit
      cstrlen = k
    end do
  end function

  subroutine eul2q(phis, psis, thetas, qs)    ! main.CUF:1352-1370
    real, intent(in) :: phis, psis, thetas
    real(c_double), intent(out) :: qs(0:3)
    real :: cy, sy, cp, sp, cr, sr
    cy = cos(phis * 0.5); sy = sin(phis * 0.5)
    cp = cos(psis * 0.5); sp = sin(psis * 0.5)
    cr = cos(thetas * 0.5); sr = sin(thetas * 0.5)
    qs(0) = cy * cp * cr + sy * sp * sr
    qs(1) = cy * cp * sr - sy * sp * cr
    qs(2) = sy * cp * sr + cy * sp * cr
    qs(3) = sy * cp * cr - cy * sp * sr
  end subroutine

  subroutine read_input_atom()
    real :: x, y, z, pvx, pvy, pvz, eul(3)
    integer :: nfile, l
    open(unit=13, file='input.xyz', status='old', form='formatted', action='read', iostat=itest)
    if (itest /= 0) then
      write(*,*) 'input.xyz] Error opening file. STOP'
      stop 1
    end if
    read(13, *) nfile
    write(*,*) 'input.xyz] Using:', numAtoms, ' particles from a file with:', nfile
    if (nfile < numAtoms) then
      write(*,*) 'Input file:input.xyz contains less particles than needed. STOP'
      stop 1
    end if
    read(13, '(a120)') string
    pos = 0.0d0
    do l = 1, numAtoms
      if (.not. lrotate) then
        read(13, *) string, x, y, z, pvx, pvy, pvz
        pos(10, l) = 1.0d0
      else
        read(13, *) string, x, y, z, pvx, pvy, pvz, eul(1), eul(2), eul(3)
        call eul2q(eul(1), eul(3), eul(2), pos(10:13, l))
      end if
      pos(1:3, l) = (/ real(x, 8), real(y, 8), real(z, 8) /)
      pos(4:6, l) = (/ real(pvx, 8), real(pvy, 8), real(pvz, 8) /)
    end do
    close(13)
  end subroutine

  ! read_and_setRed_Blue, main.CUF:2444-2479: unformatted, 6-int extent header
  subroutine read_rho_files()
    allocate(rhoR(cfg%nx, cfg%ny, cfg%nz), rhoB(cfg%nx, cfg%ny, cfg%nz))
    call read_init_bin('rhoR.init.bin', 'rhoR', rhoR, cfg%nx, cfg%ny, cfg%nz)
    call read_init_bin('rhoB.init.bin', 'rhoB', rhoB, cfg%nx, cfg%ny, cfg%nz)
  end subroutine

  ! read_and_setPops, main.CUF:2481-2529: popR.<l>.init.bin, popB.<l>.init.bin
  subroutine read_pops_files()
    integer :: l
    allocate(popR(cfg%nx, cfg%ny, cfg%nz, 0:18), popB(cfg%nx, cfg%ny, cfg%nz, 0:18))
    do l = 0, 18
      call read_init_bin('popR.' // fmtnumb(l) // '.init.bin', 'popsR', popR(:,:,:,l), cfg%nx, cfg%ny, cfg%nz)
      call read_init_bin('popB.' // fmtnumb(l) // '.init.bin', 'popsB', popB(:,:,:,l), cfg%nx, cfg%ny, cfg%nz)
    end do
  end subroutine

  ! exact state image of the library (lbm_checkpoint_save), when it has one
  subroutine save_checkpoint(path)
    character(len=*), intent(in) :: path
    integer(c_int64_t) :: nbytes
    integer(c_int8_t), allocatable, target :: img(:)
    integer :: u
    if (lbm_checkpoint_size(ctx, nbytes) /= 0) return
    allocate(img(nbytes))
    rc = lbm_checkpoint_save(ctx, c_loc(img), nbytes)
    call check(rc, 'lbm_checkpoint_save')
    open(newunit=u, file=path, status='replace', access='stream', form='unformatted', action='write')
    write(u) img
    close(u)
  end subroutine

  ! initType 4 with a state.ckpt beside the restart files: continue exactly
  subroutine load_checkpoint(path)
    character(len=*), intent(in) :: path
    integer(c_int64_t) :: nbytes
    integer(c_int8_t), allocatable, target :: img(:)
    integer :: u
    inquire(file=path, size=nbytes)
    allocate(img(nbytes))
    open(newunit=u, file=path, status='old', access='stream', form='unformatted', action='read')
    read(u) img
    close(u)
    rc = lbm_checkpoint_load(ctx, c_loc(img), nbytes)
    call check(rc, 'lbm_checkpoint_load')
    write(*,*) 'state.ckpt] Continuing from the saved state'
  end subroutine

  ! OutputVTK, main.CUF:1812-1929: colour densities as the device keeps them
  ! (rhoR_d, rhoB_d), phase and velocity from the populations (moments2Fl)
  subroutine output_vtk(step)
    integer, intent(in) :: step
    real(c_float), allocatable, target :: rr(:,:,:), rb(:,:,:), ph(:,:,:), vel(:,:,:,:)
    real(c_double), allocatable, target :: p(:,:)
    integer(8) :: t0, t1
    call system_clock(t0)
    allocate(rr(cfg%nx, cfg%ny, cfg%nz), rb(cfg%nx, cfg%ny, cfg%nz), ph(cfg%nx, cfg%ny, cfg%nz), &
             vel(3, cfg%nx, cfg%ny, cfg%nz))
    rc = lbm_get_moments(ctx, c_null_ptr, c_null_ptr, c_loc(vel), c_loc(ph))
    call check(rc, 'lbm_get_moments')
    rc = lbm_get_rho(ctx, c_loc(rr), c_loc(rb))
    call check(rc, 'lbm_get_rho')
    call write_vti('rhoB', step, 'rho2', 1, rb, cfg%nx, cfg%ny, cfg%nz)
    call write_vti('phase', step, 'phase', 1, ph, cfg%nx, cfg%ny, cfg%nz)
    call write_vti('rho', step, 'rho1', 1, rr, cfg%nx, cfg%ny, cfg%nz)
    call write_vti('vel', step, 'vel', 3, vel, cfg%nx, cfg%ny, cfg%nz)
    if (withParticles) then
      allocate(p(13, numAtoms))
      rc = lbm_get_particles(ctx, c_loc(p), c_null_ptr)
      call check(rc, 'lbm_get_particles')
      call write_particle_vtk('particle', step, p, numAtoms, lrotate)
    end if
    call system_clock(t1)
    write(*,'(a,g10.2,a)') 'Time for VTK:', real(t1 - t0, 8) / real(crate, 8), ' secs'
  end subroutine

  ! initType 4 input files of the current state, in restart_<step>/
  subroutine write_restart(step)
    integer, intent(in) :: step
    character(len=:), allocatable :: dir
    real(c_float), allocatable, target :: a(:,:,:), b(:,:,:), q(:,:,:,:)
    real(c_double), allocatable, target :: p(:,:)
    integer :: l, color
    dir = 'restart_' // fmtnumb(step)
    call execute_command_line('mkdir -p ' // dir)
    allocate(a(cfg%nx, cfg%ny, cfg%nz), b(cfg%nx, cfg%ny, cfg%nz), q(cfg%nx, cfg%ny, cfg%nz, 0:18))
    rc = lbm_get_rho(ctx, c_loc(a), c_loc(b))
    call check(rc, 'lbm_get_rho')
    call write_init_bin(dir // '/rhoR.init.bin', a, cfg%nx, cfg%ny, cfg%nz)
    call write_init_bin(dir // '/rhoB.init.bin', b, cfg%nx, cfg%ny, cfg%nz)
    do color = 0, 1
      rc = lbm_get_pops(ctx, int(color, c_int32_t), c_loc(q))
      call check(rc, 'lbm_get_pops')
      do l = 0, 18
        call write_init_bin(dir // '/pop' // merge('R', 'B', color == 0) // '.' // fmtnumb(l) // '.init.bin', &
                            q(:,:,:,l), cfg%nx, cfg%ny, cfg%nz)
      end do
    end do
    if (withParticles) then
      allocate(p(13, numAtoms))
      rc = lbm_get_particles(ctx, c_loc(p), c_null_ptr)
      call check(rc, 'lbm_get_particles')
      call write_particles_restart(dir, p, numAtoms)
    end if
    call save_checkpoint(dir // '/state.ckpt')
    write(*,'(a,i10,a)') ' restart files of step', step, ' written to ' // dir
  end subroutine

end program lbcuda_host
