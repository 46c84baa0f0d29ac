! lbcuda_io.f90 -- output and restart files of the Fortran host program.
!
! Field output keeps the reference's file formats so its post-processing
! (ParaView, LBsoft-style readers) opens them unchanged:
!   * XML ImageData (.vti), one Float32 array in raw appended data: '_',
!     a 4-byte byte count, then the values x fastest (writeImageDataVTI,
!     writeImageDataVTI_3d, write_output.f90:115-391, :393-560, serial branch);
!   * legacy ASCII POLYDATA (.vtk) for the spheres: positions, index, velocity
!     and, with rotation, the body axes q e_i q* (writeParticleVTK,
!     write_output.f90:674-746, take_rotversor* :785-837).
! File names follow the reference: output/<name>_<rank>.<step>.vti with
! six-digit zero-padded numbers (write_fmtnumb, write_output.f90:858-875) and
! output/particle_<step>.vtk.
!
! Restart files are the ones initType 4 reads (read_and_setRed_Blue,
! read_and_setPops, main.CUF:2444-2529): unformatted sequential records, a
! 6-integer extent header (1,nx,1,ny,1,nz) then the (nx,ny,nz) array, for
! rhoR.init.bin, rhoB.init.bin and popR/popB.<l>.init.bin, l = 0..18.  The
! reference has no writer for them and no particle restart; here the sphere
! state is written twice: input.xyz in the format read_input_atom reads
! (main.CUF:1230-1350; angular velocity is not part of that format) and
! particles.init.bin with the full FP64 (13,n) state.
module lbcuda_io
  use iso_c_binding
  implicit none
  private
  public :: fmtnumb, write_vti, write_particle_vtk, write_init_bin, read_init_bin, &
            write_particles_restart, read_particles_restart, body_axis, q2eul

  character(len=1), parameter :: nl = char(10)

contains

  ! at least six digits, zero padded
  function fmtnumb(n) result(s)
    integer, intent(in) :: n
    character(len=:), allocatable :: s
    character(len=24) :: b
    write(b, '(i0.6)') n
    s = trim(b)
  end function

  ! One ImageData file holding ncomp Float32 components per node.
  subroutine write_vti(fname, step, varname, ncomp, a, nx, ny, nz)
    character(len=*), intent(in) :: fname, varname
    integer, intent(in) :: step, ncomp, nx, ny, nz
    real(c_float), intent(in) :: a(*)
    character(len=:), allocatable :: path, extent, comps
    character(len=12) :: nc
    integer :: u
    integer(c_int32_t) :: nbytes
    integer(8) :: nval

    path = 'output/' // trim(fname) // '_' // fmtnumb(0) // '.' // fmtnumb(step) // '.vti'
    extent = fmtnumb(1) // ' ' // fmtnumb(nx) // ' ' // fmtnumb(1) // ' ' // fmtnumb(ny) // ' ' // &
             fmtnumb(1) // ' ' // fmtnumb(nz)
    comps = ''
    if (ncomp > 1) then
      write(nc, '(i0)') ncomp
      comps = ' NumberOfComponents="' // trim(nc) // '"'
    end if
    nval = int(ncomp, 8) * nx * ny * nz
    nbytes = int(4 * nval, c_int32_t)

    open(newunit=u, file=path, status='replace', access='stream', form='unformatted', action='write')
    write(u) ' <VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" >' // nl, &
             '  <ImageData WholeExtent="' // extent // '" >' // nl, &
             '  <Piece Extent="' // extent // '">' // nl, &
             '    <PointData>' // nl, &
             '     <DataArray type="Float32" Name="' // trim(varname) // '"' // comps // &
             ' format="appended" offset="0" />' // nl, &
             '    </PointData>' // nl, '  </Piece>' // nl, '  </ImageData>' // nl, &
             '  <AppendedData encoding="raw">' // nl, '_'
    write(u) nbytes, a(1:nval)
    write(u) nl // '  </AppendedData>' // nl // ' </VTKFile >' // nl
    close(u)
  end subroutine

  ! Hamilton product
  pure function qmul(a, b) result(c)
    real(8), intent(in) :: a(0:3), b(0:3)
    real(8) :: c(0:3)
    c(0) = a(0) * b(0) - a(1) * b(1) - a(2) * b(2) - a(3) * b(3)
    c(1) = a(0) * b(1) + b(0) * a(1) + a(2) * b(3) - a(3) * b(2)
    c(2) = a(0) * b(2) + b(0) * a(2) + a(3) * b(1) - a(1) * b(3)
    c(3) = a(0) * b(3) + b(0) * a(3) + a(1) * b(2) - a(2) * b(1)
  end function

  ! body axis i (1..3) of orientation q: the vector part of q e_i q*
  pure function body_axis(q, i) result(v)
    real(8), intent(in) :: q(0:3)
    integer, intent(in) :: i
    real(8) :: v(3), e(0:3), qc(0:3), r(0:3)
    e = 0.0d0
    e(i) = 1.0d0
    qc(0) = q(0)
    qc(1:3) = -q(1:3)
    r = qmul(q, qmul(e, qc))
    v = r(1:3)
  end function

  ! Inverse of the host's eul2q (main.CUF:1352-1370): eul2q(phi, psi, theta)
  ! builds the z-y-x rotation with yaw phi, pitch psi and roll theta, and
  ! input.xyz lists the angles as (phi, theta, psi) (main.CUF:1284).
  pure function q2eul(q) result(eul)
    real(8), intent(in) :: q(0:3)
    real(8) :: eul(3), w, x, y, z, s
    w = q(0); x = q(1); y = q(2); z = q(3)
    s = max(-1.0d0, min(1.0d0, 2.0d0 * (w * y - z * x)))
    eul(1) = atan2(2.0d0 * (w * z + x * y), 1.0d0 - 2.0d0 * (y * y + z * z))
    eul(2) = atan2(2.0d0 * (w * x + y * z), 1.0d0 - 2.0d0 * (x * x + y * y))
    eul(3) = asin(s)
  end function

  ! Spheres as legacy ASCII polydata (pos is the (13,n) state of lbm_get_particles).
  subroutine write_particle_vtk(fname, step, pos, n, lrotate)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: step, n
    real(c_double), intent(in) :: pos(13, n)
    logical, intent(in) :: lrotate
    integer :: u, i, ax
    character(len=5), parameter :: axname(3) = (/ 'Vectx', 'Vecty', 'Vectz' /)

    open(newunit=u, file='output/' // trim(fname) // '_' // fmtnumb(step) // '.vtk', &
         status='replace', action='write', form='formatted')
    write(u, '(a)') '# vtk DataFile Version 2.0', 'lbcuda_amd spheres', 'ASCII', 'DATASET POLYDATA'
    write(u, '(a,i12,a)') 'POINTS ', n, ' float'
    do i = 1, n
      write(u, '(3es16.5)') real(pos(1:3, i))
    end do
    write(u, '(a)') ''
    write(u, '(a,i12)') 'POINT_DATA ', n
    write(u, '(a)') 'SCALARS Type float 1', 'LOOKUP_TABLE default'
    do i = 1, n
      write(u, '(i12)') i
    end do
    write(u, '(a)') '', 'VECTORS vel float'
    do i = 1, n
      write(u, '(3es16.7)') real(pos(4:6, i))
    end do
    if (lrotate) then
      do ax = 1, 3
        write(u, '(a)') '', 'VECTORS ' // axname(ax) // ' float'
        do i = 1, n
          write(u, '(3f20.8)') real(body_axis(pos(10:13, i), ax))
        end do
      end do
    end if
    close(u)
  end subroutine

  subroutine write_init_bin(path, a, nx, ny, nz)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nx, ny, nz
    real(c_float), intent(in) :: a(nx, ny, nz)
    integer :: u
    open(newunit=u, file=path, form='unformatted', action='write', status='replace')
    write(u) 1, nx, 1, ny, 1, nz
    write(u) a
    close(u)
  end subroutine

  ! Reads one initType 3/4 file; stops on a wrong extent as the reference does.
  subroutine read_init_bin(path, what, a, nx, ny, nz)
    character(len=*), intent(in) :: path, what
    integer, intent(in) :: nx, ny, nz
    real(c_float), intent(out) :: a(nx, ny, nz)
    integer :: u, minx, maxx, miny, maxy, minz, maxz
    open(newunit=u, file=path, form='unformatted', action='read', status='old')
    read(u) minx, maxx, miny, maxy, minz, maxz
    if (minx /= 1 .or. miny /= 1 .or. minz /= 1 .or. maxx /= nx .or. maxy /= ny .or. maxz /= nz) then
      write(*, '(a,6i8)') 'Wrong start ' // what // ' file', minx, maxx, miny, maxy, minz, maxz
      stop 1
    end if
    read(u) a
    close(u)
  end subroutine

  subroutine write_particles_restart(dir, pos, n)
    character(len=*), intent(in) :: dir
    integer, intent(in) :: n
    real(c_double), intent(in) :: pos(13, n)
    integer :: u, i
    real(8) :: eul(3)
    open(newunit=u, file=dir // '/particles.init.bin', form='unformatted', action='write', status='replace')
    write(u) n
    write(u) pos
    close(u)
    open(newunit=u, file=dir // '/input.xyz', status='replace', action='write', form='formatted')
    write(u, '(i8)') n
    write(u, '(a)') ' restart: x y z vx vy vz e1 e2 e3 (orientation eul2q(e1, e3, e2))'
    do i = 1, n
      eul = q2eul(pos(10:13, i))
      write(u, '(a,9es24.15)') 'C ', pos(1:6, i), eul
    end do
    close(u)
  end subroutine

  ! Full sphere state of a restart; false if the file is absent.
  logical function read_particles_restart(path, pos, n) result(found)
    character(len=*), intent(in) :: path
    integer, intent(in) :: n
    real(c_double), intent(out) :: pos(13, n)
    integer :: u, nfile
    inquire(file=path, exist=found)
    if (.not. found) return
    open(newunit=u, file=path, form='unformatted', action='read', status='old')
    read(u) nfile
    if (nfile /= n) then
      write(*, '(a,2i8)') 'particles.init.bin holds a different sphere count', nfile, n
      stop 1
    end if
    read(u) pos
    close(u)
  end function

end module lbcuda_io
