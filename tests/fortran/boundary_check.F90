! Host-side check of the Fortran binding (no device calls): an invalid
! configuration is rejected with EMG_ERR_ARG and a message, and the
! reference-seeded initial particles come back in type(oh_particle) records
! (lib/ohhelp/ohhelp-1.1.1/oh_type.F90:15-28), written to particles.bin.
program boundary_check
  use iso_c_binding
  use emg_gpu
  implicit none
  type :: oh_particle
    sequence
    real(8) :: x, y, z, vx, vy, vz
    integer(4) :: pid, preside, nid, spec
  end type oh_particle
  type(emg_config) :: cfg
  type(c_ptr) :: ctx
  type(emg_host_species) :: sp(2)
  type(oh_particle), allocatable :: pbuf(:)
  integer(c_int32_t) :: n(3) = [8, 6, 4]
  real(c_double) :: t(9) = [1d0, 0d0, 0d0, 0d0, 1d0, 0d0, 0d0, 0d0, 1d0]
  integer :: rc, u

  cfg%nspec = 0
  rc = emg_create(cfg, ctx)
  if (rc /= EMG_ERR_ARG) stop 1
  print '(a)', trim(emg_error_message())

  sp(1)%npin = 300
  sp(1)%peth = 1d0
  sp(1)%path = 1d0
  sp(2)%npin = 201
  sp(2)%peth = 0.1d0
  sp(2)%path = 0.2d0
  sp(2)%vdx = 0.5d0
  allocate(pbuf(501))
  rc = emg_host_inipcl(sp, 2_c_int32_t, n, t, 10_c_int32_t, pbuf)
  if (rc /= EMG_OK) stop 2
  open(newunit=u, file='particles.bin', access='stream', form='unformatted', status='replace')
  write(u) pbuf
  close(u)
end program boundary_check
