! emg_gpu.F90 -- iso_c_binding interface of libemgpu.so (include/emgpu.h) for
! the Fortran driver of EMSES.
!
! The reference's particle subroutines psolve2(ps, ustep), psolve1(ps, ustep,
! func), charge(ps, func), positn(ps, ustep), spsort(ps) and energy
! (src/pcl2field/psolve2.F90:4, psolve1.F90:4, charge.F90:4, src/pcl/positn.F90:4,
! src/etc/spsort.F90:4, src/etc/energy.F90:4) keep their names and argument
! lists; with -DEMSES_GPU their bodies call the functions below.  Particle
! records are passed as the reference's own type(oh_particle) arrays
! (lib/ohhelp/ohhelp-1.1.1/oh_type.F90:15-28), whose 64-byte layout equals
! emg_particle, through assumed-type dummies.  Every function returns 0 on
! success or a negative EMG_ERR_* code; emg_last_error gives the message.
module emg_gpu
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: EMG_MAX_SPECIES = 8
  integer(c_int), parameter :: EMG_OK = 0
  integer(c_int), parameter :: EMG_ERR_ARG = -1, EMG_ERR_HIP = -2, EMG_ERR_NOMEM = -3
  integer(c_int), parameter :: EMG_ERR_STATE = -4, EMG_ERR_DOMAIN = -5
  integer(c_int), parameter :: EMG_ERR_UNSUPPORTED = -6, EMG_ERR_CAPACITY = -7
  integer(c_int32_t), parameter :: EMG_PLAIN_FP64 = 0, EMG_FP32_FIXED = 1
  integer(c_int32_t), parameter :: EMG_KERNEL_PUSH = 0, EMG_KERNEL_DEPOSIT = 1
  integer(c_int32_t), parameter :: EMG_KERNEL_SORT = 2, EMG_KERNEL_RELOCATE = 3
  integer(c_int32_t), parameter :: EMG_KERNEL_ENERGY = 4

  ! emg_config: grid, species constants (S-units after rescal) and the form
  type, bind(C) :: emg_config
    integer(c_int32_t) :: n(3) = 0
    integer(c_int32_t) :: nspec = 0
    integer(c_int32_t) :: precision = EMG_PLAIN_FP64
    integer(c_int32_t) :: tile = 0
    integer(c_int32_t) :: track_pid = 0
    integer(c_int32_t) :: sort_every = 0
    integer(c_int64_t) :: capacity(EMG_MAX_SPECIES) = 0
    real(c_double) :: q(EMG_MAX_SPECIES) = 0d0
    real(c_double) :: qm(EMG_MAX_SPECIES) = 0d0
    real(c_double) :: rm(EMG_MAX_SPECIES) = 0d0
    real(c_double) :: e0(3) = 0d0
    real(c_double) :: b0(3) = 0d0
  end type emg_config

  ! emg_host_species: inipcl's per-species parameters (EMSES-U units)
  type, bind(C) :: emg_host_species
    integer(c_int64_t) :: npin = 0
    real(c_double) :: peth = 0d0, path = 0d0, spe = 0d0, speth = 0d0, spa = 0d0
    real(c_double) :: vdri = 0d0, vdthz = 0d0, vdthxy = 0d0
    real(c_double) :: vdx = 0d0, vdy = 0d0, vdz = 0d0
  end type emg_host_species

  ! emg_device_species: seeded synthetic particles generated on the device
  type, bind(C) :: emg_device_species
    integer(c_int64_t) :: count = 0
    real(c_double) :: vth_perp = 0d0, vth_para = 0d0
    real(c_double) :: drift(3) = 0d0
    real(c_double) :: t(9) = 0d0
    real(c_double) :: ramp_ratio = 0d0
  end type emg_device_species

  interface
    integer(c_int) function emg_create(cfg, ctx) bind(C, name="emg_create")
      import :: c_int, c_ptr, emg_config
      type(emg_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function emg_create

    integer(c_int) function emg_destroy(ctx) bind(C, name="emg_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function emg_destroy

    integer(c_int) function emg_last_error(buf, len) bind(C, name="emg_last_error")
      import :: c_int, c_char, c_size_t
      character(kind=c_char), dimension(*), intent(out) :: buf
      integer(c_size_t), value :: len
    end function emg_last_error

    integer(c_int) function emg_set_stream(ctx, stream) bind(C, name="emg_set_stream")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream
    end function emg_set_stream

    integer(c_int) function emg_synchronize(ctx) bind(C, name="emg_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function emg_synchronize

    ! pbuf: type(oh_particle) records, species-contiguous (psolve2.F90:102-115)
    integer(c_int) function emg_upload_particles(ctx, pbuf, counts, nspec) &
        bind(C, name="emg_upload_particles")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx
      type(*), dimension(*), intent(in) :: pbuf
      integer(c_int64_t), dimension(*), intent(in) :: counts
      integer(c_int32_t), value :: nspec
    end function emg_upload_particles

    integer(c_int) function emg_download_particles(ctx, pbuf, counts) &
        bind(C, name="emg_download_particles")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      type(*), dimension(*) :: pbuf
      integer(c_int64_t), dimension(*), intent(out) :: counts
    end function emg_download_particles

    integer(c_int) function emg_particle_counts(ctx, counts) bind(C, name="emg_particle_counts")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), dimension(*), intent(out) :: counts
    end function emg_particle_counts

    ! eb(6, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3)), as allocated in ohinit.F90:131-135
    integer(c_int) function emg_set_fields(ctx, eb, lo, hi) bind(C, name="emg_set_fields")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(in) :: eb
      integer(c_int32_t), dimension(3), intent(in) :: lo, hi
    end function emg_set_fields

    integer(c_int) function emg_positn(ctx, ustep) bind(C, name="emg_positn")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ustep
    end function emg_positn

    integer(c_int) function emg_push_es(ctx, ustep) bind(C, name="emg_push_es")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ustep
    end function emg_push_es

    integer(c_int) function emg_deposit_rho(ctx, func) bind(C, name="emg_deposit_rho")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: func
    end function emg_deposit_rho

    ! rho(1, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3)), as allocated in ohinit.F90:155-159
    integer(c_int) function emg_get_rho(ctx, rho, lo, hi) bind(C, name="emg_get_rho")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: rho
      integer(c_int32_t), dimension(3), intent(in) :: lo, hi
    end function emg_get_rho

    integer(c_int) function emg_get_rhodg(ctx, is, rho, lo, hi) bind(C, name="emg_get_rhodg")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: is
      real(c_double), dimension(*), intent(out) :: rho
      integer(c_int32_t), dimension(3), intent(in) :: lo, hi
    end function emg_get_rhodg

    ! engk(3, nspec): 0.5 rm(is) sum v_d^2 / (slx sly slz)
    integer(c_int) function emg_energy(ctx, engk) bind(C, name="emg_energy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: engk
    end function emg_energy

    integer(c_int) function emg_push_em(ctx, ustep) bind(C, name="emg_push_em")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ustep
    end function emg_push_em

    ! aj(3, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3)), as allocated in ohinit.F90:150-154
    integer(c_int) function emg_get_current(ctx, aj, lo, hi) bind(C, name="emg_get_current")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: aj
      integer(c_int32_t), dimension(3), intent(in) :: lo, hi
    end function emg_get_current

    ! ajdg of species is (0-based): q(is) J(is), folded (psolve1.F90:803-817)
    integer(c_int) function emg_get_ajdg(ctx, is, aj, lo, hi) bind(C, name="emg_get_ajdg")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: is
      real(c_double), dimension(*), intent(out) :: aj
      integer(c_int32_t), dimension(3), intent(in) :: lo, hi
    end function emg_get_ajdg

    integer(c_int) function emg_sort(ctx) bind(C, name="emg_sort")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function emg_sort

    integer(c_int) function emg_init_device(ctx, spec, nspec, seed) bind(C, name="emg_init_device")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, emg_device_species
      type(c_ptr), value :: ctx
      type(emg_device_species), dimension(*), intent(in) :: spec
      integer(c_int32_t), value :: nspec
      integer(c_int64_t), value :: seed
    end function emg_init_device

    integer(c_int) function emg_host_inipcl(spec, nspec, n, t, seed, pbuf) &
        bind(C, name="emg_host_inipcl")
      import :: c_int, c_int32_t, c_double, emg_host_species
      type(emg_host_species), dimension(*), intent(in) :: spec
      integer(c_int32_t), value :: nspec
      integer(c_int32_t), dimension(3), intent(in) :: n
      real(c_double), dimension(9), intent(in) :: t
      integer(c_int32_t), value :: seed
      type(*), dimension(*) :: pbuf
    end function emg_host_inipcl

    integer(c_int) function emg_enable_timing(ctx, on) bind(C, name="emg_enable_timing")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function emg_enable_timing

    integer(c_int) function emg_kernel_stats(ctx, which, total_ms, launches) &
        bind(C, name="emg_kernel_stats")
      import :: c_int, c_ptr, c_int32_t, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which
      real(c_double), intent(out) :: total_ms
      integer(c_int64_t), intent(out) :: launches
    end function emg_kernel_stats

    integer(c_int) function emg_reset_kernel_stats(ctx) bind(C, name="emg_reset_kernel_stats")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function emg_reset_kernel_stats
    ! adjoint window of the ES push (EMG_PLAIN_FP64 contexts); no reference counterpart
    integer(c_int) function emg_state_save(ctx, slot) bind(C, name="emg_state_save")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: slot
    end function emg_state_save

    integer(c_int) function emg_state_load(ctx, slot) bind(C, name="emg_state_load")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: slot
    end function emg_state_load

    integer(c_int) function emg_adj_reset(ctx) bind(C, name="emg_adj_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function emg_adj_reset

    integer(c_int) function emg_adj_step(ctx, ustep) bind(C, name="emg_adj_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ustep
    end function emg_adj_step

    integer(c_int) function emg_adj_get_field(ctx, lam_e) bind(C, name="emg_adj_get_field")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: lam_e
    end function emg_adj_get_field

    integer(c_int) function emg_adj_add_density(ctx, g, qfac) bind(C, name="emg_adj_add_density")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(in) :: g, qfac
    end function emg_adj_add_density

    integer(c_int) function emg_adj_velocity_sums(ctx, out) bind(C, name="emg_adj_velocity_sums")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: out
    end function emg_adj_velocity_sums

    ! device-resident EM fields: bfield (bfield.F90:38-52), efield
    ! (efield.F90:55-72), grdcrt on the device copy, read-back, field energy
    integer(c_int) function emg_bfield(ctx) bind(C, name="emg_bfield")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function emg_bfield

    integer(c_int) function emg_efield(ctx, tcs, dlt) bind(C, name="emg_efield")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: tcs, dlt
    end function emg_efield

    integer(c_int) function emg_grdcrt(ctx) bind(C, name="emg_grdcrt")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function emg_grdcrt

    integer(c_int) function emg_get_fields(ctx, eb, lo, hi) bind(C, name="emg_get_fields")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: eb
      integer(c_int32_t), dimension(3), intent(in) :: lo, hi
    end function emg_get_fields

    integer(c_int) function emg_field_energy(ctx, cs, engeb) bind(C, name="emg_field_energy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cs
      real(c_double), dimension(2), intent(out) :: engeb
    end function emg_field_energy

    ! device ES field solve: esfld1 + esfld2 + grdcrt (poissn.F90, esfild.F90)
    integer(c_int) function emg_solve_es(ctx, rhobk) bind(C, name="emg_solve_es")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: rhobk
    end function emg_solve_es

    integer(c_int) function emg_get_potential(ctx, phi) bind(C, name="emg_get_potential")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: phi
    end function emg_get_potential

    ! background charge per periodic node, rhobk(1,0:nx-1,0:ny-1,0:nz-1,3)
    integer(c_int) function emg_set_background(ctx, rhobk) bind(C, name="emg_set_background")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(in) :: rhobk
    end function emg_set_background
  end interface

contains

  ! Message of the last failing call as a Fortran string.
  function emg_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char) :: buf(512)
    integer(c_int) :: rc
    integer :: i
    buf = c_null_char
    rc = emg_last_error(buf, int(size(buf), c_size_t))
    msg = ' '
    do i = 1, size(buf)
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
  end function emg_error_message

  ! The body of charge(ps, func) with the periodic fold, for the reference's
  ! rho(1, -1:n+1, ..., ps) and rhodg(2 nspec + 1, -1:n+1, ..., ps)
  ! (charge.F90:44-50, 59-143; esses.F90:555-562, 594-601): func 0 fills rho,
  ! func 2 rho and rhodg(1:nspec), func 1 only rhodg and leaves rho as the
  ! caller had it -- the EM loop calls charge(1, 1) right after rhoupd(1),
  ! whose continuity-updated rho must survive (emses.F90:474-476).  rhodg is
  ! zeroed first, as the reference zeroes it.
  function emg_charge(ctx, func, nspec, n, rho, rhodg) result(rc)
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: func, nspec, n(3)
    real(c_double), intent(inout) :: rho(-1:n(1) + 1, -1:n(2) + 1, -1:n(3) + 1)
    real(c_double), intent(inout) :: rhodg(:, :, :, :)
    integer(c_int) :: rc
    integer(c_int32_t) :: lo(3), hi(3), is
    real(c_double), allocatable :: tmp(:, :, :)
    lo = -1
    hi = n + 1
    rc = emg_deposit_rho(ctx, func)
    if (rc /= EMG_OK) return
    if (func == 0 .or. func == 2) then
      rc = emg_get_rho(ctx, rho, lo, hi)
      if (rc /= EMG_OK) return
    end if
    if (func == 1 .or. func == 2) then
      rhodg = 0d0
      allocate(tmp(-1:n(1) + 1, -1:n(2) + 1, -1:n(3) + 1))
      do is = 1, nspec
        rc = emg_get_rhodg(ctx, is - 1_c_int32_t, tmp, lo, hi)
        if (rc /= EMG_OK) exit
        rhodg(is, :, :, :) = tmp
      end do
      deallocate(tmp)
    end if
  end function emg_charge

end module emg_gpu
