! ddns_mod.f90 -- iso_c_binding interface of libddns.so (include/ddns.h) for Fortran
! host programs.  The derived types are bind(C) copies of the C structs, so their
! layout (including alignment padding) is the C compiler's; the interfaces bind the
! C entry points by name.  The reference passes everything by reference and stops on
! errors (turb.f90, partlib.f90); here every call returns a status that the caller
! checks with ddns_check.
module ddns_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: DDNS_OK = 0, DDNS_WARN_GMRES = 1
  integer(c_int), parameter, public :: DDNS_FP64 = 0, DDNS_FP32_MIXED = 1
  integer(c_int), parameter, public :: DDNS_NBIN = 180, DDNS_MAXSET = 10
  integer(c_int), parameter, public :: DDNS_SPEC_KOLMOGOROV = 0, DDNS_SPEC_K4_GAUSS = 1
  integer(c_int), parameter, public :: DDNS_RADII_SETS = 0, DDNS_RADII_LOGNORMAL = 1

  ! parameter.d fields in read order (turb.f90:103-164), then hard-coded constants
  type, bind(C), public :: ddns_params
    integer(c_int) :: nx, ny, nz
    integer(c_int) :: nprocY
    integer(c_int) :: nstep
    real(c_double) :: dt
    real(c_double) :: rnu
    real(c_double) :: time
    integer(c_int) :: new_flow
    integer(c_int) :: idp
    integer(c_int) :: iseed_m0
    integer(c_int) :: forcing
    real(c_double) :: force(2)
    integer(c_int) :: iseedf
    integer(c_int) :: ieout
    integer(c_int) :: ispec
    integer(c_int) :: divstat
    real(c_double) :: rho_water
    real(c_double) :: rho_air
    real(c_double) :: etk_DNS, tauk_DNS
    real(c_double) :: st
    real(c_double) :: ediss_CLOUD
    real(c_double) :: viscosity_CLOUD
    real(c_double) :: gravity
    integer(c_int) :: new_particle
    integer(c_int) :: nonoverlap
    integer(c_int) :: hdi_included
    real(c_double) :: hdi_trunc
    integer(c_int) :: npart
    integer(c_int) :: ndrag
    integer(c_int) :: nset
    integer(c_int) :: pad0
    real(c_double) :: radius(10)
    integer(c_int) :: precision
    integer(c_int) :: nstep_grow
    real(c_double) :: ou_tf
    real(c_double) :: ou_var
    integer(c_int) :: gmres_maxiter
    integer(c_int) :: capacity
    real(c_double) :: gmres_tol
    integer(c_int) :: relocate_seed
    integer(c_int) :: istep0
  end type ddns_params

  type, bind(C), public :: ddns_collision_counts
    integer(c_int) :: icc(3)
    integer(c_int) :: ncoll
    integer(c_int) :: npairs
    integer(c_int) :: nset
    integer(c_int) :: colltype(10, 10)
    integer(c_int) :: nir(10, 10)
    integer(c_int64_t) :: pairs_tested
    integer(c_int64_t) :: pairs_in_dstat
    integer(c_int) :: noverlap_sweeps
    integer(c_int) :: noverlap_relocated
    integer(c_int) :: gmres_iters
    integer(c_int) :: pad1
    real(c_double) :: gmres_residual
  end type ddns_collision_counts

  type, bind(C), public :: ddns_monitors
    real(c_double) :: eta, tmse, xintls, tk, et, xl, uprime
    real(c_double) :: qq, dissp, ree, resol, cfl, vskew, vflat
    real(c_double) :: pmax
    real(c_double) :: vdiv1, vdiv2, vort1, vort2, divvt
  end type ddns_monitors

  type, bind(C), public :: ddns_timing
    real(c_double) :: flow_ms, interp_ms, cells_ms, hdi_ms, advance_ms, colldet_ms, overlap_ms, total_ms
    integer(c_int) :: nsteps, pad
    real(c_double) :: flow_kernel_ms(6)
  end type ddns_timing

  public :: ddns_params_default, ddns_read_parameter_file, ddns_create, ddns_destroy
  public :: ddns_last_error_c, ddns_nek, ddns_set_flow, ddns_get_flow
  public :: ddns_set_forcing_state, ddns_get_forcing_state, ddns_init_flow
  public :: ddns_get_physical_velocity, ddns_get_spectrum, ddns_set_droplets, ddns_get_droplets
  public :: ddns_get_spectra, ddns_get_monitors, ddns_get_spectrum_device, ddns_get_tuning
  public :: ddns_init_droplets, ddns_overlap_fix, ddns_step, ddns_get_collision_counts
  public :: ddns_get_collision_bins, ddns_get_collision_pairs, ddns_get_contact_records, ddns_get_timing
  public :: ddns_fft_r2c, ddns_fft_c2r, ddns_check
  ! batched realisations: R parameter sets advanced together, one member handle each
  public :: ddns_batch_create, ddns_batch_destroy, ddns_batch_size, ddns_batch_member
  public :: ddns_batch_step, ddns_batch_get_timing

  interface
    integer(c_int) function ddns_params_default(p) bind(C, name="ddns_params_default")
      import :: c_int, ddns_params
      type(ddns_params), intent(out) :: p
    end function
    integer(c_int) function ddns_read_parameter_file(path, p) bind(C, name="ddns_read_parameter_file")
      import :: c_int, c_char, ddns_params
      character(kind=c_char), intent(in) :: path(*)
      type(ddns_params), intent(inout) :: p
    end function
    integer(c_int) function ddns_create(p, h) bind(C, name="ddns_create")
      import :: c_int, c_ptr, ddns_params
      type(ddns_params), intent(in) :: p
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function ddns_destroy(h) bind(C, name="ddns_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function ddns_last_error_c(h) bind(C, name="ddns_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function ddns_nek(h) bind(C, name="ddns_nek")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function ddns_set_flow(h, vx, vy, vz, ox, oy, oz) bind(C, name="ddns_set_flow")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: vx(*), vy(*), vz(*), ox(*), oy(*), oz(*)
    end function
    integer(c_int) function ddns_get_flow(h, vx, vy, vz, ox, oy, oz) bind(C, name="ddns_get_flow")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: vx(*), vy(*), vz(*), ox(*), oy(*), oz(*)
    end function
    integer(c_int) function ddns_set_forcing_state(h, b, iseedf, iyf, ivf) bind(C, name="ddns_set_forcing_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: b(*)
      integer(c_int), value :: iseedf, iyf
      integer(c_int), intent(in) :: ivf(*)
    end function
    integer(c_int) function ddns_get_forcing_state(h, b, iseedf, iyf, ivf) bind(C, name="ddns_get_forcing_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: b(*)
      integer(c_int), intent(out) :: iseedf, iyf
      integer(c_int), intent(out) :: ivf(*)
    end function
    integer(c_int) function ddns_init_flow(h, kind, seed, k0, urms) bind(C, name="ddns_init_flow")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: kind, seed
      real(c_double), value :: k0, urms
    end function
    integer(c_int) function ddns_get_physical_velocity(h, ux, uy, uz) bind(C, name="ddns_get_physical_velocity")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: ux(*), uy(*), uz(*)
    end function
    integer(c_int) function ddns_get_spectrum(h, ek) bind(C, name="ddns_get_spectrum")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: ek(*)
    end function
    integer(c_int) function ddns_get_spectrum_device(h, ek_dev, stream) bind(C, name="ddns_get_spectrum_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, ek_dev, stream
    end function
    integer(c_int) function ddns_get_spectra(h, ek, dk, tk) bind(C, name="ddns_get_spectra")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: ek(*), dk(*), tk(*)
    end function
    integer(c_int) function ddns_get_monitors(h, divstat, m) bind(C, name="ddns_get_monitors")
      import :: c_int, c_ptr, ddns_monitors
      type(c_ptr), value :: h
      integer(c_int), value :: divstat
      type(ddns_monitors), intent(out) :: m
    end function
    integer(c_int) function ddns_set_droplets(h, n, yp, fvp, fdp, pertvel) bind(C, name="ddns_set_droplets")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n
      real(c_double), intent(in) :: yp(*), fvp(*), fdp(*), pertvel(*)
    end function
    integer(c_int) function ddns_get_droplets(h, n, yp, fvp, fdp, pertvel, vp, dmove) bind(C, name="ddns_get_droplets")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: n
      real(c_double), intent(out) :: yp(*), fvp(*), fdp(*), pertvel(*), vp(*), dmove(*)
    end function
    integer(c_int) function ddns_init_droplets(h, seed, radii_kind, median_um, sigma_ln, rmin_um, rmax_um) &
        bind(C, name="ddns_init_droplets")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: seed, radii_kind
      real(c_double), value :: median_um, sigma_ln, rmin_um, rmax_um
    end function
    integer(c_int) function ddns_overlap_fix(h) bind(C, name="ddns_overlap_fix")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function ddns_step(h, nsteps) bind(C, name="ddns_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function ddns_get_collision_counts(h, c) bind(C, name="ddns_get_collision_counts")
      import :: c_int, c_ptr, ddns_collision_counts
      type(c_ptr), value :: h
      type(ddns_collision_counts), intent(out) :: c
    end function
    integer(c_int) function ddns_get_collision_bins(h, nirbin, wrbin, wxbin, wybin, wzbin, urbin, thetabin, wrtanbin) &
        bind(C, name="ddns_get_collision_bins")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: nirbin(*)
      real(c_double), intent(out) :: wrbin(*), wxbin(*), wybin(*), wzbin(*), urbin(*), thetabin(*), wrtanbin(*)
    end function
    integer(c_int) function ddns_get_collision_pairs(h, triples, cap, npairs) bind(C, name="ddns_get_collision_pairs")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: triples(*)
      integer(c_int), value :: cap
      integer(c_int), intent(out) :: npairs
    end function
    integer(c_int) function ddns_get_contact_records(h, rec, cap, n) bind(C, name="ddns_get_contact_records")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: rec(*)
      integer(c_int), value :: cap
      integer(c_int), intent(out) :: n
    end function
    integer(c_int) function ddns_get_tuning(h, shapes, ms) bind(C, name="ddns_get_tuning")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: shapes(*)
      real(c_double), intent(out) :: ms(*)
    end function
    integer(c_int) function ddns_get_timing(h, t) bind(C, name="ddns_get_timing")
      import :: c_int, c_ptr, ddns_timing
      type(c_ptr), value :: h
      type(ddns_timing), intent(out) :: t
    end function
    integer(c_int) function ddns_fft_r2c(h, v) bind(C, name="ddns_fft_r2c")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: v(*)
    end function
    integer(c_int) function ddns_fft_c2r(h, v) bind(C, name="ddns_fft_c2r")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: v(*)
    end function
    integer(c_int) function ddns_batch_create(params, nreal, b) bind(C, name="ddns_batch_create")
      import :: c_int, c_ptr, ddns_params
      type(ddns_params), intent(in) :: params(*)
      integer(c_int), value :: nreal
      type(c_ptr), intent(out) :: b
    end function
    integer(c_int) function ddns_batch_destroy(b) bind(C, name="ddns_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function
    integer(c_int) function ddns_batch_size(b) bind(C, name="ddns_batch_size")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function
    type(c_ptr) function ddns_batch_member(b, r) bind(C, name="ddns_batch_member")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: r
    end function
    integer(c_int) function ddns_batch_step(b, nsteps) bind(C, name="ddns_batch_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function ddns_batch_get_timing(b, t) bind(C, name="ddns_batch_get_timing")
      import :: c_int, c_ptr, ddns_timing
      type(c_ptr), value :: b
      type(ddns_timing), intent(out) :: t
    end function
    function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: c_strlen
    end function
  end interface

contains

  ! Stops the program on a fatal status, as the reference does (mpi_finalize; STOP);
  ! prints warnings and carries on.
  subroutine ddns_check(st, h, where)
    integer(c_int), intent(in) :: st
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (st == DDNS_OK) return
    p = ddns_last_error_c(h)
    n = 0
    if (c_associated(p)) n = int(c_strlen(p))
    if (n > 0) then
      call c_f_pointer(p, msg, [n])
      write(*, '(a, a, i4, a, 1000a)') where, ': status ', st, ': ', msg(1:n)
    else
      write(*, '(a, a, i4)') where, ': status ', st
    end if
    if (st < 0) stop 1
  end subroutine ddns_check

end module ddns_mod
