! turb_gpu.f90 -- Fortran host driver of the one-GPU DNS + droplet step.
!
! Mirrors the reference driver (turb.f90): parameter.d is read with the same
! list-directed statements in the same order (turb.f90:101-166), HDI forces the
! non-overlap option (turb.f90:171), the flow and the droplets are initialised
! (turb.f90:418-454: PART_INIT + PERIODIC + OVERLAP_DET, initial_flow or rest for the
! stochastic forcing) and the time loop advances one step per call (turb.f90:492-700).
! The whole step runs inside libddns.so; this program only moves parameters and
! diagnostics across the C ABI.
!
! Usage: turb_gpu [parameter-file] [fp64|fp32]
program turb_gpu
  use, intrinsic :: iso_c_binding
  use ddns_mod
  implicit none

  type(ddns_params) :: p
  type(ddns_collision_counts) :: cc
  type(ddns_timing) :: tm
  type(c_ptr) :: h
  character(len=256) :: pfile, arg, str
  integer :: nx, ny, nz, nprocY, nstep, idp, iseed_m0, forcing, iseedf, ieout, ispec
  integer :: npart, ndrag, nset, il, istep, nek
  logical :: new_flow, divstat, new_particle, nonoverlap, hdi_included
  double precision :: dt, rnu, time, force(2), rho_water, rho_air, etk_DNS, tauk_DNS, st
  double precision :: ediss_CLOUD, viscosity_CLOUD, gravity, hdi_trunc, radius(10)
  double precision :: t_total
  double precision, allocatable :: ek(:)
  integer(c_int) :: stat, prec
  integer(8) :: c0, c1, crate

  pfile = 'parameter.d'
  prec = DDNS_FP64
  if (command_argument_count() >= 1) call get_command_argument(1, pfile)
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    if (trim(arg) == 'fp32') prec = DDNS_FP32_MIXED
  end if

  ! ---- parameter.d, turb.f90:102-166
  radius = 0.d0
  open(90, file=trim(pfile), status='old')
  read(90,*) nx, ny, nz
  read(90,*) str
  read(90,*) nprocY
  read(90,*) str
  read(90,*) nstep
  read(90,*) str
  read(90,*) dt
  read(90,*) str
  read(90,*) rnu
  read(90,*) str
  read(90,*) time
  read(90,*) str
  read(90,*) new_flow
  read(90,*) str
  read(90,*) idp
  read(90,*) str
  read(90,*) iseed_m0
  read(90,*) str
  read(90,*) forcing
  read(90,*) str
  read(90,*) force
  read(90,*) str
  read(90,*) iseedf
  read(90,*) str
  read(90,*) ieout
  read(90,*) str
  read(90,*) ispec
  read(90,*) str
  read(90,*) divstat
  read(90,*) str
  read(90,*) rho_water
  read(90,*) str
  read(90,*) rho_air
  read(90,*) str
  read(90,*) etk_DNS, tauk_DNS
  read(90,*) str
  read(90,*) st
  read(90,*) str
  read(90,*) ediss_CLOUD
  read(90,*) str
  read(90,*) viscosity_CLOUD
  read(90,*) str
  read(90,*) gravity
  read(90,*) str
  read(90,*) new_particle
  read(90,*) str
  read(90,*) nonoverlap
  read(90,*) str
  read(90,*) hdi_included
  read(90,*) str
  read(90,*) hdi_trunc
  read(90,*) str
  read(90,*) npart
  read(90,*) str
  read(90,*) ndrag
  read(90,*) str
  read(90,*) nset
  read(90,*) str
  do il = 1, nset
    read(90,*) radius(il)
  end do
  close(90)
  if (hdi_included) nonoverlap = .true.                      ! turb.f90:171

  stat = ddns_params_default(p)
  p%nx = nx; p%ny = ny; p%nz = nz; p%nprocY = nprocY; p%nstep = nstep
  p%dt = dt; p%rnu = rnu; p%time = time
  p%new_flow = merge(1, 0, new_flow); p%idp = idp; p%iseed_m0 = iseed_m0
  p%forcing = forcing; p%force = force; p%iseedf = iseedf
  p%ieout = ieout; p%ispec = ispec; p%divstat = merge(1, 0, divstat)
  p%rho_water = rho_water; p%rho_air = rho_air; p%etk_DNS = etk_DNS; p%tauk_DNS = tauk_DNS
  p%st = st; p%ediss_CLOUD = ediss_CLOUD; p%viscosity_CLOUD = viscosity_CLOUD; p%gravity = gravity
  p%new_particle = merge(1, 0, new_particle); p%nonoverlap = merge(1, 0, nonoverlap)
  p%hdi_included = merge(1, 0, hdi_included); p%hdi_trunc = hdi_trunc
  p%npart = npart; p%ndrag = ndrag; p%nset = nset; p%radius = radius
  p%precision = prec

  h = c_null_ptr
  stat = ddns_create(p, h)
  call ddns_check(stat, h, 'ddns_create')
  nek = ddns_nek(h)
  allocate(ek(nek))

  ! ---- initial flow, turb.f90:438-454
  if (.not. new_flow) then
    write(*,*) 'new_flow = .FALSE.: restart files are written by the caller through ddns_set_flow'
    stop 1
  end if
  if (forcing == 1) then
    stat = ddns_init_flow(h, DDNS_SPEC_KOLMOGOROV, iseed_m0, 4.d0, 1.d0)
  else
    stat = ddns_init_flow(h, -1, 0, 0.d0, 0.d0)                ! start from rest
  end if
  call ddns_check(stat, h, 'ddns_init_flow')

  ! ---- droplets: PART_INIT (seed 290710, partlib.f90:96), PERIODIC, OVERLAP_DET
  if (npart > 0) then
    if (.not. new_particle) then
      write(*,*) 'new_particle = .FALSE.: droplet files are loaded by the caller through ddns_set_droplets'
      stop 1
    end if
    stat = ddns_init_droplets(h, 290710_c_int, DDNS_RADII_SETS, 0.d0, 0.d0, 0.d0, 0.d0)
    call ddns_check(stat, h, 'ddns_init_droplets')
  end if

  open(40, file='collision.dat', status='unknown')
  open(41, file='spectrum.dat', status='unknown')
  t_total = 0.d0
  call system_clock(c0, crate)
  do istep = 1, nstep
    stat = ddns_step(h, 1_c_int)
    call ddns_check(stat, h, 'ddns_step')
    stat = ddns_get_timing(h, tm)
    t_total = t_total + tm%total_ms
    time = time + dt
    if (npart > 0) then
      stat = ddns_get_collision_counts(h, cc)
      write(40, '(i8, 1pe16.8, 5i10)') istep, time, cc%icc, cc%ncoll, cc%npairs
    end if
    if (ieout > 0 .and. mod(istep, ieout) == 0) then
      stat = ddns_get_spectrum(h, ek)
      write(*, '(a, i8, a, 1pe14.6, a, 1pe14.6, a, 0pf9.3, a)') ' istep', istep, '  time', time, &
          '  energy', sum(ek), '  step', tm%total_ms, ' ms'
    end if
    if (ispec > 0 .and. mod(istep, ispec) == 0) then
      stat = ddns_get_spectrum(h, ek)
      do il = 1, nek
        write(41, '(i8, i6, 1pe20.10)') istep, il, ek(il)
      end do
    end if
  end do
  call system_clock(c1)
  close(40)
  close(41)
  write(*,*) 'total time              =', dble(c1 - c0) / dble(crate)
  write(*,*) 'GPU time in ddns_step   =', t_total / 1000.d0
  write(*,*) 'Next run starts at time =', time
  stat = ddns_destroy(h)
  write(*,*) 'Successfully completed'
end program turb_gpu
