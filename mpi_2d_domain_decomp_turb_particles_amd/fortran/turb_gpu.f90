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
! Restarts use the reference's files (load_save.f90): vel2D<idp>.<rank> (u, omega-hat
! history, six unformatted records of the rank's spectral block), force.<idp> (OU state)
! and part.<rank> (droplets).  A run written by P ranks (nprocY from parameter.d) is
! assembled block by block; this driver writes one-rank files, index idp+1.
!
! Usage: turb_gpu [parameter-file] [fp64|fp32] [ranks-of-the-restart-files]
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
  double precision, allocatable :: ek(:), dk(:), tk(:)
  type(ddns_monitors) :: mon
  double precision, allocatable :: vel(:,:,:,:), blk(:,:,:), yp(:,:), fvp(:,:,:), fdp(:,:,:), pert(:,:)
  double precision, allocatable :: vpd(:,:), dmv(:,:)
  double precision :: bforce(450)
  integer :: ivf(32), iyf, isf, nin, nprocY_in, nps_r, nps_tot, ir, c
  ! COLLISION_DET's statistics files (partlib.f90:1641-1797)
  integer, allocatable :: nirb(:,:,:), nirb0(:,:,:), ubin(:,:,:), uvp(:,:)
  double precision, allocatable :: wb(:,:,:,:), wb0(:,:,:,:), crec(:,:)
  integer :: i1, i2, k, itot
  integer(c_int) :: ncrec
  character(len=32) :: fname, fmt125, fmt126
  character(len=8), parameter :: binname(8) = [character(len=8) :: 'nirbin', 'wrbin', 'wxbin', 'wybin', &
                                                'wzbin', 'urbin', 'thetabin', 'wrtanbin']
  integer(c_int) :: stat, prec, np_c
  integer(8) :: c0, c1, crate

  pfile = 'parameter.d'
  prec = DDNS_FP64
  if (command_argument_count() >= 1) call get_command_argument(1, pfile)
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    if (trim(arg) == 'fp32') prec = DDNS_FP32_MIXED
  end if
  nin = 1
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) nin
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
  allocate(ek(nek), dk(nek), tk(nek))

  ! ---- initial flow, turb.f90:438-454
  nprocY_in = 1
  if (nin > 1) nprocY_in = nprocY
  if (new_flow) then
    if (forcing == 1) then
      stat = ddns_init_flow(h, DDNS_SPEC_KOLMOGOROV, iseed_m0, 4.d0, 1.d0)
    else
      stat = ddns_init_flow(h, -1, 0, 0.d0, 0.d0)              ! start from rest
    end if
    call ddns_check(stat, h, 'ddns_init_flow')
  else
    allocate(vel(nx + 2, nx + 2, nx, 6))
    call read_velocity(idp, nin, nprocY_in, nx, vel)
    stat = ddns_set_flow(h, vel(:,:,:,1), vel(:,:,:,2), vel(:,:,:,3), vel(:,:,:,4), vel(:,:,:,5), vel(:,:,:,6))
    call ddns_check(stat, h, 'ddns_set_flow')
    if (forcing == 2) then
      call read_forcing(idp, isf, ivf, iyf, bforce)
      stat = ddns_set_forcing_state(h, bforce, isf, iyf, ivf)
      call ddns_check(stat, h, 'ddns_set_forcing_state')
    end if
    deallocate(vel)
  end if

  ! ---- droplets: PART_INIT (seed 290710, partlib.f90:96), PERIODIC, OVERLAP_DET, or
  !      LOAD_PARTICLE (load_save.f90:99-124) from every rank's file
  if (npart > 0) then
    if (new_particle) then
      stat = ddns_init_droplets(h, 290710_c_int, DDNS_RADII_SETS, 0.d0, 0.d0, 0.d0, 0.d0)
      call ddns_check(stat, h, 'ddns_init_droplets')
    else
      nps_tot = 0
      do ir = 0, nin - 1
        open(12, file=rank_name('part.', ir), status='old', form='unformatted')
        read(12) nps_r
        close(12)
        nps_tot = nps_tot + nps_r
      end do
      allocate(yp(8, nps_tot), fvp(3, 0:2, nps_tot), fdp(3, 0:3, nps_tot), pert(3, nps_tot))
      pert = 0.d0
      nps_tot = 0
      do ir = 0, nin - 1
        open(12, file=rank_name('part.', ir), status='old', form='unformatted')
        read(12) nps_r
        read(12) yp(:, nps_tot + 1:nps_tot + nps_r), fvp(:, :, nps_tot + 1:nps_tot + nps_r), &
                 fdp(:, :, nps_tot + 1:nps_tot + nps_r)
        close(12)
        nps_tot = nps_tot + nps_r
      end do
      stat = ddns_set_droplets(h, int(nps_tot, c_int), yp, fvp, fdp, pert)
      call ddns_check(stat, h, 'ddns_set_droplets')
      deallocate(yp, fvp, fdp, pert)
    end if
  end if

  open(40, file='collision.dat', status='unknown')
  if (npart > 0) then
    ! turb.f90:351-395: the pair statistics of every set pair i1 <= i2
    open(51, file='detectpairs.dat', status='unknown')
    allocate(nirb(nset, nset, DDNS_NBIN), nirb0(nset, nset, DDNS_NBIN), ubin(8, nset, nset), uvp(nset, nset))
    allocate(wb(nset, nset, DDNS_NBIN, 7), wb0(nset, nset, DDNS_NBIN, 7), crec(12, 1024))
    nirb = 0
    wb = 0.d0
    do i1 = 1, nset
      do i2 = i1, nset
        do k = 1, 8
          write(fname, '(a, i0, i0, a)') trim(binname(k)), i1, i2, '.dat'
          open(newunit=ubin(k, i1, i2), file=trim(fname), status='unknown')
        end do
        write(fname, '(a, i0, i0, a)') 'Vp', i1, i2, '.dat'
        open(newunit=uvp(i1, i2), file=trim(fname), position='append', status='unknown')
      end do
    end do
    write(fmt125, '(a, i0, a)') '(', DDNS_NBIN, '(i6))'
    write(fmt126, '(a, i0, a)') '(', DDNS_NBIN, '(1pe12.4))'
  end if
  ! the reference's statistics files (turb.f90:356-360), written from the state at the
  ! start of step jstep as POSTPROC / TRANSPOSTPROC see it (turb.f90:628-656)
  open(41, file='spectrum.dat', status='unknown')
  open(42, file='transfer.dat', status='unknown')
  open(43, file='monitora.dat', status='unknown')
  open(44, file='monitorb.dat', status='unknown')
  if (divstat) open(45, file='monitorc.dat', status='unknown')
  t_total = 0.d0
  call system_clock(c0, crate)
  do istep = 1, nstep
    if (ispec > 0 .and. mod(istep, ispec) == 0) then
      stat = ddns_get_spectra(h, ek, dk, tk)
      call ddns_check(stat, h, 'ddns_get_spectra')
    end if
    if (ieout > 0 .and. mod(istep, ieout) == 0) then                  ! POSTPROC, postprocessing.f90:76-272
      stat = ddns_get_monitors(h, merge(1_c_int, 0_c_int, divstat), mon)
      call ddns_check(stat, h, 'ddns_get_monitors')
      if (ispec > 0 .and. mod(istep, ispec) == 0) then                ! 157-168
        write(41, *) istep / ispec - 2
        do il = 1, nek
          write(41, *) il, ek(il), dk(il)
        end do
      end if
      write(43, 601) time, mon%eta, mon%tmse, mon%xintls, mon%tk, mon%et, mon%xl, mon%uprime
      write(44, 601) time, mon%qq, mon%dissp, mon%ree, mon%resol, mon%cfl, mon%vskew, mon%vflat
      if (divstat) write(45, 602) time, mon%vdiv1, mon%vdiv2, mon%vort1, mon%vort2, mon%divvt
    end if
    if (ispec > 0 .and. mod(istep, ispec) == 0) then                  ! TRANSPOSTPROC, 294-313
      write(42, *) istep / ispec - 1
      do il = 1, nek
        write(42, *) il, tk(il)
      end do
    end if
    stat = ddns_step(h, 1_c_int)
    call ddns_check(stat, h, 'ddns_step')
    stat = ddns_get_timing(h, tm)
    t_total = t_total + tm%total_ms
    if (npart > 0) call collision_files()
    time = time + dt                                                  ! turb.f90:693
    if (ieout > 0 .and. mod(istep, ieout) == 0) then
      stat = ddns_get_spectrum(h, ek)
      write(*, '(a, i8, a, 1pe14.6, a, 1pe14.6, a, 0pf9.3, a)') ' istep', istep, '  time', time, &
          '  energy', sum(ek), '  step', tm%total_ms, ' ms'
    end if
  end do
  call system_clock(c1)
  close(40)
  if (npart > 0) then
    close(51)
    do i1 = 1, nset
      do i2 = i1, nset
        do k = 1, 8
          close(ubin(k, i1, i2))
        end do
        close(uvp(i1, i2))
      end do
    end do
  end if
  close(41)
  close(42)
  close(43)
  close(44)
  if (divstat) close(45)
601 format ( 2x,8(1pe15.5) )
602 format ( 2x,7(1pe15.5) )
  write(*,*) 'total time              =', dble(c1 - c0) / dble(crate)
  write(*,*) 'GPU time in ddns_step   =', t_total / 1000.d0
  write(*,*) 'Next run starts at time =', time

  ! ---- final velocity, forcing and droplets as one-rank restart files (turb.f90:722-729)
  allocate(vel(nx + 2, nx + 2, nx, 6))
  stat = ddns_get_flow(h, vel(:,:,:,1), vel(:,:,:,2), vel(:,:,:,3), vel(:,:,:,4), vel(:,:,:,5), vel(:,:,:,6))
  call ddns_check(stat, h, 'ddns_get_flow')
  open(10, file=dump_name('vel2D', idp + 1, 0), status='unknown', form='unformatted')
  do c = 1, 6
    write(10) vel(:,:,:,c)
  end do
  close(10)
  deallocate(vel)
  if (forcing == 2) then
    stat = ddns_get_forcing_state(h, bforce, isf, iyf, ivf)
    open(10, file=force_name(idp + 1), status='unknown', form='unformatted')
    write(10) isf
    write(10) ivf
    write(10) iyf
    write(10) bforce
    close(10)
  end if
  if (npart > 0) then
    allocate(yp(8, npart), fvp(3, 0:2, npart), fdp(3, 0:3, npart), pert(3, npart), vpd(3, npart), dmv(3, npart))
    stat = ddns_get_droplets(h, np_c, yp, fvp, fdp, pert, vpd, dmv)
    call ddns_check(stat, h, 'ddns_get_droplets')
    open(12, file=rank_name('part.', 0), status='unknown', form='unformatted')
    write(12) int(np_c)
    write(12) yp(:, 1:np_c), fvp(:, :, 1:np_c), fdp(:, :, 1:np_c)
    close(12)
  end if
  stat = ddns_destroy(h)
  write(*,*) 'Successfully completed'

contains

  ! COLLISION_DET's output after the step, at the step's start time (partlib.f90:1641-1797):
  ! the bin increments are summed and written every fifth step (mod(istep+1, 5) = 0), with
  ! thetabin as the mean angle per pair; collision.dat and detectpairs.dat get a line when
  ! they have a count; each pair in contact adds a line to Vp<index1><index2>.dat.
  subroutine collision_files()
    stat = ddns_get_collision_counts(h, cc)
    call ddns_check(stat, h, 'ddns_get_collision_counts')
    stat = ddns_get_collision_bins(h, nirb0, wb0(:,:,:,1), wb0(:,:,:,2), wb0(:,:,:,3), wb0(:,:,:,4), &
                                   wb0(:,:,:,5), wb0(:,:,:,6), wb0(:,:,:,7))
    call ddns_check(stat, h, 'ddns_get_collision_bins')
    nirb = nirb + nirb0                                                ! 1649-1656
    wb = wb + wb0
    if (mod(istep + 1, 5) == 0) then                                   ! 1658-1712
      where (nirb /= 0) wb(:,:,:,6) = wb(:,:,:,6) / real(nirb)
      do i1 = 1, nset
        do i2 = i1, nset
          write(ubin(1, i1, i2), fmt125) nirb(i1, i2, :)
          do k = 1, 7
            write(ubin(k + 1, i1, i2), fmt126) wb(i1, i2, :, k)
          end do
        end do
      end do
      nirb = 0
      wb = 0.d0
    end if
    itot = cc%icc(1) + cc%icc(2) + cc%icc(3)                           ! 1714-1727
    if (itot > 0) write(40, '(1pe15.5, 7i10)') time, cc%icc, cc%colltype(1, 1), cc%colltype(1, 2), &
                                                cc%colltype(2, 2)
    itot = sum(cc%nir(1:nset, 1:nset))                                 ! 1729-1739
    if (itot > 0) write(51, '(1pe15.5, 3i10)') time, cc%nir(1, 1), cc%nir(1, 2), cc%nir(2, 2)
    stat = ddns_get_contact_records(h, crec, int(size(crec, 2), c_int), ncrec)
    call ddns_check(stat, h, 'ddns_get_contact_records')
    if (ncrec > size(crec, 2)) then
      deallocate(crec)
      allocate(crec(12, ncrec))
      stat = ddns_get_contact_records(h, crec, ncrec, ncrec)
      call ddns_check(stat, h, 'ddns_get_contact_records')
    end if
    do k = 1, ncrec                                                    ! 1775-1797
      i1 = int(crec(1, k))
      i2 = int(crec(2, k))
      write(uvp(i1, i2), '(f10.6, 10(1x, f11.5))') time, crec(3:12, k)
    end do
  end subroutine collision_files

  ! 'part.0007', load_save.f90:107-116
  function rank_name(stem, id) result(f)
    character(len=*), intent(in) :: stem
    integer, intent(in) :: id
    character(len=64) :: f
    write(f, '(a, i4.4)') stem, id
  end function rank_name

  ! 'vel2D3.0007' / 'vel2D12.0007', load_save.f90:15-31
  function dump_name(stem, idump, id) result(f)
    character(len=*), intent(in) :: stem
    integer, intent(in) :: idump, id
    character(len=64) :: f
    if (idump < 10) then
      write(f, '(a, i1, a, i4.4)') stem, idump, '.', id
    else
      write(f, '(a, i2, a, i4.4)') stem, idump, '.', id
    end if
  end function dump_name

  ! 'force.3' / 'force.12', load_save.f90:71-77
  function force_name(idump) result(f)
    integer, intent(in) :: idump
    character(len=64) :: f
    if (idump < 10) then
      write(f, '(a, i1)') 'force.', idump
    else
      write(f, '(a, i2)') 'force.', idump
    end if
  end function force_name

  ! input_output (load_save.f90:2-56) over the ranks of a run: rank id holds the kx
  ! pair columns of Y-rank mod(id, nprocY) and the ky planes of Z-rank id / nprocY
  ! (turb.f90:282-295), each as six records of v(nx+2, ly+lyext, lz)
  subroutine read_velocity(idump, nranks, npy, n, v)
    integer, intent(in) :: idump, nranks, npy, n
    double precision, intent(out) :: v(n + 2, n + 2, n, 6)
    integer :: id, npz, ly, lz, lyx, iy0, iz0, f
    npz = nranks / npy
    ly = n / npy
    lz = n / npz
    v = 0.d0
    do id = 0, nranks - 1
      lyx = 0
      if (mod(id, npy) + 1 == npy) lyx = 2
      iy0 = mod(id, npy) * ly
      iz0 = (id / npy) * lz
      allocate(blk(n + 2, ly + lyx, lz))
      open(10, file=dump_name('vel2D', idump, id), status='old', form='unformatted')
      do f = 1, 6
        read(10) blk
        v(:, iy0 + 1:iy0 + ly + lyx, iz0 + 1:iz0 + lz, f) = blk
      end do
      close(10)
      deallocate(blk)
    end do
  end subroutine read_velocity

  ! input_outputf (load_save.f90:59-95)
  subroutine read_forcing(idump, iseed, iv, iy, b)
    integer, intent(in) :: idump
    integer, intent(out) :: iseed, iv(32), iy
    double precision, intent(out) :: b(450)
    open(10, file=force_name(idump), status='old', form='unformatted')
    read(10) iseed
    read(10) iv
    read(10) iy
    read(10) b
    close(10)
  end subroutine read_forcing

end program turb_gpu
