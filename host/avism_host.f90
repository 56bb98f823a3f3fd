!-----------------------------------------------------------------------
! avism_host -- Fortran host of the MI355X grid-and-void path.
!
! Drop-in for the grid-and-void path of PROGRAM AVISM (avism.f90): reads
! config/voids.dat with the same positional READs (avism.f90:176-253), reads
! input_data/bin_file_partNNNNN (reader.f90:410-450, FLAG_DATA=1) or
! input_data/bin_file_gridNNNNN (reader.f90:521-540, FLAG_DATA=2), runs the
! path through the C ABI (host/avism_c.f90) for the levels LEVMIN..LEVMAX
! (avism.f90:742-2181) and writes output_files/voidsNNNNN, mapNNNNN and, with
! FLAG_WRITE_CUBES=1, cubesNNNNN as avism.f90:707-722 and 2028-2085 do, one
! block per level.
!
! Void-in-void levels (particle input): level IR has NCOX*2**(IR-LEVMIN)
! cells per axis (avism.f90:755-758); avism_next_level hands MARCAP and REQP
! over (avism.f90:2123-2161) and the same particles are interpolated again
! (the reference frees them inside its level loop, avism.f90:1613-1629).
! The catalogue's parent radius column is REQP of the parent void (the
! reference prints REQ(INDP), the current level's array at the parent's id,
! avism.f90:2076-2078); NVOIDP is 0 at the first level (uninitialised in the
! reference, avism.f90:70).
!
! Usage:  avism_host.x [--learned weights.f32]
!   --learned: growth with the half-width proposal network whose fp32
!   parameters (19064 values, raw little-endian) are in the given file; the
!   catalogue is the same as without it.
!
! Not covered (the library returns or this program stops with a message):
! LEVMAX > LEVMIN with grid input, survey mode, MASCLET/GADGET input, an input
! grid size different from NCOX.
!-----------------------------------------------------------------------
program avism_host
  use iso_c_binding
  use avism_c
  implicit none

  ! voids.dat
  integer :: first, last, every, ncox, ncoy, ncoz, levmin, levmax, flag_period
  real :: ache, omega0, lado0
  integer :: nvoid_max
  real :: dens_thre, dens_thre2, grad_thre, div_thre, rmin, rmin_sub
  integer :: flag_data, nhyx, nhyy, nhyz, files_per_snap, parttypex, flag_gadget_reader, flag_survey
  real :: massdm
  integer(8) :: partired
  integer :: flag_vel_available, flag_vel_interp, flag_dens_interp, kneighbours, flag_write_cubes
  integer :: npalev, nlevels, namrx, namry, namrz, flag_div, flag_dens

  type(c_ptr) :: ctx
  type(avism_void_params) :: vp
  type(avism_level_stats) :: st
  integer(c_int) :: rc
  integer(c_int32_t) :: ncubes, nvoids, growth
  integer :: niter, ifi, iter, n, nlev, ir, nvoidp
  real(c_float), allocatable :: reqp(:)
  real :: zeta
  real(c_float), allocatable, target :: params(:)
  character(len=512) :: arg, weights
  character(len=:), allocatable :: dir

  growth = AVISM_GROW_PLAIN
  weights = ''
  if (command_argument_count() >= 2) then
    call get_command_argument(1, arg)
    if (trim(arg) == '--learned') then
      call get_command_argument(2, weights)
      growth = AVISM_GROW_LEARNED
    end if
  end if

  call read_voids_dat('config/voids.dat')
  call check_supported()
  dir = 'output_files'

  rc = avism_create(ctx, 0_c_int32_t)
  if (rc /= AVISM_OK) then
    write(*,*) '     ERROR: no HIP device for the AVISM library (code', rc, ')'
    stop 1
  end if
  call avism_check(ctx, avism_set_grid(ctx, int(ncox, c_int32_t), real(lado0, c_float), &
                                    int(flag_period, c_int32_t)), 'avism_set_grid')
  vp = avism_void_params(dens_thre, dens_thre2, grad_thre, div_thre, rmin, rmin_sub, nvoid_max)
  call avism_check(ctx, avism_set_void_params(ctx, vp), 'avism_set_void_params')
  if (growth == AVISM_GROW_LEARNED) then
    allocate(params(AVISM_CNN_NPARAM))
    open(unit=21, file=trim(weights), form='unformatted', access='stream', status='old', action='read')
    read(21) params
    close(21)
    call avism_check(ctx, avism_cnn_set_params(ctx, params, AVISM_HOST), 'avism_cnn_set_params')
  end if

  niter = int((last - first) / every) + 1
  nlev = levmax - levmin + 1
  do ifi = 1, niter
    iter = first + every * (ifi - 1)
    if (ifi > 1) call avism_check(ctx, avism_set_grid(ctx, int(ncox, c_int32_t), real(lado0, c_float), &
                                                   int(flag_period, c_int32_t)), 'avism_set_grid')
    call open_outputs(iter)
    nvoidp = 0
    allocate(reqp(0))
    do ir = levmin, levmax
      n = ncox * 2**(ir - levmin)
      if (ir == levmin) then
        if (flag_data == 1) then
          call run_particles(iter, zeta)
        else
          call run_grid(iter, zeta)
        end if
      else
        call avism_check(ctx, avism_next_level(ctx), 'avism_next_level')
        call particle_fields()
      end if
      call avism_check(ctx, avism_find_voids(ctx, growth, ncubes), 'avism_find_voids')
      call avism_check(ctx, avism_merge_and_post(ctx, nvoids), 'avism_merge_and_post')
      call write_level(ir, n, ncubes, nvoidp, reqp)
      write(*,*) 'ITER', iter, ' LEVEL', ir, ': cubes', ncubes, ' voids', nvoids
      nvoidp = ncubes
    end do
    deallocate(reqp)
    close(10)
    close(11)
    if (flag_write_cubes == 1) close(12)
  end do
  rc = avism_destroy(ctx)

contains

  subroutine skip(u, k)
    integer, intent(in) :: u, k
    integer :: i
    do i = 1, k
      read(u, *)
    end do
  end subroutine

  ! positional read, one label (or banner) line before each value line
  subroutine read_voids_dat(path)
    character(len=*), intent(in) :: path
    open(1, file=path, status='old', action='read')
    call skip(1, 5); read(1, *) first, last, every
    call skip(1, 1); read(1, *) ncox, ncoy, ncoz
    call skip(1, 1); read(1, *) levmin, levmax
    call skip(1, 1); read(1, *) ache, omega0
    call skip(1, 1); read(1, *) lado0
    call skip(1, 1); read(1, *) flag_period
    call skip(1, 4); read(1, *) nvoid_max
    call skip(1, 1); read(1, *) dens_thre
    call skip(1, 1); read(1, *) dens_thre2
    call skip(1, 1); read(1, *) grad_thre
    call skip(1, 1); read(1, *) div_thre
    call skip(1, 1); read(1, *) rmin
    call skip(1, 1); read(1, *) rmin_sub
    call skip(1, 4); read(1, *) flag_data
    call skip(1, 1); read(1, *) nhyx, nhyy, nhyz
    call skip(1, 1); read(1, *) files_per_snap, parttypex, massdm
    call skip(1, 1); read(1, *) flag_gadget_reader
    call skip(1, 1); read(1, *) flag_survey
    call skip(1, 4); read(1, *) partired
    call skip(1, 1); read(1, *) flag_vel_available
    call skip(1, 1); read(1, *) flag_vel_interp
    call skip(1, 1); read(1, *) flag_dens_interp
    call skip(1, 1); read(1, *) kneighbours
    call skip(1, 4); read(1, *) flag_write_cubes
    call skip(1, 4); read(1, *) npalev, nlevels, namrx, namry, namrz
    call skip(1, 1); read(1, *) flag_div
    call skip(1, 1); read(1, *) flag_dens
    close(1)
  end subroutine

  subroutine check_supported()
    if (ncox /= ncoy .or. ncox /= ncoz) then
      write(*,*) '     ERROR: NCOX, NCOY and NCOZ must be equal'
      stop 1
    end if
    if (levmax < levmin) then
      write(*,*) '     ERROR: LEVMAX must not be below LEVMIN'
      stop 1
    end if
    if (levmax /= levmin .and. flag_data /= 1) then
      write(*,*) '     ERROR: void-in-void levels (LEVMAX > LEVMIN) need particle input (FLAG_DATA = 1) in this build'
      stop 1
    end if
    if (flag_data /= 1 .and. flag_data /= 2) then
      write(*,*) '     ERROR: only FLAG_DATA = 1 (BinPart) and 2 (BinGrid) are supported by this build'
      stop 1
    end if
    if (flag_survey == 1 .and. flag_period == 1) then  ! avism.f90:278-281
      write(*,*) 'PBCs considered in voids.dat but SURVEY/MASKED conditions activated'
      stop 1
    end if
    if (flag_survey == 1 .and. (flag_data /= 1 .or. flag_vel_available /= 0)) then
      write(*,*) '     ERROR: survey mode needs particle input without velocities in this build'
      stop 1
    end if
    if (flag_vel_available == 0 .and. flag_data /= 1) then
      write(*,*) '     ERROR: velocity-free input (FLAG_VEL_AVAILABLE = 0) needs particle input in this build'
      stop 1
    end if
    if (flag_data == 2 .and. (nhyx /= ncox .or. nhyy /= ncoy .or. nhyz /= ncoz)) then
      write(*,*) '     ERROR: the input grid must have NCOX^3 cells in this build'
      stop 1
    end if
    if (kneighbours < 0) then
      write(*,*) '     ERROR: KNEIGHBOURS must be positive'
      stop 1
    end if
  end subroutine

  function iter_tag(it) result(nom)
    integer, intent(in) :: it
    character(len=5) :: nom
    write(nom, '(I5.5)') it
  end function

  subroutine run_particles(it, z)
    integer, intent(in) :: it
    real, intent(out) :: z
    integer(8) :: np
    real(c_float), allocatable :: x(:), y(:), zz(:), vx(:), vy(:), vz(:), m(:)
    open(unit=41, file='input_data/bin_file_part'//iter_tag(it), form='unformatted', status='old', &
         action='read')
    read(41) np, z
    if (np > partired) then
      write(*,*) '     ERROR: NPARTT > PARTIRED'
      stop 1
    end if
    allocate(x(np), y(np), zz(np), vx(np), vy(np), vz(np), m(np))
    read(41) x
    read(41) y
    read(41) zz
    if (flag_vel_available == 1) then  ! READ_BINARY_PART_1 (reader.f90:447-456)
      read(41) vx
      read(41) vy
      read(41) vz
    else  ! READ_BINARY_PART_2 (reader.f90:506-509): no velocity records
      vx = 0.0
      vy = 0.0
      vz = 0.0
    end if
    read(41) m
    close(41)
    write(*,*) '     TOTAL PARTICLES IN ITER=', np
    call avism_check(ctx, avism_set_cosmology(ctx, ache, omega0, z), 'avism_set_cosmology')
    call avism_check(ctx, avism_load_particles(ctx, int(np, c_int64_t), x, y, zz, vx, vy, vz, m, AVISM_HOST), &
                  'avism_load_particles')
    if (flag_survey == 1) call load_survey()
    call particle_fields()
  end subroutine

  ! the field stage on the loaded particles: with velocities (avism.f90:1159-1257)
  ! or without (FLAG_VEL_AVAILABLE = 0, avism.f90:1259-1300, 1646-1650)
  ! survey mode inputs (avism.f90:403-418, 1431-1441): input_data/mask.dat
  ! (stream: NMASK, three REAL*4, NMASK^3 INTEGER*1; NMASK = NCOX) and
  ! input_data/rands.dat (sequential: NRAND, then x, y, z, masses)
  subroutine load_survey()
    integer :: nmask
    integer(c_int32_t) :: nrand
    real :: mlow, mup, mside
    integer(c_int8_t), allocatable :: sm(:)
    real(c_float), allocatable :: rx(:), ry(:), rz(:), rm(:)
    open(unit=16, file='input_data/mask.dat', access='stream', form='unformatted', status='old', action='read')
    read(16) nmask
    if (nmask /= ncox) then
      write(*,*) 'MASK grid does not coincide with AVISM coarse grid', nmask, ncox
      stop 1
    end if
    read(16) mlow, mup, mside
    allocate(sm(int(nmask, 8)**3))
    read(16) sm
    close(16)
    call avism_check(ctx, avism_set_mask(ctx, sm, AVISM_HOST), 'avism_set_mask')
    open(unit=17, file='input_data/rands.dat', form='unformatted', status='old', action='read')
    read(17) nrand
    allocate(rx(nrand), ry(nrand), rz(nrand), rm(nrand))
    read(17) rx
    read(17) ry
    read(17) rz
    read(17) rm
    close(17)
    call avism_check(ctx, avism_load_randoms(ctx, int(nrand, c_int64_t), rx, ry, rz, rm, AVISM_HOST), &
                     'avism_load_randoms')
  end subroutine

  subroutine particle_fields()
    if (flag_survey == 1) then
      call avism_check(ctx, avism_survey_density(ctx, int(kneighbours, c_int32_t), &
                                              int(flag_dens_interp, c_int32_t)), 'avism_survey_density')
    else if (flag_vel_available == 1) then
      call avism_check(ctx, avism_particles_to_grid(ctx, int(kneighbours, c_int32_t), &
                                                 int(flag_vel_interp, c_int32_t), &
                                                 int(flag_dens_interp, c_int32_t)), 'avism_particles_to_grid')
    else
      call avism_check(ctx, avism_particles_to_density(ctx, int(kneighbours, c_int32_t), &
                                                    int(flag_dens_interp, c_int32_t)), 'avism_particles_to_density')
    end if
  end subroutine

  subroutine run_grid(it, z)
    integer, intent(in) :: it
    real, intent(out) :: z
    real(c_float), allocatable :: d(:,:,:), u(:,:,:), v(:,:,:), w(:,:,:)
    allocate(d(nhyx, nhyy, nhyz), u(nhyx, nhyy, nhyz), v(nhyx, nhyy, nhyz), w(nhyx, nhyy, nhyz))
    open(unit=41, file='input_data/bin_file_grid'//iter_tag(it), form='unformatted', status='old', &
         action='read')
    read(41) z
    read(41) d
    read(41) u
    read(41) v
    read(41) w
    close(41)
    call avism_check(ctx, avism_set_cosmology(ctx, ache, omega0, z), 'avism_set_cosmology')
    call avism_check(ctx, avism_load_grid_fields(ctx, d, u, v, w, AVISM_HOST), 'avism_load_grid_fields')
  end subroutine

  ! file headers (avism.f90:707-722); the map file takes one block per level
  subroutine open_outputs(it)
    integer, intent(in) :: it
    open(unit=10, file=dir//'/voids'//iter_tag(it), status='replace')
    write(10, *) nlev, levmin, levmax, ncox, ncoy, ncoz, lado0
    open(unit=11, file=dir//'/map'//iter_tag(it), form='unformatted', access='stream', status='replace')
    if (flag_write_cubes == 1) then
      open(unit=12, file=dir//'/cubes'//iter_tag(it), status='replace')
      write(12, *) nlev, levmin, levmax
    end if
  end subroutine

  ! one level's block of the catalogue, cubes and map files (avism.f90:2028-2085);
  ! on return reqp holds this level's REQP (REQ of the main voids, by id) for the next
  subroutine write_level(lev, nn, nv, nvp, reqp)
    integer, intent(in) :: lev, nn, nvp
    integer(c_int32_t), intent(in) :: nv
    real(c_float), allocatable, intent(inout) :: reqp(:)
    integer(c_int32_t), allocatable, target :: indice(:), uvoid(:), ncellv(:), marca(:), ci(:,:)
    real(c_float), allocatable, target :: vf(:,:), u1co(:), div(:), cf(:,:)
    integer(c_int32_t), allocatable, target :: father(:)
    integer :: iv, ind, nvs
    integer(8) :: ncell
    real :: volm, req, reqpp
    integer :: indp

    nvs = max(int(nv), 1)
    allocate(indice(nvs), uvoid(nvs), ncellv(nvs), vf(nvs, 12))
    call avism_check(ctx, avism_get_catalogue(ctx, c_loc(indice), c_loc(uvoid), c_loc(ncellv), c_loc(vf), st), &
                  'avism_get_catalogue')
    allocate(ci(nvs, 9), cf(nvs, 7), father(nvs))
    call avism_check(ctx, avism_get_cubes(ctx, c_loc(ci), c_loc(cf), c_loc(father)), 'avism_get_cubes')
    ncell = int(nn, 8)**3
    allocate(marca(ncell), u1co(ncell), div(ncell))
    call avism_check(ctx, avism_get_marca(ctx, c_loc(marca), 1_c_int32_t, AVISM_HOST), 'avism_get_marca')
    call avism_check(ctx, avism_get_fields(ctx, c_loc(u1co), c_loc(div), c_null_ptr, c_null_ptr, c_null_ptr, &
                                        AVISM_HOST), 'avism_get_fields')

    ! level header (avism.f90:2047), one line per main void
    write(10, *) lev, int(nv), int(st%nvoidt), nvp, real(st%volt_clean / (real(lado0, 8)**3)), st%meandens
    if (flag_write_cubes == 1) write(12, *) lev, int(nv), int(st%nvoidt)
    do iv = 1, int(nv)
      ind = indice(iv)
      if (flag_write_cubes == 1) then
        write(12, *) ind, uvoid(ind), ci(ind, 1), ci(ind, 4), ci(ind, 2), ci(ind, 5), ci(ind, 3), ci(ind, 6), &
                     cf(ind, 1), cf(ind, 4), cf(ind, 2), cf(ind, 5), cf(ind, 3), cf(ind, 6)
      end if
      if (uvoid(ind) /= -1) cycle
      volm = vf(ind, 7)
      req = vf(ind, 12)
      indp = 0
      reqpp = 0.0
      if (lev > levmin) then
        indp = father(ind)
        if (indp >= 1 .and. indp <= size(reqp)) reqpp = reqp(indp)
      end if
      write(10, *) ind, vf(ind, 4), vf(ind, 5), vf(ind, 6), vf(ind, 1), vf(ind, 2), vf(ind, 3), volm, req, &
                   vf(ind, 8), vf(ind, 10), vf(ind, 11), indp, reqpp, vf(ind, 9)
    end do

    ! map: MARCA, delta and divergence on the inner grid (avism.f90:2027-2037)
    write(11) marca
    write(11) u1co - 1.0
    write(11) div

    ! REQP(IND) = REQ(IND) for the main voids (avism.f90:2138-2144)
    deallocate(reqp)
    allocate(reqp(int(nv)))
    do ind = 1, int(nv)
      reqp(ind) = 0.0
      if (uvoid(ind) == -1) reqp(ind) = vf(ind, 12)
    end do
  end subroutine

end program avism_host
