!-----------------------------------------------------------------------
! avism_c -- ISO_C_BINDING interfaces to libavism_hip.so (include/avism.h).
!
! Every entry point returns 0 or a negative AVISM_ERR_* code; avism_check()
! turns a failure into the reference's behaviour (message, then STOP).
!-----------------------------------------------------------------------
module avism_c
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: AVISM_OK = 0
  integer(c_int32_t), parameter :: AVISM_HOST = 0, AVISM_DEVICE = 1
  integer(c_int32_t), parameter :: AVISM_GROW_PLAIN = 0, AVISM_GROW_PROPOSAL = 1, AVISM_GROW_LEARNED = 2
  integer(c_int32_t), parameter :: AVISM_CNN_NPARAM = 19064

  type, bind(C) :: avism_grid_info
    integer(c_int32_t) :: n, periodic, nplus, low1, low2, next
    real(c_float) :: L0, lado0plus, dx, rx1
  end type

  type, bind(C) :: avism_void_params
    real(c_float) :: dens_thre, dens_thre2, grad_thre, div_thre, rmin, rmin_sub
    integer(c_int32_t) :: nvoid_max
  end type

  type, bind(C) :: avism_level_stats
    integer(c_int64_t) :: ncand
    integer(c_int32_t) :: ncubes, nvoid2, nvoidt, nperiodic
    real(c_double) :: volt_clean
    real(c_float) :: meandens, ff
  end type

  interface
    integer(c_int) function avism_create(ctx, device) bind(C, name="avism_create")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr) :: ctx
      integer(c_int32_t), value :: device
    end function
    integer(c_int) function avism_destroy(ctx) bind(C, name="avism_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function avism_last_error(ctx) bind(C, name="avism_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function avism_set_grid(ctx, n, L0, periodic) bind(C, name="avism_set_grid")
      import :: c_ptr, c_int, c_int32_t, c_float
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n, periodic
      real(c_float), value :: L0
    end function
    integer(c_int) function avism_get_grid(ctx, info, rad_ext) bind(C, name="avism_get_grid")
      import :: c_ptr, c_int, avism_grid_info
      type(c_ptr), value :: ctx
      type(avism_grid_info) :: info
      type(c_ptr), value :: rad_ext
    end function
    integer(c_int) function avism_set_cosmology(ctx, ache, omega0, zeta) bind(C, name="avism_set_cosmology")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float), value :: ache, omega0, zeta
    end function
    integer(c_int) function avism_set_void_params(ctx, p) bind(C, name="avism_set_void_params")
      import :: c_ptr, c_int, avism_void_params
      type(c_ptr), value :: ctx
      type(avism_void_params) :: p
    end function
    integer(c_int) function avism_load_particles(ctx, np, x, y, z, vx, vy, vz, m, where) &
        bind(C, name="avism_load_particles")
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_float
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: np
      real(c_float) :: x(*), y(*), z(*), vx(*), vy(*), vz(*), m(*)
      integer(c_int32_t), value :: where
    end function
    integer(c_int) function avism_particles_to_grid(ctx, k, vel_interp, dens_interp) &
        bind(C, name="avism_particles_to_grid")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: k, vel_interp, dens_interp
    end function
    integer(c_int) function avism_particles_to_density(ctx, k, dens_interp) &
        bind(C, name="avism_particles_to_density")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: k, dens_interp
    end function
    integer(c_int) function avism_set_mask(ctx, smask, where) bind(C, name="avism_set_mask")
      import :: c_ptr, c_int, c_int8_t, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int8_t) :: smask(*)
      integer(c_int32_t), value :: where
    end function
    integer(c_int) function avism_load_randoms(ctx, nrand, x, y, z, m, where) &
        bind(C, name="avism_load_randoms")
      import :: c_ptr, c_int, c_int64_t, c_int32_t, c_float
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nrand
      real(c_float) :: x(*), y(*), z(*), m(*)
      integer(c_int32_t), value :: where
    end function
    integer(c_int) function avism_survey_density(ctx, k, dens_interp) &
        bind(C, name="avism_survey_density")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: k, dens_interp
    end function
    integer(c_int) function avism_load_grid_fields(ctx, delta, vx, vy, vz, where) &
        bind(C, name="avism_load_grid_fields")
      import :: c_ptr, c_int, c_int32_t, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: delta(*), vx(*), vy(*), vz(*)
      integer(c_int32_t), value :: where
    end function
    integer(c_int) function avism_get_fields(ctx, u1co, div, u, v, w, where) bind(C, name="avism_get_fields")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, u1co, div, u, v, w
      integer(c_int32_t), value :: where
    end function
    integer(c_int) function avism_find_voids(ctx, growth_mode, ncubes) bind(C, name="avism_find_voids")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: growth_mode
      integer(c_int32_t) :: ncubes
    end function
    integer(c_int) function avism_get_cubes(ctx, ci, cf, father) bind(C, name="avism_get_cubes")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, ci, cf, father
    end function
    integer(c_int) function avism_merge_and_post(ctx, nvoids) bind(C, name="avism_merge_and_post")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: nvoids
    end function
    integer(c_int) function avism_get_catalogue(ctx, indice, uvoid, ncellv, vf, stats) &
        bind(C, name="avism_get_catalogue")
      import :: c_ptr, c_int, avism_level_stats
      type(c_ptr), value :: ctx, indice, uvoid, ncellv, vf
      type(avism_level_stats) :: stats
    end function
    integer(c_int) function avism_get_marca(ctx, marca, inner_only, where) bind(C, name="avism_get_marca")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, marca
      integer(c_int32_t), value :: inner_only, where
    end function
    integer(c_int) function avism_next_level(ctx) bind(C, name="avism_next_level")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    ! refined levels in windows around the parent voids (include/avism.h)
    integer(c_int) function avism_refine_plan(ctx, margin, max_cells, natlas) bind(C, name="avism_refine_plan")
      import :: c_ptr, c_int, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: margin
      integer(c_int64_t), value :: max_cells
      integer(c_int32_t) :: natlas
    end function
    integer(c_int) function avism_refine_load(ctx, atlas, k) bind(C, name="avism_refine_load")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: atlas, k
    end function
    integer(c_int) function avism_refine_windows(ctx, tab, cap, n, info) bind(C, name="avism_refine_windows")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, tab, info
      integer(c_int32_t), value :: cap
      integer(c_int32_t) :: n
    end function
    integer(c_int) function avism_refine_pop(ctx) bind(C, name="avism_refine_pop")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function avism_cnn_set_params(ctx, params, where) bind(C, name="avism_cnn_set_params")
      import :: c_ptr, c_int, c_int32_t, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: params(*)
      integer(c_int32_t), value :: where
    end function
  end interface

contains

  ! message of the last failed call on ctx
  function avism_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    p = avism_last_error(ctx)
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

  ! the reference STOPs on a failed check; do the same with the library's message
  subroutine avism_check(ctx, rc, what)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= AVISM_OK) then
      write(*,*) '     ERROR in ', what, ' (code', rc, '): ', avism_message(ctx)
      stop 1
    end if
  end subroutine

end module avism_c
