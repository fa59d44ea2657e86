! kg_hip_mod.f90 -- ISO_C_BINDING layer between a Fortran driver of
! FlavioRosales/KG and the MI355X engine (include/kg.h, kg_amd/libkg_hip.so).
!
! The reference driver holds `type(scalar_field) :: MoL` and calls
! `MoL%pirk2()` once per step (main/evolve.f90:22-26); the state is
! `MoL%u(nvars, iL-g:iR+g, jL-g:jR+g, kL:kR)`, complex(8)
! (MoL/lines_method_lib.f90:27,135-138).  `type(kg_lines_method)` below keeps
! that shape of use: construct it with the arguments of
! lines_method_constructor (:49-51), hand it `u` with `upload`, call
! `pirk2`/`pirk3`/`rk3`/`rk4` or `step(n)`, and read `u` back with
! `download` when the diagnostics need it.  Every nonzero status prints the
! engine's message and stops, the reference's error convention
! (lines_method_lib.f90:529-533,744-748).
module kg_hip_mod
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: KG_OK = 0
  integer(c_int32_t), parameter, public :: KG_SCHEME_PIRK2 = 0, KG_SCHEME_PIRK3 = 1, KG_SCHEME_RK3 = 2, KG_SCHEME_RK4 = 3
  integer(c_int32_t), parameter, public :: KG_METRIC_MINKOWSKI = 0, KG_METRIC_KERR_SCHILD = 1, &
                                           KG_METRIC_MKS = 2, KG_METRIC_SCHW_EF = 3
  integer(c_int32_t), parameter, public :: KG_BC_OUTFLOW = 0, KG_BC_INFLOW = 1, KG_BC_SOMMERFELD = 2
  integer(c_int32_t), parameter, public :: KG_LAYOUT_FORTRAN = 0, KG_LAYOUT_SOA = 1
  integer(c_int32_t), parameter, public :: KG_OPS_REF2 = 0, KG_OPS_FD4 = 1

  type, bind(C), public :: kg_desc
    integer(c_int32_t) :: nvars, g_pts, Nx, Ny, Nz
    integer(c_int32_t) :: scheme, metric, bc_rmax, is_complex, op_set
    real(c_double) :: xmin, xmax, tmin, CFL, dt, a, M_bh, m
  end type kg_desc

  ! initial data (include/kg.h: struct kg_ic_desc); seed holds the bits of a uint64_t
  integer(c_int32_t), parameter, public :: KG_IC_SPHERICAL_SHELL = 0, KG_IC_RANDOM = 1, KG_IC_RANDOM_SHELL = 2
  type, bind(C), public :: kg_ic_desc
    integer(c_int32_t) :: type, ncart
    real(c_double) :: mu, sigma, p0
    integer(c_int64_t) :: seed
    type(c_ptr) :: uniforms
  end type kg_ic_desc

  ! two refinement levels (include/kg.h: struct kg_refine_desc)
  type, bind(C), public :: kg_refine_desc
    real(c_double) :: r_patch
    integer(c_int32_t) :: buffer, reserved
  end type kg_refine_desc

  interface
    function kg_create(desc, h) bind(C, name="kg_create") result(st)
      import :: kg_desc, c_ptr, c_int32_t
      type(kg_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: h
      integer(c_int32_t) :: st
    end function kg_create
    function kg_destroy(h) bind(C, name="kg_destroy") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t) :: st
    end function kg_destroy
    function kg_last_error(h) bind(C, name="kg_last_error") result(msg)
      import :: c_ptr
      type(c_ptr), value :: h
      type(c_ptr) :: msg
    end function kg_last_error
    function kg_get_time(h, t, dt) bind(C, name="kg_get_time") result(st)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(out) :: t, dt
      integer(c_int32_t) :: st
    end function kg_get_time
    function kg_set_time(h, t) bind(C, name="kg_set_time") result(st)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), value :: t
      integer(c_int32_t) :: st
    end function kg_set_time
    function kg_upload_state(h, u, layout) bind(C, name="kg_upload_state") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h, u
      integer(c_int32_t), value :: layout
      integer(c_int32_t) :: st
    end function kg_upload_state
    function kg_download_state(h, u, layout) bind(C, name="kg_download_state") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h, u
      integer(c_int32_t), value :: layout
      integer(c_int32_t) :: st
    end function kg_download_state
    function kg_upload_aux(h, which, u, layout) bind(C, name="kg_upload_aux") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h, u
      integer(c_int32_t), value :: which, layout
      integer(c_int32_t) :: st
    end function kg_upload_aux
    function kg_download_aux(h, which, u, layout) bind(C, name="kg_download_aux") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h, u
      integer(c_int32_t), value :: which, layout
      integer(c_int32_t) :: st
    end function kg_download_aux
    function kg_fill_ghosts(h) bind(C, name="kg_fill_ghosts") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t) :: st
    end function kg_fill_ghosts
    function kg_pirk2(h) bind(C, name="kg_pirk2") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t) :: st
    end function kg_pirk2
    function kg_pirk3(h) bind(C, name="kg_pirk3") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t) :: st
    end function kg_pirk3
    function kg_rk3(h) bind(C, name="kg_rk3") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t) :: st
    end function kg_rk3
    function kg_rk4(h) bind(C, name="kg_rk4") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t) :: st
    end function kg_rk4
    function kg_step(h, nsteps) bind(C, name="kg_step") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsteps
      integer(c_int32_t) :: st
    end function kg_step
    function kg_sync(h) bind(C, name="kg_sync") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t) :: st
    end function kg_sync
    function kg_batch_step(hs, n, nsteps) bind(C, name="kg_batch_step") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), intent(in) :: hs(*)
      integer(c_int32_t), value :: n, nsteps
      integer(c_int32_t) :: st
    end function kg_batch_step
    function kg_initial_data(h, desc, amp) bind(C, name="kg_initial_data") result(st)
      import :: kg_ic_desc, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: h
      type(kg_ic_desc), intent(in) :: desc
      real(c_double), intent(out) :: amp
      integer(c_int32_t) :: st
    end function kg_initial_data
    function kg_get_dims(h, dims) bind(C, name="kg_get_dims") result(st)
      import :: c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: dims(8)
      integer(c_int32_t) :: st
    end function kg_get_dims
    function kg_refined_create(base, desc, r) bind(C, name="kg_refined_create") result(st)
      import :: kg_desc, kg_refine_desc, c_ptr, c_int32_t
      type(kg_desc), intent(in) :: base
      type(kg_refine_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: r
      integer(c_int32_t) :: st
    end function kg_refined_create
    function kg_refined_levels(r, hb, hf) bind(C, name="kg_refined_levels") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: r
      type(c_ptr), intent(out) :: hb, hf
      integer(c_int32_t) :: st
    end function kg_refined_levels
    function kg_refined_step(r, nsteps) bind(C, name="kg_refined_step") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: r
      integer(c_int32_t), value :: nsteps
      integer(c_int32_t) :: st
    end function kg_refined_step
    function kg_refined_destroy(r) bind(C, name="kg_refined_destroy") result(st)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: r
      integer(c_int32_t) :: st
    end function kg_refined_destroy
    function kg_refined_last_error(r) bind(C, name="kg_refined_last_error") result(msg)
      import :: c_ptr
      type(c_ptr), value :: r
      type(c_ptr) :: msg
    end function kg_refined_last_error
    function kg_ic_last_error() bind(C, name="kg_ic_last_error") result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function kg_ic_last_error
  end interface

  public :: kg_initial_data, kg_ic_last_error
  public :: kg_create, kg_destroy, kg_last_error, kg_get_time, kg_set_time, kg_upload_state, &
            kg_download_state, kg_upload_aux, kg_download_aux, kg_fill_ghosts, kg_pirk2, kg_pirk3, &
            kg_rk3, kg_rk4, kg_step, kg_sync, kg_batch_step, kg_lines_method_batch_step

  ! The reference's lines_method as seen by its driver.
  type, public :: kg_lines_method
    type(c_ptr) :: h = c_null_ptr
    integer :: nvars, g_pts, Nx, Ny, Nz
    integer :: iL, iR, jL, jR, kL, kR
    real(kind=8) :: tmin, tmax, CFL, dt, t
    integer :: Nt
  contains
    procedure :: upload => klm_upload
    procedure :: download => klm_download
    procedure :: boundary => klm_boundary
    procedure :: pirk2 => klm_pirk2
    procedure :: pirk3 => klm_pirk3
    procedure :: rk3 => klm_rk3
    procedure :: rk4 => klm_rk4
    procedure :: step => klm_step
    procedure :: initial_conditions => klm_initial_conditions
    procedure :: destroy => klm_destroy
  end type kg_lines_method

  public :: kg_lines_method_constructor

  ! Two fixed refinement levels (no counterpart in the reference): `base` and
  ! `fine` are views of the two levels for upload/download and diagnostics;
  ! `step(n)` advances both (n base steps, two fine steps each).
  type, public :: kg_refined_lines_method
    type(c_ptr) :: r = c_null_ptr
    type(kg_lines_method) :: base, fine
  contains
    procedure :: step => krl_step
    procedure :: destroy => krl_destroy
  end type kg_refined_lines_method

  public :: kg_refined_lines_method_constructor

contains

  subroutine kg_check(st, h)
    integer(c_int32_t), intent(in) :: st
    type(c_ptr), intent(in) :: h
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (st == KG_OK) return
    p = kg_last_error(h)
    call c_f_pointer(p, msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    print*, 'kg engine error ', st, ': ', transfer(msg(1:n), repeat(' ', n))
    stop
  end subroutine kg_check

  ! lines_method_constructor(nvars, g_pts, xmin, xmax, Nx, Ny, Nz, tmin, tmax,
  ! CFL, boundary_type_rmax, metric, a, M_bh, m) -- lines_method_lib.f90:49-129
  ! op_set (optional): KG_OPS_REF2 (default, the reference's stencils) or
  ! KG_OPS_FD4 (its 4th-order set; rk3 and rk4)
  function kg_lines_method_constructor(nvars, g_pts, xmin, xmax, Nx, Ny, Nz, tmin, tmax, CFL, &
                                       boundary_type_rmax, metric, a, M_bh, m, scheme, is_complex, op_set) result(this)
    integer, intent(in) :: nvars, g_pts, Nx, Ny, Nz
    real(kind=8), intent(in) :: xmin, xmax, tmin, tmax, CFL, a, M_bh, m
    character(len=*), intent(in) :: boundary_type_rmax, metric
    integer, intent(in) :: scheme
    logical, intent(in) :: is_complex
    integer, intent(in), optional :: op_set
    type(kg_lines_method) :: this
    type(kg_desc) :: d
    type(c_ptr) :: h
    d = make_desc(nvars, g_pts, xmin, xmax, Nx, Ny, Nz, tmin, CFL, boundary_type_rmax, metric, a, M_bh, m, &
                  scheme, is_complex, op_set)
    call kg_check(kg_create(d, h), c_null_ptr)
    call view_init(this, h, tmin, tmax, CFL)
  end function kg_lines_method_constructor

  ! the engine's descriptor from lines_method_constructor's arguments
  function make_desc(nvars, g_pts, xmin, xmax, Nx, Ny, Nz, tmin, CFL, boundary_type_rmax, metric, a, M_bh, m, &
                     scheme, is_complex, op_set) result(d)
    integer, intent(in) :: nvars, g_pts, Nx, Ny, Nz
    real(kind=8), intent(in) :: xmin, xmax, tmin, CFL, a, M_bh, m
    character(len=*), intent(in) :: boundary_type_rmax, metric
    integer, intent(in) :: scheme
    logical, intent(in) :: is_complex
    integer, intent(in), optional :: op_set
    type(kg_desc) :: d
    d%nvars = nvars; d%g_pts = g_pts; d%Nx = Nx; d%Ny = Ny; d%Nz = Nz
    d%scheme = scheme
    if (metric == 'Kerr_Schild') then
      d%metric = KG_METRIC_KERR_SCHILD
    else if (metric == 'Schw_EF') then
      d%metric = KG_METRIC_SCHW_EF
    else if (metric == 'MKS') then
      d%metric = KG_METRIC_MKS
    else
      d%metric = KG_METRIC_MINKOWSKI
    end if
    if (boundary_type_rmax == 'outflow') then
      d%bc_rmax = KG_BC_OUTFLOW
    else if (boundary_type_rmax == 'inflow') then
      d%bc_rmax = KG_BC_INFLOW
    else if (boundary_type_rmax == 'sommerfeld') then
      d%bc_rmax = KG_BC_SOMMERFELD     ! radiative outer boundary (not in the reference)
    else
      print*, 'This boundary condition type for r_max is not implemented yet'
      print*, 'try with: outflow or inflow'
      stop
    end if
    ! u is complex(8) in the reference (lines_method_lib.f90:27) and the
    ! upload/download procedures below pass it as such, so a real handle
    ! would mix the re/im parts of neighbouring values: refused
    if (.not. is_complex) then
      print*, 'kg_lines_method_constructor: the state u is complex(8); is_complex must be .true.'
      stop
    end if
    d%is_complex = 1
    d%op_set = KG_OPS_REF2
    if (present(op_set)) d%op_set = int(op_set, c_int32_t)
    d%xmin = xmin; d%xmax = xmax; d%tmin = tmin; d%CFL = CFL; d%dt = 0.0d0
    d%a = a; d%M_bh = M_bh; d%m = m
  end function make_desc

  ! the Fortran view of handle h: sizes and index bounds from the engine
  subroutine view_init(this, h, tmin, tmax, CFL)
    type(kg_lines_method), intent(inout) :: this
    type(c_ptr), intent(in) :: h
    real(kind=8), intent(in) :: tmin, tmax, CFL
    integer(c_int64_t) :: dims(8)
    real(c_double) :: t, dt
    this%h = h
    call kg_check(kg_get_dims(h, dims), h)
    this%nvars = 5; this%Nx = int(dims(1)); this%Ny = int(dims(2)); this%Nz = int(dims(3))
    this%g_pts = int(dims(4))
    this%iL = 0; this%iR = this%Nx; this%jL = 0; this%jR = this%Ny; this%kL = 0; this%kR = this%Nz
    call kg_check(kg_get_time(h, t, dt), h)
    this%t = t; this%dt = dt
    this%tmin = tmin; this%tmax = tmax; this%CFL = CFL
    this%Nt = int((tmax - tmin) / dt)           ! lines_method_lib.f90:115-116
    if (mod(this%Nt, 2) /= 0) this%Nt = this%Nt + 1
  end subroutine view_init

  ! The base grid of lines_method_constructor's arguments with a 2:1 radial
  ! patch on [xmin, r_patch] and `buffer` fine columns beyond it (16 if 0).
  function kg_refined_lines_method_constructor(nvars, g_pts, xmin, xmax, Nx, Ny, Nz, tmin, tmax, CFL, &
                                               boundary_type_rmax, metric, a, M_bh, m, scheme, is_complex, &
                                               r_patch, buffer, op_set) result(this)
    integer, intent(in) :: nvars, g_pts, Nx, Ny, Nz
    real(kind=8), intent(in) :: xmin, xmax, tmin, tmax, CFL, a, M_bh, m, r_patch
    character(len=*), intent(in) :: boundary_type_rmax, metric
    integer, intent(in) :: scheme, buffer
    logical, intent(in) :: is_complex
    integer, intent(in), optional :: op_set
    type(kg_refined_lines_method) :: this
    type(kg_desc) :: d
    type(kg_refine_desc) :: rd
    type(c_ptr) :: hb, hf
    d = make_desc(nvars, g_pts, xmin, xmax, Nx, Ny, Nz, tmin, CFL, boundary_type_rmax, metric, a, M_bh, m, &
                  scheme, is_complex, op_set)
    rd%r_patch = r_patch
    rd%buffer = int(buffer, c_int32_t)
    rd%reserved = 0
    if (kg_refined_create(d, rd, this%r) /= KG_OK) call refined_fail(this%r)
    if (kg_refined_levels(this%r, hb, hf) /= KG_OK) call refined_fail(this%r)
    call view_init(this%base, hb, tmin, tmax, CFL)
    call view_init(this%fine, hf, tmin, tmax, CFL)
  end function kg_refined_lines_method_constructor

  subroutine refined_fail(r)
    type(c_ptr), intent(in) :: r
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    call c_f_pointer(kg_refined_last_error(r), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    print*, 'kg engine error (refinement): ', transfer(msg(1:n), repeat(' ', n))
    stop
  end subroutine refined_fail

  subroutine krl_step(this, nsteps)
    class(kg_refined_lines_method), intent(inout) :: this
    integer, intent(in) :: nsteps
    if (kg_refined_step(this%r, int(nsteps, c_int32_t)) /= KG_OK) call refined_fail(this%r)
    call sync_time(this%base)
    call sync_time(this%fine)
  end subroutine krl_step

  ! frees both levels with the pair
  subroutine krl_destroy(this)
    class(kg_refined_lines_method), intent(inout) :: this
    integer(c_int32_t) :: st
    if (.not. c_associated(this%r)) return
    st = kg_refined_destroy(this%r)
    this%r = c_null_ptr
    this%base%h = c_null_ptr
    this%fine%h = c_null_ptr
  end subroutine krl_destroy

  subroutine klm_upload(this, u)
    class(kg_lines_method), intent(inout) :: this
    complex(kind=8), intent(in), target :: u(:,:,:,:)
    call kg_check(kg_upload_state(this%h, c_loc(u), KG_LAYOUT_FORTRAN), this%h)
  end subroutine klm_upload

  subroutine klm_download(this, u)
    class(kg_lines_method), intent(inout) :: this
    complex(kind=8), intent(inout), target :: u(:,:,:,:)
    call kg_check(kg_download_state(this%h, c_loc(u), KG_LAYOUT_FORTRAN), this%h)
  end subroutine klm_download

  subroutine klm_boundary(this)
    class(kg_lines_method), intent(inout) :: this
    call kg_check(kg_fill_ghosts(this%h), this%h)
  end subroutine klm_boundary

  subroutine sync_time(this)
    class(kg_lines_method), intent(inout) :: this
    real(c_double) :: t, dt
    call kg_check(kg_get_time(this%h, t, dt), this%h)
    this%t = t
  end subroutine sync_time

  ! Several grids advanced together (no counterpart in the reference, which
  ! runs one grid per process): nsteps steps of their common scheme, one
  ! kernel launch per stage over all members; each ends as if stepped alone.
  subroutine kg_lines_method_batch_step(members, nsteps)
    type(kg_lines_method), intent(inout) :: members(:)
    integer, intent(in) :: nsteps
    type(c_ptr), allocatable :: hs(:)
    integer :: q
    allocate(hs(size(members)))
    do q = 1, size(members)
      hs(q) = members(q)%h
    end do
    call kg_check(kg_batch_step(hs, int(size(members), c_int32_t), int(nsteps, c_int32_t)), members(1)%h)
    do q = 1, size(members)
      call sync_time(members(q))
    end do
  end subroutine kg_lines_method_batch_step

  subroutine klm_pirk2(this)
    class(kg_lines_method), intent(inout) :: this
    call kg_check(kg_pirk2(this%h), this%h)
    call sync_time(this)
  end subroutine klm_pirk2

  subroutine klm_pirk3(this)
    class(kg_lines_method), intent(inout) :: this
    call kg_check(kg_pirk3(this%h), this%h)
    call sync_time(this)
  end subroutine klm_pirk3

  subroutine klm_rk3(this)
    class(kg_lines_method), intent(inout) :: this
    call kg_check(kg_rk3(this%h), this%h)
    call sync_time(this)
  end subroutine klm_rk3

  subroutine klm_rk4(this)
    class(kg_lines_method), intent(inout) :: this
    call kg_check(kg_rk4(this%h), this%h)
    call sync_time(this)
  end subroutine klm_rk4

  subroutine klm_step(this, n)
    class(kg_lines_method), intent(inout) :: this
    integer, intent(in) :: n
    call kg_check(kg_step(this%h, int(n, c_int32_t)), this%h)
    call sync_time(this)
  end subroutine klm_step

  ! initial_conditions (main/initial_conditions.f90:1-220); the globals
  ! initial_conditions_type, mu and sigma become arguments, with the Cartesian
  ! grid size of the random types (the reference fixes 256) and their seed
  subroutine klm_initial_conditions(this, initial_conditions_type, mu, sigma, ncart, seed, amp)
    class(kg_lines_method), intent(inout) :: this
    character(len=*), intent(in) :: initial_conditions_type
    real(kind=8), intent(in) :: mu, sigma
    integer, intent(in) :: ncart
    integer(kind=8), intent(in) :: seed
    real(kind=8), intent(out) :: amp
    type(kg_ic_desc) :: d
    character(kind=c_char), pointer :: msg(:)
    integer(c_int32_t) :: st
    integer :: n
    select case (trim(initial_conditions_type))
    case ('spherical_shell'); d%type = KG_IC_SPHERICAL_SHELL
    case ('random');          d%type = KG_IC_RANDOM
    case ('random_shell');    d%type = KG_IC_RANDOM_SHELL
    case default
      print*, 'Intial condition is not implemented yet'
      stop
    end select
    d%ncart = int(ncart, c_int32_t); d%mu = mu; d%sigma = sigma; d%p0 = 1.0d0
    d%seed = int(seed, c_int64_t); d%uniforms = c_null_ptr
    st = kg_initial_data(this%h, d, amp)
    if (st /= KG_OK) then
      call c_f_pointer(kg_ic_last_error(), msg, [512])
      n = 0
      do while (n < 512)
        if (msg(n + 1) == c_null_char) exit
        n = n + 1
      end do
      print*, 'kg engine error ', st, ': ', transfer(msg(1:n), repeat(' ', n))
      stop
    end if
  end subroutine klm_initial_conditions

  subroutine klm_destroy(this)
    class(kg_lines_method), intent(inout) :: this
    integer(c_int32_t) :: st
    if (c_associated(this%h)) st = kg_destroy(this%h)
    this%h = c_null_ptr
  end subroutine klm_destroy

end module kg_hip_mod
