! kg_evolve.f90 -- a Fortran driver in the shape of the reference's evolve
! loop (main/evolve.f90:22-26), with the step done by the MI355X engine.
!
! usage: kg_evolve Nx Ny Nz xmin xmax m scheme nsteps in.bin out.bin
!   scheme: pirk2 | pirk3 | rk3 | rk4, or rk3_fd4 | rk4_fd4 for the 4th-order stencils
!   in.bin/out.bin: u(5, -g:Nx+g, -g:Ny+g, 0:Nz) complex(8), raw stream
!   in.bin may instead be ic=<type>,<mu>,<sigma>,<ncart>,<seed>: the initial
!   data of main/initial_conditions.f90 built by the engine
!   optional further arguments: batch=<K>: K copies of the grid advanced
!   together in batched launches (the driver stops with an error unless all K
!   agree); bc=<outflow|inflow|sommerfeld>: the outer boundary (outflow);
!   refine=<r_patch>,<fine_in>,<fine_out>: a 2:1 radial patch on
!   [xmin, r_patch] (buffer 16): fine_in/fine_out hold the fine level's u
!   (same layout, its own Nx), in.bin/out.bin the base level's
program kg_evolve
  use kg_hip_mod
  implicit none
  type(kg_lines_method) :: MoL
  type(kg_lines_method), allocatable :: batch(:)
  type(kg_refined_lines_method) :: pair
  complex(kind=8), allocatable :: u(:,:,:,:), v(:,:,:,:), uf(:,:,:,:)
  character(len=256) :: arg, fin, fout, scheme, ffin, ffout
  character(len=16) :: bc
  integer :: ia, ic
  integer :: Nx, Ny, Nz, nsteps, l, g, isch, ncart, ops, nb, q
  integer(kind=8) :: seed
  real(kind=8) :: xmin, xmax, m, mu, sigma, amp, r_patch
  character(len=64) :: ictype

  g = 2
  call get_command_argument(1, arg); read(arg, *) Nx
  call get_command_argument(2, arg); read(arg, *) Ny
  call get_command_argument(3, arg); read(arg, *) Nz
  call get_command_argument(4, arg); read(arg, *) xmin
  call get_command_argument(5, arg); read(arg, *) xmax
  call get_command_argument(6, arg); read(arg, *) m
  call get_command_argument(7, scheme)
  call get_command_argument(8, arg); read(arg, *) nsteps
  call get_command_argument(9, fin)
  call get_command_argument(10, fout)
  nb = 1
  bc = 'outflow'
  r_patch = -1.0d0
  do ia = 11, command_argument_count()
    call get_command_argument(ia, arg)
    if (arg(1:6) == 'batch=') read(arg(7:), *) nb
    if (arg(1:3) == 'bc=') bc = trim(arg(4:))
    if (arg(1:7) == 'refine=') then
      arg = arg(8:)
      ic = index(arg, ',')
      read(arg(1:ic-1), *) r_patch
      arg = arg(ic+1:)
      ic = index(arg, ',')
      ffin = arg(1:ic-1)
      ffout = arg(ic+1:)
    end if
  end do

  ops = KG_OPS_REF2
  select case (trim(scheme))
  case ('pirk2'); isch = KG_SCHEME_PIRK2
  case ('pirk3'); isch = KG_SCHEME_PIRK3
  case ('rk3');   isch = KG_SCHEME_RK3
  case ('rk3_fd4'); isch = KG_SCHEME_RK3; ops = KG_OPS_FD4
  case ('rk4_fd4'); isch = KG_SCHEME_RK4; ops = KG_OPS_FD4
  case default;   isch = KG_SCHEME_RK4
  end select

  allocate(u(5, -g:Nx+g, -g:Ny+g, 0:Nz))
  if (r_patch > 0.0d0) then
    ! the base grid with its refined patch, Berger-Oliger sub-stepping
    pair = kg_refined_lines_method_constructor(5, g, xmin, xmax, Nx, Ny, Nz, 0.0d0, 1.0d0, 0.25d0, &
                                               trim(bc), 'Kerr_Schild', 0.0d0, 1.0d0, m, isch, .true., &
                                               r_patch, 16, ops)
    allocate(uf(5, -g:pair%fine%Nx+g, -g:Ny+g, 0:Nz))
    open(unit=10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) u
    close(10)
    open(unit=10, file=trim(ffin), access='stream', form='unformatted', status='old')
    read(10) uf
    close(10)
    call pair%base%upload(u)
    call pair%fine%upload(uf)
    call pair%step(nsteps)
    call pair%base%download(u)
    call pair%fine%download(uf)
    open(unit=11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) u
    close(11)
    open(unit=11, file=trim(ffout), access='stream', form='unformatted', status='replace')
    write(11) uf
    close(11)
    print '(A,I8,A,I6,A,ES24.16,A,ES24.16)', 'refined steps ', nsteps, ' fine Nx ', pair%fine%Nx, ' t ', &
          pair%base%t, ' dt ', pair%base%dt
    call pair%destroy()
    stop
  end if
  MoL = kg_lines_method_constructor(5, g, xmin, xmax, Nx, Ny, Nz, 0.0d0, 1.0d0, 0.25d0, &
                                    trim(bc), 'Kerr_Schild', 0.0d0, 1.0d0, m, isch, .true., ops)
  if (fin(1:3) == 'ic=') then
    read(fin(4:), *) ictype, mu, sigma, ncart, seed
    call MoL%initial_conditions(trim(ictype), mu, sigma, ncart, seed, amp)
    print '(A,A,A,ES24.16)', 'initial data ', trim(ictype), ' Amp ', amp
  else
    open(unit=10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) u
    close(10)
    call MoL%upload(u)
  end if

  if (nb > 1) then
    allocate(batch(nb))
    batch(1) = MoL
    do q = 2, nb
      batch(q) = kg_lines_method_constructor(5, g, xmin, xmax, Nx, Ny, Nz, 0.0d0, 1.0d0, 0.25d0, &
                                             trim(bc), 'Kerr_Schild', 0.0d0, 1.0d0, m, isch, .true., ops)
      if (fin(1:3) == 'ic=') then
        call batch(q)%initial_conditions(trim(ictype), mu, sigma, ncart, seed, amp)
      else
        call batch(q)%upload(u)
      end if
    end do
    call kg_lines_method_batch_step(batch, nsteps)
    MoL = batch(1)
    call MoL%download(u)
    allocate(v, mold=u)
    do q = 2, nb
      call batch(q)%download(v)
      if (any(v /= u)) then
        print '(A,I4,A)', 'batch member ', q, ' differs from member 1'
        stop 1
      end if
      call batch(q)%destroy()
    end do
    print '(A,I4,A)', 'batch of ', nb, ' members agree'
  else
    do l = 1, nsteps
      select case (isch)
      case (KG_SCHEME_PIRK2); call MoL%pirk2()
      case (KG_SCHEME_PIRK3); call MoL%pirk3()
      case (KG_SCHEME_RK3);   call MoL%rk3()
      case default;    call MoL%rk4()
      end select
    end do
    call MoL%download(u)
  end if

  open(unit=11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) u
  close(11)
  print '(A,I8,A,ES24.16,A,ES24.16)', 'steps ', nsteps, ' t ', MoL%t, ' dt ', MoL%dt
  call MoL%destroy()
end program kg_evolve
