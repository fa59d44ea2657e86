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
!   agree); bc=<outflow|inflow|sommerfeld>: the outer boundary (outflow)
program kg_evolve
  use kg_hip_mod
  implicit none
  type(kg_lines_method) :: MoL
  type(kg_lines_method), allocatable :: batch(:)
  complex(kind=8), allocatable :: u(:,:,:,:), v(:,:,:,:)
  character(len=256) :: arg, fin, fout, scheme
  character(len=16) :: bc
  integer :: ia
  integer :: Nx, Ny, Nz, nsteps, l, g, isch, ncart, ops, nb, q
  integer(kind=8) :: seed
  real(kind=8) :: xmin, xmax, m, mu, sigma, amp
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
  do ia = 11, command_argument_count()
    call get_command_argument(ia, arg)
    if (arg(1:6) == 'batch=') read(arg(7:), *) nb
    if (arg(1:3) == 'bc=') bc = trim(arg(4:))
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

  MoL = kg_lines_method_constructor(5, g, xmin, xmax, Nx, Ny, Nz, 0.0d0, 1.0d0, 0.25d0, &
                                    trim(bc), 'Kerr_Schild', 0.0d0, 1.0d0, m, isch, .true., ops)
  allocate(u(5, -g:Nx+g, -g:Ny+g, 0:Nz))
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
