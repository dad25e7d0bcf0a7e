! karma3d.f90 -- the Karma model program of the MI355X build.
!
! The run of the reference's program (karma/karma3d.f90:70-314): read the
! namelist karma.nml (group and entry names as the reference's, each entry
! optional, the reference's defaults otherwise), print the dimensional
! scales, set the 1/8 seed at the corner, then advance the coupled phi/U step
! on the GPU (pf_karma_run) up to each output step -- every nout steps and
! the last one -- where it locates the tip on the x axis (the largest i with
! phi(i,1,1) > 0) and on the diagonal (phi(i,i,i) > 0), prints the progress
! line, writes output/karma_<it, I7.7>.dat (Tecplot POINT: X, Y, Z in m,
! (1+phi)/2, TP and c) and stops once the x-axis tip passes tip_stop*N.
! Console lines and files use the reference's formats.  nout <= 0 writes
! only at the last step.
program karma3d
  use, intrinsic :: iso_c_binding
  use pf_capi, only: pf_check
  use pf_karma_capi
  implicit none

  integer :: N = 200
  double precision :: lambda = 4.d0, eps4 = 0.02d0, dx = 0.8d0, dt = 0.008d0, R0 = 8.d0
  double precision :: Fnoise = 0.d0
  integer :: nsteps = 100000, nout = 2000
  double precision :: tip_stop = 0.9d0
  double precision :: k = 0.14d0, Tm = 933.3d0, mliq = 640.d0, C0 = 0.0196d0
  double precision :: TP = 909.5d0, sigma = 0.093d0, dH = 0.891002d9, Dl = 3.0d-9
  namelist /karma/ N, lambda, eps4, dx, dt, R0, Fnoise, nsteps, nout, tip_stop, &
                   k, Tm, mliq, C0, TP, sigma, dH, Dl

  double precision, parameter :: a1 = 0.8839d0, a2 = 0.6267d0  ! thin-interface constants
  double precision :: D, cl0, Omega, Gam, d0, W, tau0
  double precision, allocatable :: phi(:,:,:), U(:,:,:)
  type(pf_karma_params_t) :: kp
  type(c_ptr) :: s = c_null_ptr
  integer :: it, nxt, u_nml, ios, tip_x, tip_d, i
  integer(c_int32_t) :: it_dev
  integer(8) :: clk0, clk, rate
  character(len=64) :: fname

  open(newunit=u_nml, file='karma.nml', status='old', action='read', iostat=ios)
  if (ios == 0) then
    read(u_nml, nml=karma, iostat=ios)
    if (ios /= 0) stop 'erro lendo karma.nml'
    close(u_nml)
  end if

  ! scales (karma3d.f90:110-116)
  D = a2 * lambda
  cl0 = (Tm - TP) / mliq
  Omega = (cl0 - C0) / ((1.d0 - k) * cl0)
  Gam = sigma * Tm / dH
  d0 = Gam / (mliq * (1.d0 - k) * cl0)
  W = lambda / a1 * d0
  tau0 = a2 * lambda * W * W / Dl

  write(*, '(A)') '### karma3d -- campo de fase quantitativo (Karma 2001 / EFKP 2004) ###'
  write(*, '(A,I0,A,F6.3,A,F8.3,A)') 'N = ', N, '^3,  dx = ', dx, ' W,  dominio = ', &
        N * dx * W * 1.d6, ' um'
  write(*, '(A,F6.3,A,F6.3,A,F6.3,A,F7.3)') 'k = ', k, '  eps4 = ', eps4, '  lambda = ', &
        lambda, '  D = ', D
  write(*, '(A,F8.2,A,F8.5,A,F6.3)') 'TP = ', TP, ' K   c_l0 = ', cl0, '   Omega = ', Omega
  write(*, '(A,ES10.3,A,ES10.3,A,ES10.3,A)') 'd0 = ', d0, ' m   W = ', W, ' m   tau0 = ', &
        tau0, ' s'
  write(*, '(A,ES10.3,A,I0,A,F7.4,A)') 'dt = ', dt * tau0, ' s   (', nsteps, &
        ' passos max)   Fnoise = ', Fnoise, ''
  ! explicit-stability margin of the two diffusion operators
  if (D * dt / (dx * dx) > 0.3d0 / 3.d0 .or. dt / (dx * dx) > 0.4d0 / 3.d0) &
    write(*, '(A)') 'AVISO: dt perto do limite de estabilidade explicita'

  kp%n = N; kp%lambda = lambda; kp%eps4 = eps4; kp%dx = dx; kp%dt = dt; kp%R0 = R0
  kp%Fnoise = Fnoise; kp%k = k; kp%Tm = Tm; kp%mliq = mliq; kp%C0 = C0; kp%TP = TP
  call pf_check(pf_karma_create(kp, s), 'pf_karma_create')
  write(*, '(A)') '### gfx950 HIP device: phi and U resident in HBM ###'
  allocate(phi(N, N, N), U(N, N, N))

  call system_clock(clk0, rate)
  it = 0
  do while (it < nsteps)
    nxt = nsteps
    if (nout > 0) nxt = min(nsteps, (it / nout + 1) * nout)
    call pf_check(pf_karma_run(s, int(nxt - it, c_int32_t)), 'pf_karma_run')
    it = nxt
    call pf_check(pf_karma_download(s, phi, U, it_dev), 'pf_karma_download')
    tip_x = 0
    tip_d = 0
    do i = 1, N
      if (phi(i, 1, 1) > 0.d0) tip_x = i
      if (phi(i, i, i) > 0.d0) tip_d = i
    end do
    call system_clock(clk)
    write(*, '(A,I8,A,F10.1,A,ES10.3,A,F7.1,A,F7.1,A,F6.3,A,F8.1,A)') 'it=', it, '  t=', &
          it * dt, ' tau0 (', it * dt * tau0, ' s)  ponta_eixo=', tip_x * dx, &
          ' W  ponta_diag=', tip_d * dx * sqrt(3.d0), ' W  razao=', &
          tip_x / max(tip_d * sqrt(3.d0), 1.d0), '  wall=', dble(clk - clk0) / dble(rate), ' s'
    flush(6)
    write(fname, '(A,I7.7,A)') 'output/karma_', it, '.dat'
    call escreve(trim(fname))
    if (tip_x > tip_stop * N) then
      write(*, '(A)') '### ponta chegou perto da parede -- fim ###'
      exit
    end if
  end do

  call pf_check(pf_karma_destroy(s), 'pf_karma_destroy')

contains

  ! Tecplot POINT file of the cells 1..N (karma3d.f90:341-360's layout):
  ! positions in m, phase (1+phi)/2, TP, and c recovered from U and phi.
  subroutine escreve(nome)
    character(len=*), intent(in) :: nome
    integer :: uo, a, b, c
    double precision :: cc
    open(newunit=uo, file=nome, status='replace', action='write')
    write(uo, '(A)') 'TITLE="Phase Field (Karma)"'
    write(uo, '(A)') 'variables= "X","Y","Z","PhaseField","Temp","Conc_Solute"'
    write(uo, '(A)') 'ZONE'
    write(uo, '(A,I0,A,I0,A,I0,A)') ' T=" " ,I= ', N, ',J= ', N, ',K= ', N, ',F=POINT'
    do c = 1, N
      do b = 1, N
        do a = 1, N
          cc = cl0 * 0.5d0 * (1.d0 + k - (1.d0 - k) * phi(a, b, c)) * (1.d0 + (1.d0 - k) * U(a, b, c))
          write(uo, '(6(1X,ES15.7E3))') (a - 1) * dx * W, (b - 1) * dx * W, (c - 1) * dx * W, &
                0.5d0 * (1.d0 + phi(a, b, c)), TP, cc
        end do
      end do
    end do
    close(uo)
  end subroutine escreve

end program karma3d
