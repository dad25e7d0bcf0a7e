! pf_io.f90 -- the driver's host I/O: the entrada.dat deck (same positional
! format and optional records as source/entrada.f90:17-89), the derived
! scalars (entrada.f90:159-185), the initial fields (entrada.f90:226-262 plus
! the seeded modes), the restart reader with the moving frame
! (entrada.f90:284-321) and the Tecplot POINT writer (saida.f90:1-66).
module pf_io
  use, intrinsic :: iso_c_binding
  use pf_capi
  use pf_state
  implicit none

contains

  ! Read entrada.dat, set the scalars, allocate and initialise the host fields.
  subroutine entrada()
    integer :: u, ios
    integer :: opt_precision, opt_n_seeds
    integer(c_int64_t) :: opt_noise_seed, opt_seed_rng
    logical :: opt_fused, opt_snapshots, opt_w0e0_trace, have
    integer :: opt_it_stop
    double precision :: opt_D
    namelist /pf3d/ opt_precision, opt_noise_seed, opt_n_seeds, opt_seed_rng, opt_fused, &
                    opt_snapshots, opt_it_stop, opt_D, opt_w0e0_trace

    open(newunit=u, file='entrada.dat', status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*, '(A)') 'pf3d: entrada.dat not found'
      error stop 1
    end if
    read(u, *); read(u, *); read(u, *)
    read(u, *) Tm;      read(u, *) T0;      read(u, *) delta_H
    read(u, *) E0_deck; read(u, *) W0_deck; read(u, *) M0_deck
    read(u, *) D_s;     read(u, *) D_l;     read(u, *) Cp_s;    read(u, *) Cp_l
    read(u, *) Nx;      read(u, *) Ny;      read(u, *) Nz
    read(u, *) dy;      read(u, *) dx;      read(u, *) dz
    read(u, *) small;   read(u, *) ANI;     read(u, *) it_max
    read(u, *) AAA_deck; read(u, *) Teta0;  read(u, *) DELTA_EPSLON
    read(u, *) DELTA_T_deck; read(u, *) D_sol; read(u, *) D_liq
    read(u, *) Ke;      read(u, *) Me;      read(u, *) Vm;      read(u, *) R
    read(u, *) C0;      read(u, *) Thin_Interface
    read(u, *) Sigma;   read(u, *) Beta;    read(u, *) ajuste
    read(u, *) stoptime; read(u, *) WriteFileD_timeCounter
    ! optional records, with the reference's defaults
    read(u, *, iostat=ios) TP_deck
    if (ios /= 0) then
      TP_deck = 919.d0
      write(*, '(A)') 'AVISO: linha TP ausente no entrada.dat -- usando TP = 919.0 K'
    end if
    read(u, *, iostat=ios) tipo_inicial, R_semente
    if (ios /= 0) then
      tipo_inicial = 0; R_semente = 0.d0
    end if
    read(u, *, iostat=ios) taxa_resfr, n_varreduras
    if (ios /= 0) then
      taxa_resfr = 0.d0; n_varreduras = 1
    end if
    n_varreduras = max(1, n_varreduras)
    read(u, *, iostat=ios) h_quintico, ruido_tipo
    if (ios /= 0) then
      h_quintico = 1; ruido_tipo = 2
    end if
    close(u)

    ! options of this build: pf3d.nml next to the deck (optional)
    opt_precision = precision; opt_noise_seed = noise_seed; opt_n_seeds = n_seeds
    opt_seed_rng = seed_rng; opt_fused = fused; opt_snapshots = write_snapshots
    opt_it_stop = it_stop; opt_D = D_thermal; opt_w0e0_trace = w0e0_trace
    inquire(file='pf3d.nml', exist=have)
    if (have) then
      open(newunit=u, file='pf3d.nml', status='old', action='read')
      read(u, nml=pf3d, iostat=ios)
      close(u)
      if (ios /= 0) then
        write(*, '(A)') 'pf3d: error reading pf3d.nml'
        error stop 1
      end if
    end if
    precision = opt_precision; noise_seed = opt_noise_seed; n_seeds = opt_n_seeds
    seed_rng = opt_seed_rng; fused = opt_fused; write_snapshots = opt_snapshots
    it_stop = opt_it_stop; D_thermal = opt_D; w0e0_trace = opt_w0e0_trace

    if (taxa_resfr > 0.d0) TP_deck = T0  ! entrada.f90:83: cooling starts at T0
    call echo_deck()

    ! derived scalars (entrada.f90:159-185)
    Lambda = Thin_Interface * dx
    DELTA_T = pf_time_step(dx, D_liq)
    AAA = pf_noise_scale(AAA_deck, DELTA_T, dx, D_liq, int(ruido_tipo, c_int32_t))
    if (ruido_tipo == 0) then
      write(*, "('AAA efetivo (escalado com dt) = ',F10.5)") AAA
    else
      write(*, "('AAA efetivo (escalado com sqrt(dt)) = ',F10.5)") AAA
    end if
    write(*, "('h(phi) quintico = ',I0,'   tipo de ruido = ',I0)") h_quintico, ruido_tipo
    write(*, '(A,I0,A,L1,A,I0,A,L1)') 'pf3d: precision = ', precision, '  fused = ', fused, &
          '  noise seed = ', noise_seed, '  W0_E0 trace = ', w0e0_trace

    prm%nx = Nx; prm%ny = Ny; prm%nz = Nz; prm%it = 0
    prm%n_varreduras = n_varreduras; prm%h_quintico = h_quintico; prm%ruido_tipo = ruido_tipo
    prm%precision = precision; prm%seed = noise_seed
    prm%dx = dx; prm%dy = dy; prm%dz = dz; prm%dt = DELTA_T; prm%D = D_thermal
    prm%delta_H = delta_H; prm%cp_s = Cp_s
    prm%TP = TP_deck
    prm%Tm = Tm; prm%Me = Me; prm%Vm = Vm; prm%R = R
    prm%M0 = M0_deck; prm%delta_eps = DELTA_EPSLON; prm%AAA = AAA
    prm%D_liq = D_liq; prm%D_sol = D_sol; prm%Ke = Ke
    call pf_equilibrium(prm%TP, prm%Tm, prm%Ke, prm%Me, prm%Cs_eq, prm%Cl_eq)
    lp%T0 = T0; lp%taxa_resfr = taxa_resfr
    lp%Sigma = Sigma; lp%Beta = Beta; lp%ajuste = ajuste; lp%time = 0.0

    allocate(phase_field_old(Nx, Ny, Nz), phase_field_new(Nx, Ny, Nz))
    allocate(conc_soluto_old(Nx, Ny, Nz), conc_soluto_new(Nx, Ny, Nz))
    allocate(TOLD(Nx, Ny, Nz), TNEW(Nx, Ny, Nz), DPP(Nx, Ny, Nz))
    call pf_check(pf_initial_fields(int(Nx, c_int32_t), int(Ny, c_int32_t), int(Nz, c_int32_t), &
                  int(merge(0, tipo_inicial, tipo_inicial == 4), c_int32_t), R_semente, T0, C0, &
                  Ke, int(n_seeds, c_int32_t), seed_rng, phase_field_old, phase_field_new, &
                  conc_soluto_old, conc_soluto_new, TOLD, TNEW, DPP), 'initial fields')
    if (tipo_inicial == 4) call ler_restart()
  end subroutine entrada

  ! The input echo of entrada.f90:92-148: the deck's values as read (E0,
  ! w0, DELTA_t, AAA from the deck, before W0_E0 and the rescales), in the
  ! reference's formats, so that logs compare line by line.  Modes 5 and 6
  ! (bottom seeds) are this build's and have lines of their own.
  subroutine echo_deck()
    write(*, "(/,'### Dados de Entrada ###',/)")
    write(*, "('Tm = ',F14.2,' K')") Tm
    write(*, "('T0 = ',F14.2,' K')") T0
    write(*, "('delta_H = ',E14.6,' J/m^3')") delta_H
    write(*, "('E0 = ',E14.6,' -')") E0_deck
    write(*, "('w0 = ',E14.6,' -')") W0_deck
    write(*, "('M0 = ',E14.6,' -')") M0_deck
    write(*, "('D_s = ',E14.6,' m^2/s')") D_s
    write(*, "('D_l = ',E14.6,' m^2/s')") D_l
    write(*, "('Cp_s = ',E14.6,' J/kg.K')") Cp_s
    write(*, "('Cp_l = ',E14.6,' J/kg.K')") Cp_l
    write(*, "('Nx = ',I5,' -')") Nx
    write(*, "('Ny = ',I5,' -')") Ny
    write(*, "('Nz = ',I5,' -')") Nz
    write(*, "('dy = ',E14.6,' m')") dy
    write(*, "('dx = ',E14.6,' m')") dx
    write(*, "('dz = ',E14.6,' m')") dz
    write(*, "('small = ',E14.6,' -')") small
    write(*, "('ANI = ',E14.6,' -')") ANI
    write(*, "('it_max = ',I10,' -')") it_max
    write(*, "('ANI = ',E14.6,' -')") ANI
    write(*, "('AAA = ',F14.4,' -')") AAA_deck
    write(*, "('Theta0_inic = ',F14.4,' -')") Teta0
    write(*, "('DELTA_EPSLON = ',F14.4,' -')") DELTA_EPSLON
    write(*, "('DELTA_t = ',E14.4,' s')") DELTA_T_deck
    write(*, "('D_sol = ',E14.4,' m^2/s')") D_sol
    write(*, "('D_liq = ',E14.4,' m^2/s')") D_liq
    write(*, "('Ke = ',F14.4,' -')") Ke
    write(*, "('Me = ',F14.4,' K/mol')") Me
    write(*, "('Mv = ',E14.4,' m^3/mol')") Vm
    write(*, "('R = ',F14.5,' J/mol K')") R
    write(*, "('C0 = ',F14.5,' molar fraction')") C0
    write(*, "('Thin_Interface = ',F14.3,' molar fraction')") Thin_Interface
    write(*, "('Sigma = ',F14.3,' J/m^2')") Sigma
    write(*, "('ajuste',F14.3,' -')") ajuste
    write(*, "('stoptime',F14.3,' s')") stoptime
    write(*, "('WriteFileD_timeCounter',I8,' -')") WriteFileD_timeCounter
    write(*, "('TP = ',F14.2,' K  (Cl_eq = ',F10.6,', TL(C0) = ',F10.2,' K)')") &
          TP_deck, (Tm - TP_deck) / Me, Tm - Me * C0
    if (taxa_resfr > 0.d0) write(*, "('Resfriamento continuo: TP = T0 - ',F8.3," // &
                                 "' K/s * t  (T0 = ',F8.2,' K)')") taxa_resfr, T0
    write(*, "('Varreduras da fase por passo = ',I0)") n_varreduras
    select case (tipo_inicial)
    case (4)
      write(*, "('Condicao inicial: continuacao a partir de restart.dat')")
    case (3)
      write(*, "('Condicao inicial: frente plana na base (k<=4), crescimento em +z')")
    case (2)
      write(*, "('Condicao inicial: 1/8 de semente no canto (3,3,3), raio = ',F6.1,' celulas')") &
            R_semente
    case (1)
      write(*, "('Condicao inicial: semente esferica central, raio = ',F6.1,' celulas')") R_semente
    case (5)
      write(*, "('Condicao inicial: semente esferica no centro da base, raio = ',F6.1," // &
               "' celulas')") R_semente
    case (6)
      write(*, "('Condicao inicial: ',I0,' sementes na base, raio = ',F6.1,' celulas')") &
            n_seeds, R_semente
    case default
      write(*, "('Condicao inicial: frente plana (j<=4)')")
    end select
    write(*, "(/,'### Fim Leitura de Entrada de Dados ###',/)")
  end subroutine echo_deck

  ! tipo_inicial = 4: continue from restart.dat, a Tecplot POINT snapshot of
  ! this grid (4 header records, then one record "x y z phi T c" per cell, x
  ! fastest, then y, then z).  The whole file is read into buffers first;
  ! the moving frame then drops the nshift = nint(R_semente) bottom layers
  ! (file layer k becomes layer k - nshift) and makes the nshift top layers
  ! liquid, each column with the concentration of the file's top layer.
  ! T comes from the deck; c_new equals c_old afterwards.
  subroutine ler_restart()
    double precision, allocatable :: phi_f(:,:,:), c_f(:,:,:)
    double precision :: rec(6)
    integer :: u, ios, nshift, q, ncell, top
    nshift = max(0, nint(R_semente))
    if (nshift >= Nz) call restart_fail('the shift is not below Nz')
    open(newunit=u, file='restart.dat', status='old', action='read', iostat=ios)
    if (ios /= 0) call restart_fail('restart.dat not found')
    read(u, '(3/)', iostat=ios)  ! the title, variables and zone records
    if (ios /= 0) call restart_fail('restart.dat has no header')
    allocate(phi_f(Nx, Ny, Nz), c_f(Nx, Ny, Nz))
    ncell = Nx * Ny * Nz
    do q = 0, ncell - 1
      read(u, *, iostat=ios) rec
      if (ios /= 0) call restart_fail('restart.dat shorter than the grid')
      phi_f(1 + mod(q, Nx), 1 + mod(q / Nx, Ny), 1 + q / (Nx * Ny)) = rec(4)
      c_f(1 + mod(q, Nx), 1 + mod(q / Nx, Ny), 1 + q / (Nx * Ny)) = rec(6)
    end do
    close(u)
    top = Nz - nshift  ! layers 1..top come from the file
    phase_field_old(:, :, 1:top) = phi_f(:, :, nshift + 1:Nz)
    conc_soluto_old(:, :, 1:top) = c_f(:, :, nshift + 1:Nz)
    do q = top + 1, Nz
      phase_field_old(:, :, q) = 0.d0
      conc_soluto_old(:, :, q) = c_f(:, :, Nz)
    end do
    conc_soluto_new = conc_soluto_old
    deallocate(phi_f, c_f)
    write(*, "('Condicao inicial: restart.dat lido (',I0,' x ',I0,' x ',I0," // &
             "'), deslocamento = ',I0,' camadas')") Nx, Ny, Nz, nshift
  end subroutine ler_restart

  subroutine restart_fail(msg)
    character(len=*), intent(in) :: msg
    write(*, '(A,A)') 'pf3d: tipo_inicial = 4: ', msg
    error stop 1
  end subroutine restart_fail

  ! Tecplot POINT snapshot of phi_new, TP and c_new in ./output/, named
  ! saida_<IT (I6)>_time_<time (f14.7)>s.dat like the reference's files.
  subroutine saida()
    character(len=10) :: t_it
    character(len=14) :: t_time
    character(len=256) :: fname
    integer :: u, i, j, k
    write(t_it, '(I6)') it
    write(t_time, '(f14.7)') lp%time
    fname = './output/saida_' // trim(t_it) // '_time_' // trim(t_time) // 's.dat'
    open(newunit=u, file=trim(fname), status='unknown', action='write')
    write(u, *) 'TITLE="Phase Field"'
    write(u, '(A)') 'variables= "X","Y","Z","PhaseField","Temp","Conc_Solute"'
    write(u, *) 'ZONE'
    write(u, '(1X,A,I4,A,I4,A,I4,A,A5)') 'T=" " ,I=', Nx, ',J=', Ny, ',K=', Nz, ',F=', 'POINT'
    do k = 1, Nz
      do j = 1, Ny
        do i = 1, Nx
          write(u, '(6(1X,ES15.7E3))') (i - 1) * dx, (j - 1) * dy, (k - 1) * dz, &
                phase_field_new(i, j, k), prm%TP, conc_soluto_new(i, j, k)
        end do
      end do
    end do
    write(u, *) '       '
    write(u, *) '       '
    close(u)
  end subroutine saida

end module pf_io
