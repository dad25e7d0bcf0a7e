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
    logical :: opt_fused, opt_snapshots, have
    integer :: opt_it_stop
    double precision :: opt_D
    namelist /pf3d/ opt_precision, opt_noise_seed, opt_n_seeds, opt_seed_rng, opt_fused, &
                    opt_snapshots, opt_it_stop, opt_D

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
    opt_it_stop = it_stop; opt_D = D_thermal
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
    it_stop = opt_it_stop; D_thermal = opt_D

    write(*, '(/,A,/)') '### Dados de Entrada ###'
    write(*, '(A,I5,A,I5,A,I5)') 'Nx x Ny x Nz = ', Nx, ' x ', Ny, ' x ', Nz
    write(*, '(A,F14.2,A,F14.2,A)') 'Tm = ', Tm, ' K   T0 = ', T0, ' K'
    write(*, '(A,F14.2,A,F10.6,A,F10.2,A)') 'TP = ', TP_deck, ' K  (Cl_eq = ', &
          (Tm - TP_deck) / Me, ', TL(C0) = ', Tm - Me * C0, ' K)'
    write(*, '(A,I0,A,I0,A,I0,A,I0)') 'tipo_inicial = ', tipo_inicial, '  n_varreduras = ', &
          n_varreduras, '  h_quintico = ', h_quintico, '  ruido_tipo = ', ruido_tipo
    write(*, '(A,I0,A,L1,A,I0)') 'precision = ', precision, '  fused = ', fused, &
          '  noise seed = ', noise_seed

    ! derived scalars (entrada.f90:159-185)
    Lambda = Thin_Interface * dx
    DELTA_T = pf_time_step(dx, D_liq)
    AAA = pf_noise_scale(AAA_deck, DELTA_T, dx, D_liq, int(ruido_tipo, c_int32_t))
    write(*, '(A,F10.5)') 'AAA efetivo = ', AAA

    prm%nx = Nx; prm%ny = Ny; prm%nz = Nz; prm%it = 0
    prm%n_varreduras = n_varreduras; prm%h_quintico = h_quintico; prm%ruido_tipo = ruido_tipo
    prm%precision = precision; prm%seed = noise_seed
    prm%dx = dx; prm%dy = dy; prm%dz = dz; prm%dt = DELTA_T; prm%D = D_thermal
    prm%delta_H = delta_H; prm%cp_s = Cp_s
    prm%TP = TP_deck
    if (taxa_resfr > 0.d0) prm%TP = T0
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
    write(*, '(A,I0,A,I0,A,I0,A,I0,A)') 'restart.dat read (', Nx, ' x ', Ny, ' x ', Nz, &
          '), shift = ', nshift, ' layers'
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
