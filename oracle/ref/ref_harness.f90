! ref_harness.f90 -- a driver for the reference's own kernels, used only to
! produce golden vectors (TEST INFRASTRUCTURE).
!
! It is linked against the reference's source/*.f90 compiled from where they
! lie (oracle/ref/Makefile) in place of source/principal.f90, and does what
! principal.f90:93-226 does between output events, then writes raw FP64 state
! so that the C oracle can be compared with the reference bit for bit.
!
! Control: a namelist file harness.nml in the working directory (next to the
! entrada.dat deck the reference's `entrada` reads):
!   mode       'run'    nsteps steps of the time loop; dumps at dump_steps
!              'unit'   one call of `unit_call` on the (optionally loaded)
!                       fields; dumps once
!              'random' pseudo_random / ruido_suave values for tuples.txt
!              'time'   wall-clock of nsteps steps after nwarm warm-up steps
!   unit_call  'temperatura' | 'fase' | 'soluto' | 'diag' | 'auxiliar'
!   init_file  raw FP64 file with the seven fields in the order
!              phi_old phi_new c_old c_new TOLD TNEW DPP (replaces entrada's)
!   it_set, tp_set, m0_set  values forced before a unit call (< -1e30: keep)
! Output: dump_<n>.bin = the seven fields, then 16 scalars:
!   it time TP M0 E0 W0 Cs_eq Cl_eq AAA DELTA_T L2 max min sum_c free_energy
!   n_varreduras
program ref_harness
  use variaveis
  use omp_lib
  implicit none
  character(len=16) :: mode, unit_call
  character(len=256) :: init_file
  integer :: nsteps, nwarm, dump_steps(64), ndump, n_sw, u, ios, nt, q
  integer :: ti, tj, tk, tit, tsalt
  double precision :: it_set, tp_set, m0_set, t0w, t1w
  double precision :: L2_phi, phi_max, phi_min, solute_total, free_energy
  namelist /harness/ mode, unit_call, init_file, nsteps, nwarm, dump_steps, &
                     it_set, tp_set, m0_set

  mode = 'run'; unit_call = ''; init_file = ''; nsteps = 1; nwarm = 0
  dump_steps = -1; it_set = -1.d31; tp_set = -1.d31; m0_set = -1.d31
  open(newunit=u, file='harness.nml', status='old', action='read')
  read(u, nml=harness)
  close(u)
  ndump = count(dump_steps >= 0)

  if (trim(mode) == 'random') then
    open(newunit=u, file='tuples.txt', status='old', action='read')
    read(u,*) nt
    open(newunit=q, file='random.bin', access='stream', form='unformatted', status='replace')
    do n_sw = 1, nt
      read(u,*) ti, tj, tk, tit, tsalt
      ran = pseudo_random(ti, tj, tk, tit, tsalt)
      write(q) pseudo_random(ti, tj, tk, tit, tsalt), ruido_suave(ti, tj, tk, tit), dble(ran)
    end do
    close(q)
    close(u)
    stop
  end if

  call entrada
  call W0_E0
  time = 0.0
  E_R = 0

  if (len_trim(init_file) > 0) then
    open(newunit=u, file=trim(init_file), access='stream', form='unformatted', &
         status='old', action='read')
    read(u) phase_field_old, phase_field_new, conc_soluto_old, conc_soluto_new, &
            TOLD, TNEW, DPP
    close(u)
  end if

  if (trim(mode) == 'unit') then
    if (it_set > -1.d30) it = nint(it_set)
    if (tp_set > -1.d30) TP = tp_set
    if (m0_set > -1.d30) M0 = m0_set
    L2_phi = 0; phi_max = 0; phi_min = 0; solute_total = 0; free_energy = 0
    select case (trim(unit_call))
    case ('temperatura')
      call temperatura
    case ('fase')
      call equacao_campo_fase
    case ('soluto')
      call concentracao_soluto
    case ('diag')
      call diagnosticos(L2_phi, phi_max, phi_min, solute_total, free_energy)
    case ('auxiliar')
      Cs_eq = ((Tm-TP)*Ke)/Me
      Cl_eq = (Tm-TP)/Me
      call auxiliar
    end select
    call dump(0)
    stop
  end if

  if (trim(mode) == 'time') then
    t0w = 0
    do it = 1, nwarm + nsteps
      if (it == nwarm + 1) t0w = omp_get_wtime()
      call one_step()
    end do
    t1w = omp_get_wtime()
    open(newunit=u, file='time.txt', status='replace')
    write(u,'(ES25.17E3,1X,I0,1X,I0)') t1w - t0w, nsteps, omp_get_max_threads()
    close(u)
    stop
  end if

  do it = 1, nsteps
    call one_step()
    if (any(dump_steps(1:max(ndump,1)) == it)) then
      call diagnosticos(L2_phi, phi_max, phi_min, solute_total, free_energy)
      call dump(it)
    end if
    phase_field_old = phase_field_new
    TOLD = TNEW
    conc_soluto_old = conc_soluto_new
  end do
  if (any(dump_steps(1:max(ndump,1)) == 0)) call dump(0)

contains

  subroutine one_step()
    time = delta_T + time
    call temperatura
    if (taxa_resfr > 0.d0) TP = T0 - taxa_resfr*DELTA_T*IT
    do n_sw = 1, n_varreduras
      if (n_sw > 1) phase_field_old = phase_field_new
      call equacao_campo_fase
    end do
    call concentracao_soluto
    call auxiliar
    if (trim(mode) == 'time') then
      phase_field_old = phase_field_new
      TOLD = TNEW
      conc_soluto_old = conc_soluto_new
    end if
  end subroutine one_step

  subroutine dump(n)
    integer, intent(in) :: n
    character(len=64) :: fn
    integer :: v
    write(fn,'(A,I0,A)') 'dump_', n, '.bin'
    open(newunit=v, file=trim(fn), access='stream', form='unformatted', status='replace')
    write(v) phase_field_old, phase_field_new, conc_soluto_old, conc_soluto_new, &
             TOLD, TNEW, DPP
    write(v) dble(it), dble(time), TP, M0, E0, W0, Cs_eq, Cl_eq, AAA, DELTA_T, &
             L2_phi, phi_max, phi_min, solute_total, free_energy, dble(n_varreduras)
    close(v)
  end subroutine dump

end program ref_harness
