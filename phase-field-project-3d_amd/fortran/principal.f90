! principal.f90 -- the Fortran driver of the MI355X build.
!
! Same run as the reference's program (source/principal.f90:32-277): read
! entrada.dat, find W0/E0, put the seven fields on the device once, advance
! the coupled step, and at every output event (it = N+1, 2N+1, ...) print the
! diagnostics, write the Tecplot snapshot and the timing block; stop on
! stoptime (REAL(4) time) or it_max.
!
! Two ways to advance the fields:
!   fused = .false.  the reference's loop body, one drop-in subroutine call
!                    per kernel (temperatura, equacao_campo_fase x n_var,
!                    concentracao_soluto, auxiliar) with buffer swaps for
!                    the old<-new copies;
!   fused = .true.   (default) pf_run_steps between output events: the same
!                    step, fused into two HBM sweeps, with the per-sweep probe
!                    lines printed after each chunk.
program principal
  use, intrinsic :: iso_c_binding
  use pf_capi
  use pf_state
  use pf_io
  implicit none
  external :: temperatura, equacao_campo_fase, concentracao_soluto, auxiliar, diagnosticos, W0_E0
  integer, parameter :: iu_log = 77
  character(len=*), parameter :: perf_log_file = 'performance_log.txt'
  integer :: n_sw, n, q, nstop, n_outputs, it_prev_output
  integer(8) :: clk0, clk_prev, clk_rate
  double precision, allocatable, target :: probes(:)
  real :: tf
  character(len=*), parameter :: fmt_diag = "('L2(phi)==>',e14.6,'  phi_max==>',e14.6," // &
      "'  phi_min==>',e14.6,'  solute_total==>',e14.6,'  free_energy==>',e14.6)"

  call entrada
  call W0_E0
  call pf_check(pf_init(prm, phase_field_old, phase_field_new, conc_soluto_old, &
                        conc_soluto_new, TOLD, TNEW, DPP, ctx), 'pf_init')
  write(*, '(/,A,/)') '### gfx950 HIP device: fields resident in HBM ###'

  call system_clock(clk0, clk_rate)
  clk_prev = clk0
  n_outputs = 0
  it_prev_output = 0
  open(iu_log, file=perf_log_file, status='replace', action='write')
  write(iu_log, '(A)') '###########################################################'
  write(iu_log, '(A)') '###  Phase-field 3D -- performance / results log        ###'
  write(iu_log, '(A)') '###########################################################'
  write(iu_log, '(A,I0,A,I0,A,I0)') 'Grid  Nx x Ny x Nz     : ', Nx, ' x ', Ny, ' x ', Nz
  write(iu_log, '(A)') 'Execution              : GPU (HIP, gfx950)'
  write(iu_log, '(A)') ''
  close(iu_log)

  it = 0
  E_R = 0
  do
    if (it >= it_max) exit
    if (it_stop > 0 .and. it >= it_stop) exit
    if (fused) then
      ! steps up to and including the next output-event step, the stop step
      ! or the step limit
      n = WriteFileD_timeCounter - E_R + 1
      if (E_R >= WriteFileD_timeCounter) n = 1
      n = min(n, it_max - it)
      if (it_stop > 0) n = min(n, it_stop - it)
      if (stoptime /= 0.d0) then
        tf = lp%time
        nstop = 0
        do while (nstop < n)
          nstop = nstop + 1
          tf = real(DELTA_T + dble(tf))
          if (dble(tf) >= stoptime) exit
        end do
        n = min(n, nstop)
      end if
      if (allocated(probes)) deallocate(probes)
      allocate(probes(3 * n_varreduras * n))
      call pf_check(pf_run_steps(ctx, prm, lp, int(n, c_int32_t), c_loc(probes)), 'pf_run_steps')
      do q = 0, n_varreduras * n - 1
        write(*, '(9(1X,e14.4))') probes(3 * q + 1), probes(3 * q + 2), probes(3 * q + 3)
      end do
      it = prm%it
      E_R = E_R + (n - 1)   ! the event checks of the first n-1 steps did not fire
    else
      it = it + 1
      prm%it = it
      lp%time = real(DELTA_T + dble(lp%time))
      call temperatura
      if (taxa_resfr > 0.d0) prm%TP = T0 - taxa_resfr * DELTA_T * it
      do n_sw = 1, n_varreduras
        if (n_sw > 1) call pf_check(pf_phase_new_to_old(ctx), 'phase old<-new')
        call equacao_campo_fase
      end do
      call concentracao_soluto
      call auxiliar
    end if

    if (E_R >= WriteFileD_timeCounter) then   ! output event
      E_R = 0
      call event_header()
      call diagnosticos(L2_phi, phi_max, phi_min, solute_total, free_energy)
      write(*, fmt_diag) L2_phi, phi_max, phi_min, solute_total, free_energy
      if (write_snapshots) then
        call download_fields()
        call saida()
      end if
      call print_timing()
    end if
    E_R = E_R + 1
    if (.not. fused) call pf_check(pf_end_step(ctx), 'end of step')

    if (stoptime /= 0.d0 .and. dble(lp%time) >= stoptime) then
      write(*, '(//,5X,A,f14.7,A,/)') 'Program terminated by stop time condition = ', stoptime, ' s'
      call event_header()
      if (E_R < WriteFileD_timeCounter .and. write_snapshots) then
        call download_fields()
        call saida()
      end if
      call print_timing()
      exit
    end if
  end do

  call pf_check(pf_finalize(ctx), 'pf_finalize')

contains

  subroutine download_fields()
    call pf_check(pf_download(ctx, phase_field_new, conc_soluto_new, TNEW), 'pf_download')
  end subroutine download_fields

  ! The event banner of principal.f90:166-175 (formats 2000-2050).
  subroutine event_header()
    write(*, '(A)') '====================================================='
    write(*, '(A)') '=============  ESCREVENDO RESULTADOS  ==============='
    write(*, '(A)') '====================================================='
    write(*, '(A)') '====================================================='
    write(*, '(A,I8,A)') 'IT==> ', it, ' iteracoes'
    write(*, '(A,d12.4,A)') 'Tempo total simulado ==> ', lp%time, ' s'
    write(*, '(A,d12.4,A)') 'Tempo de parada (stoptime) ==> ', stoptime, ' s'
    write(*, '(A,d12.4,A)') 'Intervalo de escrita em arquivo  ==> ', &
          WriteFileD_timeCounter * DELTA_T, ' s'
    write(*, '(A)') '====================================================='
  end subroutine event_header

  ! The wall-clock block of principal.f90:308-408, in its formats: the log
  ! file gets the event banner and the diagnostics line, then the console and
  ! the log get the timing block.  Progress is time/stoptime, or it/it_max
  ! when stoptime is 0.  The host side of this driver is one thread.
  subroutine print_timing()
    integer(8) :: clk
    double precision :: el, dl, avg_step, avg_output, pct, eta
    integer :: u, ku, units(2), nunits, ios, ndev32
    integer(c_int32_t) :: ndev, devid
    call system_clock(clk)
    el = dble(clk - clk0) / dble(clk_rate)
    dl = dble(clk - clk_prev) / dble(clk_rate)
    n_outputs = n_outputs + 1
    avg_step = el / dble(max(it, 1))
    avg_output = el / dble(max(n_outputs, 1))
    if (stoptime > 0.d0) then
      pct = 100.d0 * dble(lp%time) / stoptime
    else
      pct = 100.d0 * dble(it) / dble(max(it_max, 1))
    end if
    eta = 0.d0
    if (pct > 0.d0) eta = el * (100.d0 - pct) / pct
    ndev = 0
    if (pf_device_count(ndev) /= 0) ndev = 0
    ndev32 = ndev
    devid = 0
    if (pf_ctx_device(ctx, devid) /= 0) devid = -1
    open(iu_log, file=perf_log_file, status='old', position='append', action='write', iostat=ios)
    nunits = 1; units(1) = 6
    if (ios == 0) then
      nunits = 2; units(2) = iu_log
      write(iu_log, '(A)') '====================================================='
      write(iu_log, '(A)') '=============  ESCREVENDO RESULTADOS  ==============='
      write(iu_log, '(A)') '====================================================='
      write(iu_log, '(A)') '====================================================='
      write(iu_log, '(A,I8,A)') 'IT==> ', it, ' iteracoes'
      write(iu_log, '(A,D12.4,A)') 'Tempo total simulado ==> ', lp%time, ' s'
      write(iu_log, '(A,D12.4,A)') 'Tempo de parada (stoptime) ==> ', stoptime, ' s'
      write(iu_log, '(A,D12.4,A)') 'Intervalo de escrita em arquivo  ==> ', &
            WriteFileD_timeCounter * DELTA_T, ' s'
      write(iu_log, '(A)') '====================================================='
      write(iu_log, fmt_diag) L2_phi, phi_max, phi_min, solute_total, free_energy
    end if
    do ku = 1, nunits
      u = units(ku)
      write(u, '(A)') '--------------------------------------------------------'
      write(u, '(A,I0)') 'Output Number          : ', n_outputs
      write(u, '(A,F0.4)') 'Simulation Time        : ', lp%time
      write(u, '(A,I0)') 'Iteration              : ', it
      write(u, '(A)') ''
      write(u, '(A)') 'Execution Device       : GPU (HIP, gfx950)'
      write(u, '(A,I0)') 'GPU Device ID          : ', devid
      write(u, '(A,I0)') 'Available Devices      : ', ndev32
      write(u, '(A,I0)') 'OpenMP Threads         : ', 1
      write(u, '(A)') ''
      write(u, '(A,F0.2,A)') 'Elapsed Time           : ', el, ' s'
      write(u, '(A,F0.2,A)') 'Time Since Last Output : ', dl, ' s'
      write(u, '(A)') ''
      write(u, '(A,F0.5,A)') 'Average Time/Step      : ', avg_step, ' s'
      write(u, '(A,F0.2,A)') 'Average Time/Output    : ', avg_output, ' s'
      write(u, '(A,I0)') 'Steps This Interval    : ', it - it_prev_output
      write(u, '(A,I0)') 'Total Outputs Written  : ', n_outputs
      write(u, '(A,F0.2,A)') 'Completed              : ', pct, ' %'
      write(u, '(A,F0.1,A)') 'Estimated Remaining    : ', eta, ' s'
      write(u, '(A)') '--------------------------------------------------------'
    end do
    if (ios == 0) close(iu_log)
    clk_prev = clk
    it_prev_output = it
  end subroutine print_timing

end program principal
