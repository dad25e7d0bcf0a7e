! pf_kernels.f90 -- the reference's hot-path subroutines, same names and
! arguments, now thin calls into libpfamd (gfx950).  A driver written for the
! reference (source/principal.f90:98-176) calls them unchanged:
!   call temperatura                      source/temperatura.f90:1
!   call equacao_campo_fase               source/equacao_campo_fase.f90:1
!   call concentracao_soluto              source/concentracao_soluto.f90:1
!   call auxiliar                         source/auxiliar.f90:1
!   call diagnosticos(L2,max,min,sc,fe)   source/diagnosticos.f90:1
!   call W0_E0                            source/W0_E0.f90:1
! State lives in module pf_state (the counterpart of module variaveis) and
! the fields in HBM behind the context `ctx`.

subroutine temperatura()
  use pf_state
  implicit none
  call pf_check(pf_temperatura(ctx, prm), 'temperatura')
end subroutine temperatura

subroutine equacao_campo_fase()
  use pf_state
  implicit none
  real(c_double) :: probe(3)
  ! Cs_eq/Cl_eq from the current TP (equacao_campo_fase.f90:21-22); auxiliar reads them
  call pf_equilibrium(prm%TP, prm%Tm, prm%Ke, prm%Me, prm%Cs_eq, prm%Cl_eq)
  call pf_check(pf_equacao_campo_fase(ctx, prm, probe), 'equacao_campo_fase')
  write(*, '(9(1X,e14.4))') probe(1), probe(2), probe(3)
end subroutine equacao_campo_fase

subroutine concentracao_soluto()
  use pf_state
  implicit none
  call pf_check(pf_concentracao_soluto(ctx, prm), 'concentracao_soluto')
end subroutine concentracao_soluto

subroutine auxiliar()
  use pf_state
  implicit none
  prm%M0 = pf_auxiliar_m0(prm%TP, prm%Cl_eq, prm%Cs_eq, prm%R, prm%Vm, prm%E0, prm%W0, &
                          lp%Sigma, lp%Beta, prm%Ke, prm%Me, prm%D_liq, lp%ajuste)
end subroutine auxiliar

subroutine diagnosticos(L2, pmax, pmin, sc, fe)
  use pf_state
  implicit none
  double precision, intent(out) :: L2, pmax, pmin, sc, fe
  real(c_double) :: out(5)
  call pf_check(pf_diagnosticos(ctx, prm, out), 'diagnosticos')
  L2 = out(1); pmax = out(2); pmin = out(3); sc = out(4); fe = out(5)
end subroutine diagnosticos

subroutine W0_E0()
  ! The search runs in the library (bit-exact); with w0e0_trace (the
  ! default) every try's |E0 - try| is printed as W0_E0.f90:19 does, then
  ! the final line of W0_E0.f90:30.
  use pf_state
  implicit none
  integer(c_int32_t) :: idx, n, q
  real(c_double) :: w, e
  real(c_double), allocatable :: diff(:)
  idx = pf_w0_e0(Lambda, ajuste, Sigma, w, e)
  if (w0e0_trace) then
    n = idx + 1
    if (idx < 0) n = 10000001
    allocate(diff(n))
    n = pf_w0_e0_tries(Lambda, ajuste, Sigma, diff, n)
    if (n < 0) n = size(diff)
    do q = 1, n
      write(*, *) "E0-EPSLON_Try==>", diff(q)
    end do
    deallocate(diff)
  end if
  prm%W0 = w
  prm%E0 = e
  write(*, *) 'W0==>', prm%W0, '     ', 'E0==>', prm%E0, '     ', 'Dt==>', DELTA_T
end subroutine W0_E0
