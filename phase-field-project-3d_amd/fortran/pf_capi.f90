! pf_capi.f90 -- ISO_C_BINDING interface to libpfamd (include/pf_capi.h).
!
! This is the binding a Fortran caller (our driver, or the reference's own
! principal.f90) uses to reach the gfx950 kernels.  Types mirror the C
! structs field by field; every function returns PF_OK (0) or a negative
! PF_E_* code, with pf_last_error() giving the message.
module pf_capi
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: PF_OK = 0, PF_E_ARG = -1, PF_E_HIP = -2, PF_E_NOMEM = -3, &
                               PF_E_STATE = -4, PF_E_NODEV = -5
  integer(c_int32_t), parameter :: PF_FP64 = 0, PF_FP32_MIXED = 1
  integer, parameter :: PF_KT_COUNT = 5

  ! struct pf_params: the kernel-visible part of module variaveis
  type, bind(C) :: pf_params_t
    integer(c_int32_t) :: nx = 0, ny = 0, nz = 0, it = 0
    integer(c_int32_t) :: n_varreduras = 1, h_quintico = 1, ruido_tipo = 2, precision = 0
    integer(c_int64_t) :: seed = 0
    real(c_double) :: dx = 0, dy = 0, dz = 0, dt = 0, D = 0, delta_H = 0, cp_s = 0, TP = 0
    real(c_double) :: Tm = 0, Me = 0, Vm = 0, R = 0, E0 = 0, W0 = 0, M0 = 0, delta_eps = 0
    real(c_double) :: AAA = 0, D_liq = 0, D_sol = 0, Cs_eq = 0, Cl_eq = 0
    real(c_float) :: Ke = 0
    integer(c_int32_t) :: reserved = 0
  end type pf_params_t

  ! struct pf_loop: what principal.f90 keeps between kernel calls
  type, bind(C) :: pf_loop_t
    real(c_double) :: T0 = 0, taxa_resfr = 0
    real(c_float) :: Sigma = 0, Beta = 0, ajuste = 0, time = 0
  end type pf_loop_t

  interface
    function pf_last_error_c() bind(C, name='pf_last_error') result(s)
      import :: c_ptr
      type(c_ptr) :: s
    end function
    integer(c_int) function pf_device_count(n) bind(C, name='pf_device_count')
      import :: c_int, c_int32_t
      integer(c_int32_t), intent(out) :: n
    end function
    integer(c_int32_t) function pf_w0_e0(lambda, ajuste, sigma, W0, E0) bind(C, name='pf_w0_e0')
      import :: c_int32_t, c_double, c_float
      real(c_double), value :: lambda
      real(c_float), value :: ajuste, sigma
      real(c_double), intent(out) :: W0, E0
    end function
    integer(c_int) function pf_ctx_device(ctx, dev) bind(C, name='pf_ctx_device')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: dev
    end function
    integer(c_int32_t) function pf_w0_e0_tries(lambda, ajuste, sigma, diff, cap) &
        bind(C, name='pf_w0_e0_tries')
      import :: c_int32_t, c_double, c_float
      real(c_double), value :: lambda
      real(c_float), value :: ajuste, sigma
      real(c_double), intent(out) :: diff(*)
      integer(c_int32_t), value :: cap
    end function
    real(c_double) function pf_auxiliar_m0(TP, Cl_eq, Cs_eq, R, Vm, E0, W0, Sigma, Beta, Ke, Me, &
                                           D_liq, ajuste) bind(C, name='pf_auxiliar_m0')
      import :: c_double, c_float
      real(c_double), value :: TP, Cl_eq, Cs_eq, R, Vm, E0, W0, Me, D_liq
      real(c_float), value :: Sigma, Beta, Ke, ajuste
    end function
    subroutine pf_equilibrium(TP, Tm, Ke, Me, Cs_eq, Cl_eq) bind(C, name='pf_equilibrium')
      import :: c_double, c_float
      real(c_double), value :: TP, Tm, Me
      real(c_float), value :: Ke
      real(c_double), intent(out) :: Cs_eq, Cl_eq
    end subroutine
    real(c_double) function pf_time_step(dx, D_liq) bind(C, name='pf_time_step')
      import :: c_double
      real(c_double), value :: dx, D_liq
    end function
    real(c_double) function pf_noise_scale(AAA, dt, dx, D_liq, ruido) bind(C, name='pf_noise_scale')
      import :: c_double, c_int32_t
      real(c_double), value :: AAA, dt, dx, D_liq
      integer(c_int32_t), value :: ruido
    end function
    integer(c_int) function pf_initial_fields(nx, ny, nz, tipo, R_semente, T0, C0, Ke, n_seeds, &
                                              seed_rng, phio, phin, co, cn, to, tn, dpp) &
                                              bind(C, name='pf_initial_fields')
      import :: c_int, c_int32_t, c_int64_t, c_double, c_float
      integer(c_int32_t), value :: nx, ny, nz, tipo, n_seeds
      real(c_double), value :: R_semente, T0, C0
      real(c_float), value :: Ke
      integer(c_int64_t), value :: seed_rng
      real(c_double), intent(out) :: phio(*), phin(*), co(*), cn(*), to(*), tn(*), dpp(*)
    end function
    integer(c_int) function pf_init(p, phio, phin, co, cn, to, tn, dpp, ctx) bind(C, name='pf_init')
      import :: c_int, c_double, c_ptr, pf_params_t
      type(pf_params_t), intent(in) :: p
      real(c_double), intent(in) :: phio(*), phin(*), co(*), cn(*), to(*), tn(*), dpp(*)
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function pf_finalize(ctx) bind(C, name='pf_finalize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pf_temperatura(ctx, p) bind(C, name='pf_temperatura')
      import :: c_int, c_ptr, pf_params_t
      type(c_ptr), value :: ctx
      type(pf_params_t), intent(in) :: p
    end function
    integer(c_int) function pf_equacao_campo_fase(ctx, p, probe) bind(C, name='pf_equacao_campo_fase')
      import :: c_int, c_ptr, c_double, pf_params_t
      type(c_ptr), value :: ctx
      type(pf_params_t), intent(in) :: p
      real(c_double), intent(out) :: probe(3)
    end function
    integer(c_int) function pf_phase_new_to_old(ctx) bind(C, name='pf_phase_new_to_old')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pf_concentracao_soluto(ctx, p) bind(C, name='pf_concentracao_soluto')
      import :: c_int, c_ptr, pf_params_t
      type(c_ptr), value :: ctx
      type(pf_params_t), intent(in) :: p
    end function
    integer(c_int) function pf_end_step(ctx) bind(C, name='pf_end_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pf_diagnosticos(ctx, p, out) bind(C, name='pf_diagnosticos')
      import :: c_int, c_ptr, c_double, pf_params_t
      type(c_ptr), value :: ctx
      type(pf_params_t), intent(in) :: p
      real(c_double), intent(out) :: out(5)
    end function
    integer(c_int) function pf_download(ctx, phin, cn, tn) bind(C, name='pf_download')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: phin(*), cn(*), tn(*)
    end function
    integer(c_int) function pf_run_steps(ctx, p, L, nsteps, probes) bind(C, name='pf_run_steps')
      import :: c_int, c_ptr, c_int32_t, pf_params_t, pf_loop_t
      type(c_ptr), value :: ctx
      type(pf_params_t), intent(inout) :: p
      type(pf_loop_t), intent(inout) :: L
      integer(c_int32_t), value :: nsteps
      type(c_ptr), value :: probes
    end function
    integer(c_int) function pf_synchronize(ctx) bind(C, name='pf_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface

contains

  ! pf_last_error() as a Fortran string
  function pf_last_error() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = pf_last_error_c()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function pf_last_error

  ! Stop with the library's message when a call fails.
  subroutine pf_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= PF_OK) then
      write(*, '(A,A,A,I0,A,A)') 'pf3d: ', where, ' failed (', rc, '): ', pf_last_error()
      error stop 2
    end if
  end subroutine pf_check

end module pf_capi
