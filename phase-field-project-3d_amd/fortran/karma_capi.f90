! karma_capi.f90 -- ISO_C_BINDING interface to the Karma model's entry points
! (include/pf_karma.h) in libpfamd.  Status codes and pf_last_error() are
! those of pf_capi.
module pf_karma_capi
  use, intrinsic :: iso_c_binding
  implicit none

  ! struct pf_karma_params
  type, bind(C) :: pf_karma_params_t
    integer(c_int32_t) :: n = 0, pad_ = 0
    real(c_double) :: lambda = 0, eps4 = 0, dx = 0, dt = 0, R0 = 0, Fnoise = 0
    real(c_double) :: k = 0, Tm = 0, mliq = 0, C0 = 0, TP = 0
  end type pf_karma_params_t

  interface
    integer(c_int) function pf_karma_create(p, s) bind(C, name='pf_karma_create')
      import :: c_int, c_ptr, pf_karma_params_t
      type(pf_karma_params_t), intent(in) :: p
      type(c_ptr), intent(out) :: s
    end function
    integer(c_int) function pf_karma_run(s, nsteps) bind(C, name='pf_karma_run')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: s
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function pf_karma_download(s, phi, U, it) bind(C, name='pf_karma_download')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double), intent(out) :: phi(*), U(*)
      integer(c_int32_t), intent(out) :: it
    end function
    integer(c_int) function pf_karma_destroy(s) bind(C, name='pf_karma_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
  end interface
end module pf_karma_capi
