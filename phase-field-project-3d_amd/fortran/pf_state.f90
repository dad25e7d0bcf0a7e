! pf_state.f90 -- the driver's state: the deck scalars, the host copies of
! the fields (used only at upload, output and restart) and the device
! context.  The field arrays themselves live in HBM for the whole run.
module pf_state
  use, intrinsic :: iso_c_binding
  use pf_capi
  implicit none

  ! deck values (entrada.dat records 4-43)
  double precision :: Tm, T0, delta_H, E0_deck, W0_deck, M0_deck, D_s, D_l, Cp_s, Cp_l
  integer :: Nx, Ny, Nz, it_max, WriteFileD_timeCounter
  double precision :: dx, dy, dz, small, ANI, AAA_deck, Teta0, DELTA_EPSLON, DELTA_T_deck
  double precision :: D_sol, D_liq, Me, Vm, R, C0, Thin_Interface, stoptime
  real :: Ke, Sigma, Beta, ajuste                    ! REAL(4), as in the reference
  double precision :: TP_deck, R_semente, taxa_resfr
  integer :: tipo_inicial, n_varreduras, h_quintico, ruido_tipo

  ! options of this build (side file pf3d.nml; the deck format is unchanged)
  integer :: precision = 0          ! 0 FP64, 1 FP32-mixed
  integer(c_int64_t) :: noise_seed = 0
  integer :: n_seeds = 0            ! tipo_inicial = 6
  integer(c_int64_t) :: seed_rng = 20261203_c_int64_t
  logical :: fused = .true.         ! pf_run_steps between output events
  logical :: write_snapshots = .true.
  integer :: it_stop = 0            ! > 0: stop after this many steps
  double precision :: D_thermal = 0.d0  ! the reference's D (never assigned there: 0)
  logical :: w0e0_trace = .true.    ! print every W0_E0 try, as W0_E0.f90:19 does

  ! derived scalars and loop state
  double precision :: DELTA_T, Lambda, AAA
  type(pf_params_t), target :: prm
  type(pf_loop_t), target :: lp
  integer :: it = 0, E_R = 0
  double precision :: L2_phi = 0, phi_max = 0, phi_min = 0, solute_total = 0, free_energy = 0

  ! host fields
  double precision, allocatable :: phase_field_old(:,:,:), phase_field_new(:,:,:)
  double precision, allocatable :: conc_soluto_old(:,:,:), conc_soluto_new(:,:,:)
  double precision, allocatable :: TOLD(:,:,:), TNEW(:,:,:), DPP(:,:,:)

  type(c_ptr) :: ctx = c_null_ptr

end module pf_state
