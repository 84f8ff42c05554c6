! sd2d_c.f90 - ISO_C_BINDING interface to the C ABI of include/sd2d.h.
!
! The derived types mirror the C structs field by field (bind(C) gives
! them the C layout, padding included); the functions are the C entry
! points.  Array arguments are assumed-size dummies, so a caller passes
! its module arrays (e.g. setup2d's Q, resid, nablaQs) directly, without
! the TARGET attribute; optional dummies become NULL when absent.
module sd2d_c
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: SD2D_OK = 0, SD2D_EINVAL = 1, &
       SD2D_EHIP = 2, SD2D_ENOMEM = 3, SD2D_ENONFINITE = 4, &
       SD2D_EUNSUPPORTED = 5
  integer(c_int32_t), parameter, public :: SD2D_FP64 = 0, SD2D_MIXED = 1
  integer(c_int32_t), parameter, public :: SD2D_TILE_DEFAULT = 0, &
       SD2D_TILE_STRIPS = 1, SD2D_TILE_BLOCKS = 2, SD2D_TILE_ROWPAIRS = 3

  type, bind(C), public :: sd2d_mesh_t
     integer(c_int32_t) :: N, ncell, nface
     type(c_ptr) :: if2c
     integer(c_int32_t) :: ninter
     type(c_ptr) :: ibfinter
     integer(c_int32_t) :: ninlet
     type(c_ptr) :: ibfinl
     integer(c_int32_t) :: noutlet
     type(c_ptr) :: ibfout
     integer(c_int32_t) :: nwall
     type(c_ptr) :: ibfwal
     integer(c_int32_t) :: nsymp
     type(c_ptr) :: ibfsymp
     integer(c_int32_t) :: nfree
     type(c_ptr) :: ibfree
     integer(c_int32_t) :: npair
     type(c_ptr) :: cyc2fpair
     type(c_ptr) :: lmat, mmat, xs, xs_low, jac, s1, s2
  end type sd2d_mesh_t

  type, bind(C), public :: sd2d_params_t
     integer(c_int32_t) :: vismode, artiv
     real(c_double) :: mu, eta0, alpha
     real(c_double) :: rinf, pinf, uinf, vinf, bxinf, byinf
     real(c_double) :: dxcycl, dycycl
     integer(c_int32_t) :: ncellglob, k_stage
     real(c_double) :: dt
     integer(c_int32_t) :: precision, tile_shape
  end type sd2d_params_t

  type, bind(C), public :: sd2d_maxima_t
     real(c_double) :: eigvmax, umax, eigvmax_g, umax_g
  end type sd2d_maxima_t

  type, bind(C), public :: sd2d_diag_t
     real(c_double) :: resid_rms, divb_rms, eigvmax_g, umax_g, c_h, epsilon0, ctime
     integer(c_int32_t) :: iter, norms_valid
  end type sd2d_diag_t

  public :: sd2d_create, sd2d_destroy, sd2d_last_error, sd2d_set_state, &
       sd2d_get_state, sd2d_set_time, sd2d_compresid, sd2d_compresid_q, &
       sd2d_compresid_noalloc, sd2d_step, sd2d_rk_stage, sd2d_get_diag, &
       sd2d_get_av, sd2d_plot_nodes, sd2d_write_restart, sd2d_read_restart, sd2d_write_checkpoint, &
       sd2d_read_checkpoint, sd2d_set_profiling, sd2d_get_profile, &
       sd2d_set_check_finite, sd2d_ens_create, sd2d_ens_destroy, &
       sd2d_ens_last_error, sd2d_ens_set_state, sd2d_ens_get_state, &
       sd2d_ens_set_time, sd2d_ens_get_time, sd2d_ens_write_checkpoint, &
       sd2d_ens_read_checkpoint, sd2d_ens_step, sd2d_ens_tune, &
       sd2d_ens_set_check_finite, sd2d_ens_nonfinite, sd2d_fwd, sd2d_vjp, &
       sd2d_compresid_vjp, sd2d_adj_save, sd2d_adj_load, sd2d_error_message

  interface
     integer(c_int) function sd2d_create(mesh, params, ctx) bind(C, name="sd2d_create")
       import :: c_int, c_ptr, sd2d_mesh_t, sd2d_params_t
       type(sd2d_mesh_t), intent(in) :: mesh
       type(sd2d_params_t), intent(in) :: params
       type(c_ptr), intent(out) :: ctx
     end function sd2d_create

     subroutine sd2d_destroy(ctx) bind(C, name="sd2d_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine sd2d_destroy

     type(c_ptr) function sd2d_last_error(ctx) bind(C, name="sd2d_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function sd2d_last_error

     integer(c_int) function sd2d_set_state(ctx, q, maxima) bind(C, name="sd2d_set_state")
       import :: c_int, c_ptr, c_double, sd2d_maxima_t
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: q(*)
       type(sd2d_maxima_t), intent(in), optional :: maxima
     end function sd2d_set_state

     integer(c_int) function sd2d_get_state(ctx, q, maxima) bind(C, name="sd2d_get_state")
       import :: c_int, c_ptr, c_double, sd2d_maxima_t
       type(c_ptr), value :: ctx
       real(c_double), intent(out), optional :: q(*)
       type(sd2d_maxima_t), intent(out), optional :: maxima
     end function sd2d_get_state

     integer(c_int) function sd2d_set_time(ctx, ctime, iter) bind(C, name="sd2d_set_time")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), value :: ctime
       integer(c_int32_t), value :: iter
     end function sd2d_set_time

     integer(c_int) function sd2d_compresid(ctx, resid, nablaqs) bind(C, name="sd2d_compresid")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out), optional :: resid(*)
       real(c_double), intent(out), optional :: nablaqs(*)
     end function sd2d_compresid

     integer(c_int) function sd2d_compresid_q(ctx, q, resid, nablaqs) &
          bind(C, name="sd2d_compresid_q")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: q(*)
       real(c_double), intent(out), optional :: resid(*)
       real(c_double), intent(out), optional :: nablaqs(*)
     end function sd2d_compresid_q

     ! no counterpart in the reference: sd2d_compresid without any device
     ! allocation, for hosts that cannot allocate (graph capture, serving loops)
     integer(c_int) function sd2d_compresid_noalloc(ctx, resid, nablaqs) &
          bind(C, name="sd2d_compresid_noalloc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out), optional :: resid(*)
       real(c_double), intent(out), optional :: nablaqs(*)
     end function sd2d_compresid_noalloc

     integer(c_int) function sd2d_step(ctx, nsteps, norm_every) bind(C, name="sd2d_step")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsteps, norm_every
     end function sd2d_step

     integer(c_int) function sd2d_rk_stage(ctx, stage) bind(C, name="sd2d_rk_stage")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: stage
     end function sd2d_rk_stage

     integer(c_int) function sd2d_get_diag(ctx, diag) bind(C, name="sd2d_get_diag")
       import :: c_int, c_ptr, sd2d_diag_t
       type(c_ptr), value :: ctx
       type(sd2d_diag_t), intent(out) :: diag
     end function sd2d_get_diag

     integer(c_int) function sd2d_get_av(ctx, sm_ind, muav, epsilon0) bind(C, name="sd2d_get_av")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out), optional :: sm_ind(*)
       real(c_double), intent(out), optional :: muav(*)
       real(c_double), intent(out), optional :: epsilon0
     end function sd2d_get_av

     ! nodal values of the Tecplot output (plot mesh from the host, CSR)
     integer(c_int) function sd2d_plot_nodes(ctx, nre, nnode, xadj, occ, muav, plotq, plotav) &
          bind(C, name="sd2d_plot_nodes")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nre, nnode
       integer(c_int32_t), intent(in) :: xadj(*), occ(*)
       real(c_double), intent(in), optional :: muav(*)
       real(c_double), intent(out) :: plotq(*)
       real(c_double), intent(out), optional :: plotav(*)
     end function sd2d_plot_nodes

     ! path: a NUL-terminated string, e.g. trim(name)//c_null_char
     integer(c_int) function sd2d_write_restart(ctx, path, icg) bind(C, name="sd2d_write_restart")
       import :: c_int, c_ptr, c_char, c_int32_t
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), intent(in), optional :: icg(*)
     end function sd2d_write_restart

     integer(c_int) function sd2d_read_restart(ctx, path) bind(C, name="sd2d_read_restart")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
     end function sd2d_read_restart

     ! no counterpart in the reference: restart file plus running-maxima sidecar
     integer(c_int) function sd2d_write_checkpoint(ctx, path, icg) &
          bind(C, name="sd2d_write_checkpoint")
       import :: c_int, c_ptr, c_char, c_int32_t
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), intent(in), optional :: icg(*)
     end function sd2d_write_checkpoint

     integer(c_int) function sd2d_read_checkpoint(ctx, path) bind(C, name="sd2d_read_checkpoint")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
     end function sd2d_read_checkpoint

     integer(c_int) function sd2d_set_profiling(ctx, on) bind(C, name="sd2d_set_profiling")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: on
     end function sd2d_set_profiling

     integer(c_int) function sd2d_get_profile(ctx, stage_ms, nlaunch) &
          bind(C, name="sd2d_get_profile")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), intent(out), optional :: stage_ms
       integer(c_int32_t), intent(out), optional :: nlaunch
     end function sd2d_get_profile

     ! opt-in NaN/Inf check of the state every `every` steps (0: off)
     integer(c_int) function sd2d_set_check_finite(ctx, every) bind(C, name="sd2d_set_check_finite")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: every
     end function sd2d_set_check_finite

     ! ensembles: meshes(nmem), params(nmem) arrays of the derived types
     integer(c_int) function sd2d_ens_create(nmem, meshes, params, ens) &
          bind(C, name="sd2d_ens_create")
       import :: c_int, c_int32_t, c_ptr, sd2d_mesh_t, sd2d_params_t
       integer(c_int32_t), value :: nmem
       type(sd2d_mesh_t), intent(in) :: meshes(*)
       type(sd2d_params_t), intent(in) :: params(*)
       type(c_ptr), intent(out) :: ens
     end function sd2d_ens_create

     subroutine sd2d_ens_destroy(ens) bind(C, name="sd2d_ens_destroy")
       import :: c_ptr
       type(c_ptr), value :: ens
     end subroutine sd2d_ens_destroy

     type(c_ptr) function sd2d_ens_last_error(ens) bind(C, name="sd2d_ens_last_error")
       import :: c_ptr
       type(c_ptr), value :: ens
     end function sd2d_ens_last_error

     ! member: 0-based index
     integer(c_int) function sd2d_ens_set_state(ens, member, q, maxima) &
          bind(C, name="sd2d_ens_set_state")
       import :: c_int, c_int32_t, c_ptr, c_double, sd2d_maxima_t
       type(c_ptr), value :: ens
       integer(c_int32_t), value :: member
       real(c_double), intent(in) :: q(*)
       type(sd2d_maxima_t), intent(in), optional :: maxima
     end function sd2d_ens_set_state

     integer(c_int) function sd2d_ens_get_state(ens, member, q, maxima) &
          bind(C, name="sd2d_ens_get_state")
       import :: c_int, c_int32_t, c_ptr, c_double, sd2d_maxima_t
       type(c_ptr), value :: ens
       integer(c_int32_t), value :: member
       real(c_double), intent(out), optional :: q(*)
       type(sd2d_maxima_t), intent(out), optional :: maxima
     end function sd2d_ens_get_state

     integer(c_int) function sd2d_ens_set_time(ens, member, ctime, iter) &
          bind(C, name="sd2d_ens_set_time")
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ens
       integer(c_int32_t), value :: member
       real(c_double), value :: ctime
       integer(c_int32_t), value :: iter
     end function sd2d_ens_set_time

     integer(c_int) function sd2d_ens_get_time(ens, member, ctime, iter) &
          bind(C, name="sd2d_ens_get_time")
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ens
       integer(c_int32_t), value :: member
       real(c_double), intent(out), optional :: ctime
       integer(c_int32_t), intent(out), optional :: iter
     end function sd2d_ens_get_time

     integer(c_int) function sd2d_ens_write_checkpoint(ens, member, path, icg) &
          bind(C, name="sd2d_ens_write_checkpoint")
       import :: c_int, c_int32_t, c_ptr, c_char
       type(c_ptr), value :: ens
       integer(c_int32_t), value :: member
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), intent(in), optional :: icg(*)
     end function sd2d_ens_write_checkpoint

     integer(c_int) function sd2d_ens_read_checkpoint(ens, member, path) &
          bind(C, name="sd2d_ens_read_checkpoint")
       import :: c_int, c_int32_t, c_ptr, c_char
       type(c_ptr), value :: ens
       integer(c_int32_t), value :: member
       character(kind=c_char), intent(in) :: path(*)
     end function sd2d_ens_read_checkpoint

     integer(c_int) function sd2d_ens_step(ens, nsteps) bind(C, name="sd2d_ens_step")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ens
       integer(c_int32_t), value :: nsteps
     end function sd2d_ens_step

     ! shapes(3): the chosen SD2D_TILE_* for N = 3, 4, 5
     integer(c_int) function sd2d_ens_tune(ens, trial_steps, shapes) bind(C, name="sd2d_ens_tune")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ens
       integer(c_int32_t), value :: trial_steps
       integer(c_int32_t), intent(inout), optional :: shapes(*)
     end function sd2d_ens_tune

     integer(c_int) function sd2d_ens_set_check_finite(ens, every) &
          bind(C, name="sd2d_ens_set_check_finite")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ens
       integer(c_int32_t), value :: every
     end function sd2d_ens_set_check_finite

     ! flags(nmem): 1 for the members the latest check found non-finite
     integer(c_int) function sd2d_ens_nonfinite(ens, flags) bind(C, name="sd2d_ens_nonfinite")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ens
       integer(c_int32_t), intent(out) :: flags(*)
     end function sd2d_ens_nonfinite

     ! adjoint: forward with checkpoints, then the vector-Jacobian product
     integer(c_int) function sd2d_fwd(ctx, nsteps, ckpt, objective) bind(C, name="sd2d_fwd")
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsteps, ckpt
       real(c_double), intent(out), optional :: objective
     end function sd2d_fwd

     integer(c_int) function sd2d_vjp(ctx, qbar_final, jbar, qbar_init, pbar, chbar) &
          bind(C, name="sd2d_vjp")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in), optional :: qbar_final(*)
       real(c_double), value :: jbar
       real(c_double), intent(out), optional :: qbar_init(*), pbar(*), chbar(*)
     end function sd2d_vjp

     ! the adjoint checkpoints of the last sd2d_fwd as files, and back
     integer(c_int) function sd2d_adj_save(ctx, prefix) bind(C, name="sd2d_adj_save")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: prefix(*)
     end function sd2d_adj_save

     integer(c_int) function sd2d_adj_load(ctx, prefix) bind(C, name="sd2d_adj_load")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: prefix(*)
     end function sd2d_adj_load

     integer(c_int) function sd2d_compresid_vjp(ctx, rbar, qbar, chbar, pbar) &
          bind(C, name="sd2d_compresid_vjp")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: rbar(*)
       real(c_double), intent(out), optional :: qbar(*)
       real(c_double), intent(out), optional :: chbar
       real(c_double), intent(out), optional :: pbar(*)
     end function sd2d_compresid_vjp
  end interface

contains

  ! The message of sd2d_last_error as a Fortran string.
  function sd2d_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: n
    interface
       integer(c_size_t) function c_strlen(p) bind(C, name="strlen")
         import :: c_size_t, c_ptr
         type(c_ptr), value :: p
       end function c_strlen
    end interface
    type(c_ptr) :: p
    p = sd2d_last_error(ctx)
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function sd2d_error_message

end module sd2d_c
