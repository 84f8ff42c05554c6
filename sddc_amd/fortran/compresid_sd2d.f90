! compresid_sd2d.f90 - drop-in replacement of the reference's argument-less
! `subroutine compresid` (compresid.f90:1-80) that runs the residual on
! the GPU through the C ABI of include/sd2d.h.
!
! Link it in place of compresid.o (with sd2d_c.f90 and libsd2d.so); the
! callers RK_S5P3 (iterations.f90:183-209) and FiveStageRk
! (iterations.f90:107-144) are unchanged.  Per call it has the module-state
! contract of the reference routine:
!   reads   Q, eigvmax, umax, eigvmax_g, umax_g (and, on the first call,
!           the mesh tables, operators and metrics of INIT_SETUP)
!   writes  resid, nablaQs, c_h, eigvmax, umax, eigvmax_g, umax_g and,
!           with vismode=1 and ARTIV=1, SM_IND, muAV, epsilon0
! The library context is created on the first call and lives until the
! program ends.  A failure prints the library's message and stops, as the
! reference does on its own set-up errors.
module compresid_sd2d_state
  use iso_c_binding
  implicit none
  type(c_ptr), save :: sd2d_ctx = c_null_ptr
contains
  ! Address of the first element of a contiguous module array (the arrays
  ! passed here are setup2d allocatables, which stay allocated and do not
  ! move, and sd2d_create copies them before it returns).
  type(c_ptr) function iaddr(a, n)
    integer(c_int32_t), target, intent(in) :: a(*)
    integer, intent(in) :: n
    iaddr = c_null_ptr
    if (n > 0) iaddr = c_loc(a(1))
  end function iaddr
  type(c_ptr) function daddr(a)
    real(c_double), target, intent(in) :: a(*)
    daddr = c_loc(a(1))
  end function daddr
end module compresid_sd2d_state

subroutine compresid_sd2d_create
  use setup2d
  use sd2d_c
  use compresid_sd2d_state
  implicit none
  type(sd2d_mesh_t) :: m
  type(sd2d_params_t) :: p
  integer(c_int) :: rc

  m%N = N
  m%ncell = NCELL
  m%nface = NFACE
  m%if2c = iaddr(IF2C, NFACE)
  m%ninter = NINTER;   m%ibfinter = iaddr(IBFINTER, NINTER)
  m%ninlet = NINLET;   m%ibfinl = iaddr(IBFINL, NINLET)
  m%noutlet = NOUTLET; m%ibfout = iaddr(IBFOUT, NOUTLET)
  m%nwall = NWALL;     m%ibfwal = iaddr(IBFWAL, NWALL)
  m%nsymp = NSYMP;     m%ibfsymp = iaddr(IBFSYMP, NSYMP)
  m%nfree = NFREE;     m%ibfree = iaddr(IBFREE, NFREE)
  m%npair = 0;         m%cyc2fpair = c_null_ptr
  if (allocated(CYC2FPAIR)) then
     m%npair = NPAIR
     m%cyc2fpair = iaddr(CYC2FPAIR, NPAIR)
  end if
  m%lmat = daddr(Lmat)
  m%mmat = daddr(Mmat)
  m%xs = daddr(Xs)
  m%xs_low = c_null_ptr
  if (allocated(Xs_low)) m%xs_low = daddr(Xs_low)
  m%jac = daddr(Jac)
  m%s1 = daddr(S1)
  m%s2 = daddr(S2)

  p%vismode = vismode
  p%artiv = ARTIV
  p%mu = mu
  p%eta0 = eta0
  p%alpha = alpha
  p%rinf = rinf; p%pinf = pinf; p%uinf = uinf; p%vinf = vinf
  p%bxinf = Bxinf; p%byinf = Byinf
  p%dxcycl = DXCYCL; p%dycycl = DYCYCL
  p%ncellglob = NCELLGLOB
  p%k_stage = k_stage
  p%dt = dt
  p%precision = SD2D_FP64
  p%tile_shape = SD2D_TILE_DEFAULT

  rc = sd2d_create(m, p, sd2d_ctx)
  if (rc /= SD2D_OK) then
     write(*,*) 'sd2d_create failed: ', sd2d_error_message(c_null_ptr)
     stop
  end if
end subroutine compresid_sd2d_create

subroutine compresid
  use setup2d
  use sd2d_c
  use compresid_sd2d_state
  implicit none
  type(sd2d_maxima_t) :: mx
  type(sd2d_diag_t) :: d
  integer(c_int) :: rc

  if (.not. c_associated(sd2d_ctx)) call compresid_sd2d_create

  mx = sd2d_maxima_t(eigvmax, umax, eigvmax_g, umax_g)
  rc = sd2d_set_state(sd2d_ctx, Q, mx)
  if (rc == SD2D_OK) rc = sd2d_compresid(sd2d_ctx, resid, nablaQs)
  if (rc == SD2D_OK) rc = sd2d_get_state(sd2d_ctx, maxima=mx)
  if (rc == SD2D_OK) rc = sd2d_get_diag(sd2d_ctx, d)
  if (rc == SD2D_OK .and. vismode == 1 .and. ARTIV == 1) &
       rc = sd2d_get_av(sd2d_ctx, SM_IND, muAV, epsilon0)
  if (rc /= SD2D_OK) then
     write(*,*) 'sd2d compresid failed: ', sd2d_error_message(sd2d_ctx)
     stop
  end if
  eigvmax = mx%eigvmax
  umax = mx%umax
  eigvmax_g = mx%eigvmax_g
  umax_g = mx%umax_g
  c_h = d%c_h
end subroutine compresid
