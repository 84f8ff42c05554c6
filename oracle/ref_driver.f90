! ref_driver.f90 - fixture generator linked against the reference solver's
! own objects (everything but main.o), built by oracle/ref.mk into
! oracle/_ref/.  Test infrastructure only: it runs in the build container,
! never on the GPU machine, and nothing in the product links it.
!
! It performs the reference's set-up chain (main.f90:16-46, the .vrt8
! nodes read when CURVE_WALL = 1 as main.f90:28 does), dumps the
! set-up arrays, then advances the state:
!   - the first NSTAGED steps stage by stage, calling the reference's
!     compresid and applying the RK_S5P3 stage updates
!     (iterations.f90:179-211), dumping every stage's inputs and outputs;
!     each staged step is replayed from a saved copy through the
!     reference's own RK_S5P3 and the two results are compared bit by bit;
!   - NRK further whole steps through RK_S5P3 / FiveStageRk.
! Control file DRIVER.INP: NSTAGED, NRK.  Output: setup.bin,
! stage_SSS.bin, final.bin in the run directory.
program ref_driver
    use setup2d
    implicit none
    include 'mpif.h'

    integer :: nstaged, nrk, istep, istage, nst, u, bitok
    double precision, allocatable :: Qk0(:,:,:,:), Qka(:,:,:,:), Qsave(:,:,:,:)
    double precision :: ctime0, dt1, dt2, dt3, dt4, sv(5)
    integer :: itsave
    character(64) :: fname
    double precision, parameter :: a30=0.355909775063327d0
    double precision, parameter :: a32=0.644090224936674d0
    double precision, parameter :: a40=0.367933791638137d0
    double precision, parameter :: a43=0.632066208361863d0
    double precision, parameter :: a52=0.237593836598569d0
    double precision, parameter :: a54=0.762406163401431d0
    double precision, parameter :: b10=0.377268915331368d0
    double precision, parameter :: b21=0.377268915331368d0
    double precision, parameter :: b32=0.242995220537396d0
    double precision, parameter :: b43=0.238458932846290d0
    double precision, parameter :: b54=0.287632146308408d0

    call MPI_INIT(ierror)
    call MPI_COMM_SIZE(MPI_COMM_WORLD, size, ierror)
    call MPI_COMM_RANK(MPI_COMM_WORLD, rank, ierror)
    NPROC = size

    open(41, file='DRIVER.INP')
    read(41,*) nstaged
    read(41,*) nrk
    close(41)

    call readinput
    call READ_CELL_DATA
    call READ_VRT_DATA
    call CHECK_MESH
    call READ_METIS
    call GLOB2LOCAL
    if (CURVE_WALL==1) call READ_VRT8_DATA
    call READ_BND_DATA
    call CONNECTIVITY
    call CONNECT_BDRY
    call MATCHPROC
    call MATCHCYCREM
    call GETIVCELL_PROC
    call GETIVCELL_CYC
    call MAPCYCLOC
    call INIT_SETUP
    ! module defaults are not guaranteed to be zero (path finding F1)
    eigvmax_g = 0.d0
    umax_g = 0.d0
    c_h = 0.d0

    open(newunit=u, file='setup.bin', access='stream', status='replace')
    call dump_i(u, 'N', [N], 1)
    call dump_i(u, 'NCELL', [NCELL], 1)
    call dump_i(u, 'NFACE', [NFACE], 1)
    call dump_i(u, 'IF2C', IF2C, 4*NFACE)
    call dump_i(u, 'IBFINTER', IBFINTER, NINTER)
    call dump_i(u, 'IBFINL', IBFINL, NINLET)
    call dump_i(u, 'IBFOUT', IBFOUT, NOUTLET)
    call dump_i(u, 'IBFWAL', IBFWAL, NWALL)
    call dump_i(u, 'IBFSYMP', IBFSYMP, NSYMP)
    call dump_i(u, 'IBFREE', IBFREE, NFREE)
    call dump_i(u, 'NPAIR', [NPAIR], 1)
    if (NPAIR > 0) call dump_i(u, 'CYC2FPAIR', CYC2FPAIR, 2*NPAIR)
    call dump_d(u, 'Xs', Xs, N)
    call dump_d(u, 'Xf', Xf, N+1)
    call dump_d(u, 'Lmat', Lmat, (N+1)*N)
    call dump_d(u, 'Mmat', Mmat, (N+1)*N)
    call dump_d(u, 'Jac', Jac, N*N*NCELL)
    call dump_d(u, 'S1', S1, 4*(N+1)*N*NCELL)
    call dump_d(u, 'S2', S2, 4*N*(N+1)*NCELL)
    if (allocated(Xs_low)) call dump_d(u, 'Xs_low', Xs_low, N-1)
    call dump_d(u, 'XXsolu', XXsolu, 2*N*N*NCELL)
    call dump_d(u, 'Q0', Q, 9*N*N*NCELL)
    call dump_d(u, 'ctime', [ctime], 1)
    close(u)

    allocate(Qk0(9,N,N,NCELL), Qka(9,N,N,NCELL), Qsave(9,N,N,NCELL))
    nst = 0
    do istep = 1, nstaged
        iter = iter + 1
        ! save the full state of the path for the RK_S5P3 replay
        Qsave = Q
        sv = [eigvmax, umax, eigvmax_g, umax_g, ctime]
        itsave = iter

        Qk0 = Q
        ctime0 = ctime
        do istage = 1, 5
            nst = nst + 1
            write(fname, '(A,I3.3,A)') 'stage_', nst, '.bin'
            open(newunit=u, file=trim(fname), access='stream', status='replace')
            call dump_d(u, 'Qin', Q, 9*N*N*NCELL)
            call dump_d(u, 'pre', [eigvmax, umax, eigvmax_g, umax_g], 4)
            call compresid
            call dump_d(u, 'resid', resid, 9*N*N*NCELL)
            call dump_d(u, 'nablaQs', nablaQs, 18*N*N*NCELL)
            call dump_d(u, 'post', [eigvmax, umax, eigvmax_g, umax_g, c_h], 5)
            if (vismode == 1 .and. ARTIV == 1) then
                call dump_d(u, 'SM_IND', SM_IND, 2*NCELL)
                call dump_d(u, 'muAV', muAV, 2*NCELL)
                call dump_d(u, 'epsilon0', [epsilon0], 1)
            end if
            close(u)
            select case (istage)
            case (1)
                Q = Q + b10*dt*resid
                dt1 = b10*dt
                ctime = ctime0 + dt1
            case (2)
                Q = Q + b21*dt*resid
                dt2 = dt1 + b21*dt
                ctime = ctime0 + dt2
                Qka = Q
            case (3)
                Q = a30*Qk0 + a32*Q + b32*dt*resid
                dt3 = a32*dt2 + b32*dt
                ctime = ctime0 + dt3
            case (4)
                Q = a40*Qk0 + a43*Q + b43*dt*resid
                dt4 = a43*dt3 + b43*dt
                ctime = ctime0 + dt4
            case (5)
                Q = a52*Qka + a54*Q + b54*dt*resid
                ctime = ctime0 + dt
            end select
        end do

        ! replay the same step through the reference's RK_S5P3
        Qk0 = Q
        Q = Qsave
        eigvmax = sv(1); umax = sv(2); eigvmax_g = sv(3); umax_g = sv(4)
        ctime = sv(5)
        call RK_S5P3
        bitok = 1
        if (any(Q /= Qk0)) bitok = 0
        write(fname, '(A,I3.3,A)') 'replay_', istep, '.bin'
        open(newunit=u, file=trim(fname), access='stream', status='replace')
        call dump_i(u, 'bitequal', [bitok], 1)
        call dump_d(u, 'ctime', [ctime], 1)
        close(u)
    end do

    do istep = 1, nrk
        iter = iter + 1
        if (k_stage == 5) then
            call FiveStageRk
        else
            call RK_S5P3
        end if
    end do

    open(newunit=u, file='final.bin', access='stream', status='replace')
    call dump_d(u, 'Q', Q, 9*N*N*NCELL)
    call dump_d(u, 'maxima', [eigvmax, umax, eigvmax_g, umax_g, c_h], 5)
    call dump_d(u, 'ctime', [ctime], 1)
    call dump_i(u, 'iter', [iter], 1)
    call dump_d(u, 'resid', resid, 9*N*N*NCELL)
    call dump_d(u, 'nablaQs', nablaQs, 18*N*N*NCELL)
    if (vismode == 1 .and. ARTIV == 1) then
        call dump_d(u, 'SM_IND', SM_IND, 2*NCELL)
        call dump_d(u, 'muAV', muAV, 2*NCELL)
        call dump_d(u, 'epsilon0', [epsilon0], 1)
    end if
    close(u)

    call MPI_FINALIZE(ierror)

contains

    ! record: 16-char name, int32 kind (1=float64, 2=int32), int32 count, data
    subroutine dump_d(un, name, a, n)
        integer, intent(in) :: un, n
        character(*), intent(in) :: name
        double precision, intent(in) :: a(*)
        character(16) :: nm
        nm = name
        write(un) nm, 1, n
        if (n > 0) write(un) a(1:n)
    end subroutine dump_d

    subroutine dump_i(un, name, a, n)
        integer, intent(in) :: un, n
        character(*), intent(in) :: name
        integer, intent(in) :: a(*)
        character(16) :: nm
        nm = name
        write(un) nm, 2, n
        if (n > 0) write(un) a(1:n)
    end subroutine dump_i

end program ref_driver
