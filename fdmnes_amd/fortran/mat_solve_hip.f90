! Drop-in replacement of the Gaussian solver file of FDMNES
! (mat_solve_gaussian.f90): the same two external procedures, mat_solve with
! its 93 dummy arguments and getSolverParams, solving on an MI355X through
! libfdmgpu's C ABI (include/fdmgpu.h) with ISO_C_BINDING.
!
! mat_solve assembles every row with the host's own calcMatRow
! (mat.f90:1034), exactly as the Gaussian and MUMPS solvers do, keeps the
! non-zero entries of each dense row buffer abvr/abvi(lb1:lb2) in compressed
! rows, and hands them with lb1/lb2 and the right-hand sides smr/smi (to which
! calcMatRow has added its incoming-wave terms) to fdmgpu_mat_solve.  The
! solution comes back in smr/smi.  A vanishing pivot produces the reference's
! write_error + message + STOP (mat_solve_gaussian.f90:266-273, :289-295).
! Time_fill and Time_tria accumulate assembly and device time.

subroutine mat_solve(Base_hexa, Basereel, Bessel, Besselr, Cal_comp, cgrad, clapl, E_comp, Eimag, Enervide, gradvr, &
        ianew, iato, ibord, icheck, ie, igrph, ii, irregular, isbord, iso, ispinin, isrt, isvois, ivois, Kar, Kari, lato, &
        lb1, lb2, lmaxso, lso, &
        mato, MPI_host_num_for_mumps, mpirank0, mso, n_ato_RI, n_RI, natome, nbm, nbord, nbordf, nbtm, Neuman, Neumanr, &
        new, newinv, ngrph, nicm, nim, nligne, nligne_i, nligneso, nlmsam,  nlmagm, nlmmax, nlmomax, nlmsa, nlmso, nlmso_i, &
        nphiato1, nphiato7, npoint, npsom, nsm, nso1, nsort, nsort_c, nsort_r, nsortf, nspin, nspino, nspinp, nspinr, nstm, &
        numia, nvois, phiato, poidsa, poidso, Relativiste, Repres_comp, rvol, smi, smr, Spinorbite, Test_reg, Time_fill, &
        Time_tria, Vr, Ylm_comp, Ylmato, Ylmso)

  use declarations
  use iso_c_binding
  implicit none

  interface
    integer(c_int) function fdmgpu_mat_solve(nligne, nlmso, cal_comp, lb1, lb2, row_ptr, col, a_re, a_im, smr, smi, &
                                             fail_row, t_fill, t_tria) bind(C, name = 'fdmgpu_mat_solve')
      import :: c_int, c_int64_t, c_double
      integer(c_int), value:: nligne, nlmso, cal_comp
      integer(c_int), dimension(*):: lb1, lb2, col
      integer(c_int64_t), dimension(*):: row_ptr
      real(c_double), dimension(*):: a_re, a_im, smr, smi
      integer(c_int):: fail_row
      real(c_double):: t_fill, t_tria
    end function
    function fdmgpu_last_error() bind(C, name = 'fdmgpu_last_error')
      import :: c_ptr
      type(c_ptr):: fdmgpu_last_error
    end function
  end interface

  ! Dummy arguments, declared as in mat_solve_gaussian.f90:17-65
  integer:: i, ia, icheck, ie, igrph, ii, ipr, ispin, ispinin, lb1i, lb1r, lb2i, lb2r, lm, &
    lmaxso, MPI_host_num_for_mumps, mpirank0, n_ato_RI, n_RI, natome, nbm, nbtm, ngrph, nicm, nim, nligne, nligne_i, &
    nligneso, nlmagm, nlmmax, nlmomax, nlmsam, nlmso, nlmso_i, nphiato1, nphiato7, npoint, &
    npsom, nsm, nso1, nsort, nsort_c, nsort_r, nsortf, nspin, nspino, nspinp, nspinr, nstm, nvois, Test_reg

  integer, dimension(0:npoint):: new
  integer, dimension(natome):: ianew, nbord, nbordf, nlmsa
  integer, dimension(nstm):: isrt
  integer, dimension(npsom):: numia
  integer, dimension(nso1,ngrph):: lso, mso, iso
  integer, dimension(nlmsam,natome,ngrph):: lato, mato, iato
  integer, dimension(nbtm,natome):: ibord, isbord
  integer, dimension(npsom,nvois):: ivois, isvois
  integer, dimension(nligne):: lb1, lb2, newinv

  character(len=2), dimension(nligne):: mletl

  complex(kind=db), dimension(nsort_c,0:lmaxso,nspinr):: Bessel, Neuman

  logical:: Base_hexa, Basereel, Cal_comp, E_comp, irregular, Relativiste, Repres_comp, Spinorbite, Ylm_comp

  real(kind=sg):: time

  real(kind=db):: Enervide, Eimag, tp1, tp2, Time_fill, Time_tria

  real(kind=db), dimension(nopsm,nspino):: Kar, Kari
  real(kind=db), dimension(nvois):: cgrad
  real(kind=db), dimension(0:nvois):: clapl
  real(kind=db), dimension(npoint,nspin):: Vr
  real(kind=db), dimension(nbm,natome):: poidsa
  real(kind=db), dimension(nim):: rvol
  real(kind=db), dimension(nsm):: poidso
  real(kind=db), dimension(nsort,nlmomax):: Ylmso
  real(kind=db), dimension(nicm,3,nspino):: gradvr
  real(kind=db), dimension(nbtm,nlmmax,natome):: Ylmato
  real(kind=db), dimension(nphiato1,nlmagm,nspinp,nspino,natome,nphiato7):: phiato
  real(kind=db), dimension(nlmso,nligne):: smr
  real(kind=db), dimension(nlmso_i,nligne_i):: smi
  real(kind=db), dimension(nsort_r,0:lmaxso,nspinr):: Besselr, Neumanr

  ! Compressed rows gathered from calcMatRow
  integer(c_int):: rc, fail_row, cal
  integer(c_int64_t):: nnz, cap, k
  integer(c_int64_t), dimension(:), allocatable:: row_ptr
  integer(c_int), dimension(:), allocatable:: col, col_t
  integer(c_int), dimension(:), allocatable:: lb1c, lb2c
  real(c_double), dimension(:), allocatable:: a_re, a_im, re_t, im_t
  real(c_double), dimension(1):: dummy
  real(c_double):: t_fill, t_tria
  real(kind=db), dimension(:), allocatable:: abvi, abvr

  if( Spinorbite ) then
    ispin = 1
  else
    ispin = ispinin
  endif

  call CPU_TIME(time)
  tp1 = real(time,db)

  cap = 16_c_int64_t * nligne
  allocate( row_ptr(nligne+1), col(cap), a_re(cap), a_im(cap) )
  row_ptr(1) = 0
  nnz = 0

  do ii = 1,nligne
    lb1r = lb1(ii)
    lb2r = lb2(ii)
    if( Cal_comp ) then
      lb1i = lb1(ii)
      lb2i = lb2(ii)
    else
      lb1i = 0
      lb2i = 0
    endif
    allocate( abvr(lb1r:lb2r) )
    abvr(:) = 0._db
    allocate( abvi(lb1i:lb2i) )
    abvi(:) = 0._db

    call calcMatRow( abvr, abvi, Base_hexa, Basereel, Bessel, Besselr, Cal_comp, cgrad, clapl, E_comp, Eimag, &
      Enervide, gradvr, ianew, iato, ibord, icheck, igrph, ii, irregular, isbord, iso, ispin, isrt, isvois, ivois, Kar, &
      Kari, lato, lb1i, lb1r, lb2i, lb2r, lmaxso, lso, mato, mletl, mso, n_ato_RI, n_RI, natome, nbm, nbord, &
      nbordf, nbtm, Neuman, Neumanr, new, newinv, ngrph, nicm, nim, nligne, nligne_i, &
      nligneso, nlmagm, nlmmax, nlmomax, nlmsa, nlmsam, nlmso, nlmso_i, nphiato1, nphiato7, npoint, npsom, nsm, nso1, &
      nsort, nsort_c, nsort_r, nsortf, nspin, nspino, nspinp, nspinr, nstm, &
      numia, nvois, phiato, poidsa, poidso, Relativiste, Repres_comp, rvol, Spinorbite,  &
      smi, smr, Test_reg, Vr, Ylm_comp, Ylmato, Ylmso, .false. )

    do i = lb1r,lb2r
      if( abvr(i) == 0._db ) then
        if( .not. Cal_comp ) cycle
        if( abvi(i) == 0._db ) cycle
      endif
      if( nnz == cap ) then
        cap = 2 * cap
        allocate( col_t(cap), re_t(cap), im_t(cap) )
        col_t(1:nnz) = col(1:nnz)
        re_t(1:nnz) = a_re(1:nnz)
        im_t(1:nnz) = a_im(1:nnz)
        call move_alloc( col_t, col )
        call move_alloc( re_t, a_re )
        call move_alloc( im_t, a_im )
      endif
      nnz = nnz + 1
      col(nnz) = i
      a_re(nnz) = abvr(i)
      if( Cal_comp ) then
        a_im(nnz) = abvi(i)
      else
        a_im(nnz) = 0._db
      endif
    end do
    row_ptr(ii+1) = nnz

    deallocate( abvr, abvi )
  end do

  call CPU_TIME(time)
  tp2 = real(time,db)
  Time_fill = Time_fill + tp2 - tp1

  allocate( lb1c(nligne), lb2c(nligne) )
  lb1c(:) = lb1(:)
  lb2c(:) = lb2(:)
  if( Cal_comp ) then
    cal = 1
  else
    cal = 0
  endif
  fail_row = 0
  t_fill = 0._db
  t_tria = 0._db
  if( Cal_comp ) then
    rc = fdmgpu_mat_solve(nligne, nlmso, cal, lb1c, lb2c, row_ptr, col, a_re, a_im, smr, smi, fail_row, t_fill, t_tria)
  else
    rc = fdmgpu_mat_solve(nligne, nlmso, cal, lb1c, lb2c, row_ptr, col, a_re, a_im, smr, dummy, fail_row, t_fill, &
                          t_tria)
  endif
  Time_fill = Time_fill + t_fill
  Time_tria = Time_tria + t_tria

  if( rc == 1 .and. mpirank0 == 0 ) then
    ii = fail_row
    i = newinv(ii)
    ia = 0
    lm = 0
    call write_error
    do ipr = 6,9,3
      write(ipr,140) ii, i, ispin, ia, lm
    end do
    stop
  elseif( rc /= 0 ) then
    call write_error
    do ipr = 6,9,3
      write(ipr,150) rc
    end do
    stop
  endif

  deallocate( row_ptr, col, a_re, a_im, lb1c, lb2c )

  return
  140  format(/' Division par zero dans mat',// &
               '  ii =',i6,', i =',i6,', ispin =',i3,', ia =',i3,', lm =',i3)
  150  format(/' Error in the GPU solver (fdmgpu), code',i4)
end

!***********************************************************************

subroutine getSolverParams(MPI_host_num_for_mumps,mpinodes0,Solver)

  use declarations
  implicit none

  integer:: MPI_host_num_for_mumps, mpinodes0

  Character(len=5):: Solver

  Solver = 'HIP  '
  MPI_host_num_for_mumps = 1

  return
end
