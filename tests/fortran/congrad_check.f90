! congrad_check.f90 -- exercises the C-ABI through the Fortran interface
! module efdc_hip the way a Fortran host calls it: builds the CALPUV2C
! system of a closed NX x NY basin (1-based L indexing, LC for land,
! EFDC/cellmap.f90 numbering), sets a right-hand side from a known
! surface P_EXACT (FPTMP = A * P_EXACT), and solves it with efdc_congrad in
! the FP64 and the mixed-precision modes.  Prints one line per mode and
! exits with status 1 if a solution misses P_EXACT by more than 1e-7.
program congrad_check
  use iso_c_binding
  use efdc_hip
  implicit none
  integer, parameter :: NX = 48, NY = 40
  integer, parameter :: LA = NX*NY + 1, LC = LA + 1, LCM = LC + 1, KC = 1, KCM = 2
  integer(c_int32_t), target :: LWC(LCM), LEC(LCM), LSC(LCM), LNC(LCM), LNWC(LCM), LSEC(LCM)
  real(c_double) :: CCC(LCM), CCS(LCM), CCW(LCM), CCE(LCM), CCN(LCM), CCCI(LCM)
  real(c_double) :: FPTMP(LCM), P(LCM), PEX(LCM), BACK(LCM)
  type(efdc_grid_t) :: GR
  type(efdc_params_t) :: PRM
  type(efdc_cg_stats_t) :: ST
  type(c_ptr) :: CTX
  integer(c_int) :: IERR
  integer :: I, J, L, MODE, NFAIL
  real(c_double) :: ERR, X, Y

  NFAIL = 0
  if (efdc_abi_version() < 1) stop 'bad ABI version'
  ! CELLMAP numbering: J outer, I inner; L = 2..LA
  LWC = LC; LEC = LC; LSC = LC; LNC = LC; LNWC = LC; LSEC = LC
  do J = 1, NY
    do I = 1, NX
      L = 2 + (I-1) + (J-1)*NX
      if (I > 1) LWC(L) = L - 1
      if (I < NX) LEC(L) = L + 1
      if (J > 1) LSC(L) = L - NX
      if (J < NY) LNC(L) = L + NX
      if (I > 1 .and. J < NY) LNWC(L) = L + NX - 1
      if (I < NX .and. J > 1) LSEC(L) = L - NX + 1
    end do
  end do
  LWC(1) = LC; LEC(1) = LC; LSC(1) = LC; LNC(1) = LC; LNWC(1) = LC; LSEC(1) = LC
  LWC(LC) = 1; LEC(LC) = 1; LSC(LC) = 1; LNC(LC) = 1; LNWC(LC) = 1; LSEC(LC) = 1

  ! symmetric, diagonally dominant 5-point operator with variable faces
  CCC = 1.0; CCS = 0.0; CCW = 0.0; CCE = 0.0; CCN = 0.0; PEX = 0.0
  do L = 2, LA
    I = mod(L-2, NX) + 1
    J = (L-2)/NX + 1
    X = real(I, c_double)/NX; Y = real(J, c_double)/NY
    if (LWC(L) /= LC) CCW(L) = -(0.5 + 0.4*sin(7.0*X + 3.0*Y))
    if (LSC(L) /= LC) CCS(L) = -(0.5 + 0.4*cos(5.0*X - 2.0*Y))
    PEX(L) = 0.3*sin(6.283185307179586*X)*cos(3.141592653589793*Y) + 0.05*X
  end do
  do L = 2, LA
    if (LEC(L) /= LC) CCE(L) = CCW(LEC(L))
    if (LNC(L) /= LC) CCN(L) = CCS(LNC(L))
  end do
  do L = 2, LA
    CCC(L) = 1.0e-3 - (CCS(L) + CCW(L) + CCE(L) + CCN(L))
  end do
  CCCI = 1.0/CCC
  FPTMP = 0.0
  do L = 2, LA
    FPTMP(L) = CCC(L)*PEX(L) + CCS(L)*PEX(LSC(L)) + CCW(L)*PEX(LWC(L)) &
             + CCE(L)*PEX(LEC(L)) + CCN(L)*PEX(LNC(L))
  end do

  GR%LA = LA; GR%LCM = LCM; GR%KC = KC; GR%KCM = KCM
  GR%LWC = c_loc(LWC); GR%LEC = c_loc(LEC); GR%LSC = c_loc(LSC); GR%LNC = c_loc(LNC)
  GR%LNWC = c_loc(LNWC); GR%LSEC = c_loc(LSEC)

  do MODE = EFDC_CG_FP64, EFDC_CG_MIXED
    call efdc_params_default(PRM)
    PRM%ITERM = 2000; PRM%RSQM = 0.0; PRM%cg_rel_tol = 1.0e-13; PRM%cg_mode = MODE
    IERR = efdc_ctx_create(GR, PRM, CTX)
    if (IERR /= EFDC_OK) then
      write(*,'(A,I3,1X,A)') 'efdc_ctx_create failed', IERR, trim(efdc_error_text(CTX))
      stop 1
    end if
    call put(EFDC_CCC, CCC); call put(EFDC_CCS, CCS); call put(EFDC_CCW, CCW)
    call put(EFDC_CCE, CCE); call put(EFDC_CCN, CCN); call put(EFDC_CCCI, CCCI)
    call put(EFDC_FPTMP, FPTMP)
    P = 0.0
    call put(EFDC_P, P)
    IERR = efdc_congrad(CTX, ST)
    if (IERR /= EFDC_OK) then
      write(*,'(A,I3,1X,A)') 'efdc_congrad failed', IERR, trim(efdc_error_text(CTX))
      stop 1
    end if
    IERR = efdc_get_array(CTX, EFDC_P, P)
    IERR = efdc_get_array(CTX, EFDC_CCW, BACK)
    ERR = maxval(abs(P(2:LA) - PEX(2:LA)))
    if (maxval(abs(BACK(2:LA) - CCW(2:LA))) /= 0.0) NFAIL = NFAIL + 1
    if (.not. (ERR <= 1.0e-7)) NFAIL = NFAIL + 1
    write(*,'(A,I2,A,I6,A,I4,A,ES10.3)') 'FORTRAN_CONGRAD mode=', MODE, ' iters=', ST%iters, &
        ' outer=', ST%outer, ' maxerr=', ERR
    IERR = efdc_ctx_destroy(CTX)
  end do
  if (NFAIL > 0) then
    write(*,'(A,I3)') 'FORTRAN_CONGRAD FAILED checks: ', NFAIL
    stop 1
  end if
  write(*,'(A)') 'FORTRAN_CONGRAD_OK'
contains
  subroutine put(ID, A)
    integer(c_int32_t), intent(in) :: ID
    real(c_double), intent(in) :: A(*)
    integer(c_int) :: E
    E = efdc_set_array(CTX, ID, A)
    if (E /= EFDC_OK) then
      write(*,'(A,I4,1X,A)') 'efdc_set_array failed for id', ID, trim(efdc_error_text(CTX))
      stop 1
    end if
  end subroutine
end program congrad_check
