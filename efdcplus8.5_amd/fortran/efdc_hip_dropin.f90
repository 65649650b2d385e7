! efdc_hip_dropin.f90 -- drop-in replacements of CALPUV2C, CONGRAD and
! CALCONC for the EFDC+ 8.5 Fortran host.  Link this file (with
! efdc_hip.f90 and -lefdc_hip) in place of calpuv2c.f90, congrad.f90 and
! calconc.f90; the callers (HDMT2T: EFDC/hdmt2t.f90:509,562) are unchanged.
!
! Each routine copies the MODULE GLOBAL inputs it needs to the device,
! runs the hot path there and copies back every output the host reads
! afterwards.  Scope: the 2TL scheme with any ISDRY (no drying, the drying
! with the re-solve loop, or drying and rewetting, with or without the
! draining of isolated cells of NDRYSTP < 0), IGRIDV=0, MDCHH=0, ISGWIE=0,
! ISDYNSTP=0 and sigma layers; salinity, temperature, dye, toxics and
! both sediment kinds transported with the standard scheme (ISCDCA < 4),
! their open-boundary concentrations and point sources; the baroclinic
! term (BSC > 0, with the host's CALEBI) and the implicit bottom and
! vegetation drag (ISITB >= 1, ISVEG > 0) formed on the host; CALHEAT,
! the sediment totals and bed model (SSEDTOX) and the two-level balances
! (ISBAL >= 1) called on the host around the transport.  Other settings
! stop at the first call with a message, like the reference's own
! unsupported-option stops.
module efdc_hip_state
  use iso_c_binding
  implicit none
  type(c_ptr), save :: HCTX = c_null_ptr
  integer, save :: NSLOT = 0
  integer, save :: SLOT_KIND(16) = 0     ! MVAR: 1 SAL, 2 TEM, 3 DYE, 5 TOX, 6 SED, 7 SND
  integer, save :: SLOT_NS(16) = 0       ! class of a TOX, SED or SND slot
  integer, save :: SLOT_MO(16) = 0       ! its MO index (the CB, CQS, CSERT column)
  integer(c_int32_t), allocatable, target, save :: MW(:), ME(:), MS(:), MN(:), MNW(:), MSE(:)
  real(c_double), allocatable, target, save :: TCPH(:)
  integer(c_int32_t), allocatable, target, save :: MASKI(:), DRYI(:)
  ! the pressure open-boundary tables (EFDC_HIP_OPENBC)
  integer(c_int32_t), allocatable, target, save :: LW_(:), LE_(:), LS_(:), LN_(:)
  integer(c_int32_t), allocatable, target, save :: IW_(:), IE_(:), IS_(:), IN_(:)
  real(c_double), allocatable, target, save :: SW_(:), SE_(:), SS_(:), SN_(:)
  real(c_double), allocatable, target, save :: CW_(:,:), CE_(:,:), CS_(:,:), CN_(:,:)
  real(c_double), allocatable, target, save :: QW_(:,:), QE_(:,:), QS_(:,:), QN_(:,:)
  ! the concentration-BC and source tables (EFDC_HIP_TABLES)
  integer(c_int32_t), allocatable, target, save :: HXLB(:,:), HXNTB(:,:), HXLQ(:)
  real(c_double), allocatable, target, save :: HXCB(:,:,:,:), HXCS(:,:,:,:), HXWT(:,:)
  real(c_double), allocatable, target, save :: HXQS(:,:), HXQC(:,:), HXCQ(:,:,:), HXCT(:,:,:)
end module efdc_hip_state

! End of the run: the device context and every table of the drop-ins are
! released (the next call of a drop-in starts over from MODULE GLOBAL).
subroutine EFDC_HIP_FINISH
  use efdc_hip
  use efdc_hip_state
  implicit none
  integer(c_int) :: HXERR
  if (c_associated(HCTX)) then
    HXERR = efdc_ctx_destroy(HCTX)
    HCTX = c_null_ptr
  end if
  NSLOT = 0
  if (allocated(MW)) deallocate(MW, ME, MS, MN, MNW, MSE)
  if (allocated(MASKI)) deallocate(MASKI, DRYI)
  if (allocated(TCPH)) deallocate(TCPH)
  if (allocated(LW_)) deallocate(LW_, LE_, LS_, LN_, IW_, IE_, IS_, IN_, SW_, SE_, SS_, SN_, &
                                 CW_, CE_, CS_, CN_, QW_, QE_, QS_, QN_)
  if (allocated(HXLB)) deallocate(HXLB, HXNTB, HXCB, HXCS, HXWT, HXLQ, HXQS, HXQC, HXCQ, HXCT)
end subroutine EFDC_HIP_FINISH

subroutine EFDC_HIP_CHECK(HXERR, WHERE)
  use efdc_hip
  use efdc_hip_state
  implicit none
  integer(c_int), intent(in) :: HXERR
  character(*), intent(in) :: WHERE
  if (HXERR /= EFDC_OK) then
    write(6,'(A,A,A,I3,A,A)') ' EFDC_HIP: ', WHERE, ' failed, status ', HXERR, ': ', &
         trim(efdc_error_text(HCTX))
    stop 'EFDC_HIP ERROR'
  end if
end subroutine

subroutine EFDC_HIP_INIT
  use GLOBAL
  use efdc_hip
  use efdc_hip_state
  implicit none
  type(efdc_grid_t) :: HXGRID
  type(efdc_params_t) :: HXPRM
  type(efdc_constituent_t) :: HXD
  integer(c_int) :: HXERR
  integer(c_int32_t) :: HXSLOT
  integer :: NS

  if (ISDRY < 0 .or. IGRIDV /= 0 .or. MDCHH /= 0 .or. ISGWIE /= 0 .or. ISDYNSTP /= 0) then
    stop 'EFDC_HIP: needs IGRIDV=0, MDCHH=0, ISGWIE=0, ISDYNSTP=0'
  end if
  do NS = 1, 7
    if (ISTRAN(NS) >= 1 .and. NS /= 4 .and. ISCDCA(NS) >= 4) &
      stop 'EFDC_HIP: only the standard transport scheme (ISCDCA < 4) is built'
  end do
  if (ISTRAN(4) >= 1) stop 'EFDC_HIP: shellfish larvae (ISTRAN(4)) are not transported here'
  if (ISTRAN(2) >= 1 .and. ISICE == 4) stop 'EFDC_HIP: frazil ice transport (ISICE=4) is not built'
  if (NTOX + NSED + NSND + 3 > 16) stop 'EFDC_HIP: at most 16 transported constituents'
  allocate(MW(LCM), ME(LCM), MS(LCM), MN(LCM), MNW(LCM), MSE(LCM))
  allocate(MASKI(LCM), DRYI(LCM))
  MW = LWC(1:LCM); ME = LEC(1:LCM); MS = LSC(1:LCM); MN = LNC(1:LCM)
  MNW = LNWC(1:LCM); MSE = LSEC(1:LCM)
  HXGRID%LA = LA; HXGRID%LCM = LCM; HXGRID%KC = KC; HXGRID%KCM = KCM
  HXGRID%LWC = c_loc(MW); HXGRID%LEC = c_loc(ME); HXGRID%LSC = c_loc(MS); HXGRID%LNC = c_loc(MN)
  HXGRID%LNWC = c_loc(MNW); HXGRID%LSEC = c_loc(MSE)
  call efdc_params_default(HXPRM)
  HXPRM%G = G; HXPRM%RITB1 = RITB1; HXPRM%SNLT = SNLT; HXPRM%HDRY = HDRY
  HXPRM%TCON_TBEGIN = TCON*TBEGIN; HXPRM%ITERM = ITERM; HXPRM%RSQM = RSQM
  HXPRM%ISDRY = ISDRY; HXPRM%NDRYSTP = NDRYSTP
  ! implicit bottom / vegetation drag: RCX/RCY formed here each step
  ! (EFDC/calpuv2c.f90:223-243) scale FUHDYE/FVHDXE on the device
  HXPRM%implicit_drag = 0
  if (ISITB >= 1 .or. ISVEG > 0) HXPRM%implicit_drag = 1
  HXERR = efdc_ctx_create(HXGRID, HXPRM, HCTX)
  call EFDC_HIP_CHECK(HXERR, 'efdc_ctx_create')
  ! geometry and layer constants (EFDC/aaefdc.f90:1212-1232,2172-2185)
  HXERR = efdc_set_array(HCTX, EFDC_DXYP, DXYP);   call EFDC_HIP_CHECK(HXERR, 'DXYP')
  HXERR = efdc_set_array(HCTX, EFDC_DXYIP, DXYIP); call EFDC_HIP_CHECK(HXERR, 'DXYIP')
  HXERR = efdc_set_array(HCTX, EFDC_DXYU, DXYU);   call EFDC_HIP_CHECK(HXERR, 'DXYU')
  HXERR = efdc_set_array(HCTX, EFDC_DXYV, DXYV);   call EFDC_HIP_CHECK(HXERR, 'DXYV')
  HXERR = efdc_set_array(HCTX, EFDC_DXIU, DXIU);   call EFDC_HIP_CHECK(HXERR, 'DXIU')
  HXERR = efdc_set_array(HCTX, EFDC_DYIU, DYIU);   call EFDC_HIP_CHECK(HXERR, 'DYIU')
  HXERR = efdc_set_array(HCTX, EFDC_DXIV, DXIV);   call EFDC_HIP_CHECK(HXERR, 'DXIV')
  HXERR = efdc_set_array(HCTX, EFDC_DYIV, DYIV);   call EFDC_HIP_CHECK(HXERR, 'DYIV')
  HXERR = efdc_set_array(HCTX, EFDC_HRUO, HRUO);   call EFDC_HIP_CHECK(HXERR, 'HRUO')
  HXERR = efdc_set_array(HCTX, EFDC_HRVO, HRVO);   call EFDC_HIP_CHECK(HXERR, 'HRVO')
  HXERR = efdc_set_array(HCTX, EFDC_SUB, SUB);     call EFDC_HIP_CHECK(HXERR, 'SUB')
  HXERR = efdc_set_array(HCTX, EFDC_SVB, SVB);     call EFDC_HIP_CHECK(HXERR, 'SVB')
  if (ISDRY > 0) then
    ! the switches the faces reopen to at each step start (calpuv2c.f90:246-253)
    HXERR = efdc_set_array(HCTX, EFDC_SUBO, SUBO); call EFDC_HIP_CHECK(HXERR, 'SUBO')
    HXERR = efdc_set_array(HCTX, EFDC_SVBO, SVBO); call EFDC_HIP_CHECK(HXERR, 'SVBO')
  end if
  if (ISDRY > 0 .and. NDRYSTP < 0) then
    ! drained volume and isolation counters continue from the host's
    ! (a restart's) values (calpuv2c.f90:838-893)
    DRYI = 0
    DRYI(2:LA) = NATDRY(2:LA)
    HXERR = efdc_set_dry_counters(HCTX, VDWASTE, DRYI)
    call EFDC_HIP_CHECK(HXERR, 'efdc_set_dry_counters')
  end if
  HXERR = efdc_set_array(HCTX, EFDC_BELV, BELV);   call EFDC_HIP_CHECK(HXERR, 'BELV')
  HXERR = efdc_set_array(HCTX, EFDC_DZC, DZC);     call EFDC_HIP_CHECK(HXERR, 'DZC')
  HXERR = efdc_set_array(HCTX, EFDC_CDZKK, CDZKK); call EFDC_HIP_CHECK(HXERR, 'CDZKK')
  HXERR = efdc_set_array(HCTX, EFDC_CDZKMK, CDZKMK); call EFDC_HIP_CHECK(HXERR, 'CDZKMK')
  HXERR = efdc_set_array(HCTX, EFDC_DZIG, DZIG);   call EFDC_HIP_CHECK(HXERR, 'DZIG')
  if (ISDRY > 0) then
    ! with drying the library scales the full-grid layer switches by SUB/SVB
    HXERR = efdc_set_array(HCTX, EFDC_SUB3D, SUB3DO); call EFDC_HIP_CHECK(HXERR, 'SUB3DO')
    HXERR = efdc_set_array(HCTX, EFDC_SVB3D, SVB3DO); call EFDC_HIP_CHECK(HXERR, 'SVB3DO')
  else
    HXERR = efdc_set_array(HCTX, EFDC_SUB3D, SUB3D); call EFDC_HIP_CHECK(HXERR, 'SUB3D')
    HXERR = efdc_set_array(HCTX, EFDC_SVB3D, SVB3D); call EFDC_HIP_CHECK(HXERR, 'SVB3D')
  end if
  ! transported constituents in the order CALCONC calls CALTRAN
  ! (EFDC/calconc.f90:159-228)
  NSLOT = 0
  if (ISTRAN(1) == 1) call ADD(1, 1, 0)
  if (ISTRAN(2) == 1) call ADD(2, 2, 0)
  if (ISTRAN(3) == 1) call ADD(3, 3, 0)
  if (ISTRAN(5) == 1) then
    do NS = 1, NTOX
      call ADD(5, MSVTOX(NS), NS)
    end do
  end if
  if (ISTRAN(6) == 1) then
    do NS = 1, NSED
      call ADD(6, MSVSED(NS), NS)
    end do
  end if
  if (ISTRAN(7) == 1) then
    do NS = 1, NSND
      call ADD(7, MSVSND(NS), NS)
    end do
  end if
contains
  subroutine ADD(MVAR, MO, NS_)
    integer, intent(in) :: MVAR, MO, NS_
    HXD%mo = MO; HXD%isadac = ISADAC(MVAR); HXD%isfct = ISFCT(MVAR)
    HXD%rkdye = 0.0; HXD%is_dye = 0
    ! settling of the sediment classes stays with the host's SSEDTOX, which
    ! also owns the bed exchange
    HXD%wset = 0.0; HXD%probdep = 0.0; HXD%wese = 0.0
    if (MVAR == 3) then
      HXD%rkdye = RKDYE; HXD%is_dye = 1
    end if
    HXERR = efdc_add_constituent(HCTX, HXD, HXSLOT)
    call EFDC_HIP_CHECK(HXERR, 'efdc_add_constituent')
    NSLOT = NSLOT + 1
    SLOT_KIND(NSLOT) = MVAR
    SLOT_NS(NSLOT) = NS_
    SLOT_MO(NSLOT) = MO
  end subroutine
end subroutine EFDC_HIP_INIT

! The concentration open boundaries (CALTRAN, EFDC/caltran.f90:302-471)
! and the point sources (CALFQC) of this step, gathered from MODULE GLOBAL
! for the registered constituents: the inflow concentrations CB*, the
! series values CSERT of each boundary cell and source, the ramp lengths
! NTSCR*, the layer weights WTCI, and QSS, QSERCELL, CQS.  The recovery
! memory CLO*/NLO* stays on the device.
subroutine EFDC_HIP_TABLES
  use GLOBAL
  use efdc_hip
  use efdc_hip_state
  implicit none
  type(efdc_conc_bc_t) :: HXC
  type(efdc_sources_t) :: HXS
  integer :: NB(4), NBM, LL, K, J, M, S, NQ
  integer(c_int) :: HXERR

  if (NSLOT == 0) return
  NB = [NCBS, NCBW, NCBE, NCBN]
  NBM = max(maxval(NB), 1)
  NQ = NQSIJ
  if (.not. allocated(HXLB)) then
    allocate(HXLB(NBM,4), HXNTB(NBM,4), HXCB(NSLOT,2,NBM,4), HXCS(NSLOT,KC,NBM,4), HXWT(2,KC))
    allocate(HXLQ(max(NQ,1)), HXQS(KC,max(NQ,1)), HXQC(KC,max(NQ,1)))
    allocate(HXCQ(NSLOT,KC,max(NQ,1)), HXCT(NSLOT,KC,max(NQ,1)))
  end if
  if (sum(NB) > 0) then
    do S = 1, 4
      do LL = 1, NB(S)
        select case (S)
        case (1)
          HXLB(LL,S) = LCBS(LL); HXNTB(LL,S) = NTSCRS(LL)
        case (2)
          HXLB(LL,S) = LCBW(LL); HXNTB(LL,S) = NTSCRW(LL)
        case (3)
          HXLB(LL,S) = LCBE(LL); HXNTB(LL,S) = NTSCRE(LL)
        case (4)
          HXLB(LL,S) = LCBN(LL); HXNTB(LL,S) = NTSCRN(LL)
        end select
        do J = 1, NSLOT
          M = SLOT_MO(J)
          select case (S)
          case (1)
            HXCB(J,1:2,LL,S) = CBS(LL,1:2,M)
            HXCS(J,1:KC,LL,S) = CSERT(1:KC,NCSERS(LL,M),M)
          case (2)
            HXCB(J,1:2,LL,S) = CBW(LL,1:2,M)
            HXCS(J,1:KC,LL,S) = CSERT(1:KC,NCSERW(LL,M),M)
          case (3)
            HXCB(J,1:2,LL,S) = CBE(LL,1:2,M)
            HXCS(J,1:KC,LL,S) = CSERT(1:KC,NCSERE(LL,M),M)
          case (4)
            HXCB(J,1:2,LL,S) = CBN(LL,1:2,M)
            HXCS(J,1:KC,LL,S) = CSERT(1:KC,NCSERN(LL,M),M)
          end select
        end do
      end do
      HXC%side(S) = efdc_cbc_side_t(NB(S), c_loc(HXLB(1,S)), c_loc(HXCB(1,1,1,S)), &
                                    c_loc(HXCS(1,1,1,S)), c_loc(HXNTB(1,S)))
    end do
    do K = 1, KC
      HXWT(1,K) = WTCI(K,1)
      HXWT(2,K) = WTCI(K,2)
    end do
    HXC%wtci = c_loc(HXWT)
    HXERR = efdc_set_conc_bc(HCTX, HXC)
    call EFDC_HIP_CHECK(HXERR, 'efdc_set_conc_bc')
  end if
  if (NQ > 0) then
    do LL = 1, NQ
      HXLQ(LL) = LQS(LL)
      do K = 1, KC
        HXQS(K,LL) = QSS(K,LL)
        HXQC(K,LL) = QSERCELL(K,LL)
        do J = 1, NSLOT
          M = SLOT_MO(J)
          HXCQ(J,K,LL) = CQS(K,LL,M)
          HXCT(J,K,LL) = CSERT(K,NCSERQ(LL,M),M)
        end do
      end do
    end do
    HXS = efdc_sources_t(NQ, c_loc(HXLQ), c_loc(HXQS), c_loc(HXQC), c_loc(HXCQ), c_loc(HXCT))
    HXERR = efdc_set_sources(HCTX, HXS)
    call EFDC_HIP_CHECK(HXERR, 'efdc_set_sources')
  end if
end subroutine EFDC_HIP_TABLES

! Pressure open boundaries of this step (the time-series part of
! EFDC/setopenbc.f90:44-48 is formed on the host; the tide is added on the
! device).
subroutine EFDC_HIP_OPENBC
  use GLOBAL
  use efdc_hip
  use efdc_hip_state
  implicit none
  type(efdc_open_bc_t) :: HXBC
  integer(c_int) :: HXERR
  integer :: LL, M

  if (.not. allocated(LW_)) then
    allocate(LW_(max(NPBW,1)), LE_(max(NPBE,1)), LS_(max(NPBS,1)), LN_(max(NPBN,1)))
    allocate(IW_(max(NPBW,1)), IE_(max(NPBE,1)), IS_(max(NPBS,1)), IN_(max(NPBN,1)))
    allocate(SW_(max(NPBW,1)), SE_(max(NPBE,1)), SS_(max(NPBS,1)), SN_(max(NPBN,1)))
    allocate(CW_(max(MTIDE,1),max(NPBW,1)), CE_(max(MTIDE,1),max(NPBE,1)))
    allocate(CS_(max(MTIDE,1),max(NPBS,1)), CN_(max(MTIDE,1),max(NPBN,1)))
    allocate(QW_(max(MTIDE,1),max(NPBW,1)), QE_(max(MTIDE,1),max(NPBE,1)))
    allocate(QS_(max(MTIDE,1),max(NPBS,1)), QN_(max(MTIDE,1),max(NPBN,1)))
    allocate(TCPH(max(MTIDE,1)))
    TCPH = 0.0
    if (MTIDE > 0) TCPH(1:MTIDE) = TCP(1:MTIDE)
  end if
  do LL = 1, NPBW
    LW_(LL) = LPBW(LL); IW_(LL) = ISPBW(LL)
    SW_(LL) = PSERT(NPSERW(LL)) + 0.5*PSERZDF(NPSERW(LL)) + PSERST(NPSERW(LL)) + 0.5*PSERZDS(NPSERW(LL))
    do M = 1, MTIDE
      CW_(M,LL) = PCBW(LL,M); QW_(M,LL) = PSBW(LL,M)
    end do
  end do
  do LL = 1, NPBE
    LE_(LL) = LPBE(LL); IE_(LL) = ISPBE(LL)
    SE_(LL) = PSERT(NPSERE(LL)) + 0.5*PSERZDF(NPSERE(LL)) + PSERST(NPSERE(LL)) + 0.5*PSERZDS(NPSERE(LL))
    do M = 1, MTIDE
      CE_(M,LL) = PCBE(LL,M); QE_(M,LL) = PSBE(LL,M)
    end do
  end do
  do LL = 1, NPBS
    LS_(LL) = LPBS(LL); IS_(LL) = ISPBS(LL)
    SS_(LL) = PSERT(NPSERS(LL)) + 0.5*PSERZDF(NPSERS(LL)) + PSERST(NPSERS(LL)) + 0.5*PSERZDS(NPSERS(LL))
    do M = 1, MTIDE
      CS_(M,LL) = PCBS(LL,M); QS_(M,LL) = PSBS(LL,M)
    end do
  end do
  do LL = 1, NPBN
    LN_(LL) = LPBN(LL); IN_(LL) = ISPBN(LL)
    SN_(LL) = PSERT(NPSERN(LL)) + 0.5*PSERZDF(NPSERN(LL)) + PSERST(NPSERN(LL)) + 0.5*PSERZDS(NPSERN(LL))
    do M = 1, MTIDE
      CN_(M,LL) = PCBN(LL,M); QN_(M,LL) = PSBN(LL,M)
    end do
  end do
  if (NPFORT >= 1) stop 'EFDC_HIP: NPFORT>=1 open-boundary blending is formed by the caller'
  HXBC%side(1) = efdc_pbc_side_t(NPBW, c_loc(LW_), c_loc(IW_), c_loc(SW_), c_loc(CW_), c_loc(QW_))
  HXBC%side(2) = efdc_pbc_side_t(NPBE, c_loc(LE_), c_loc(IE_), c_loc(SE_), c_loc(CE_), c_loc(QE_))
  HXBC%side(3) = efdc_pbc_side_t(NPBS, c_loc(LS_), c_loc(IS_), c_loc(SS_), c_loc(CS_), c_loc(QS_))
  HXBC%side(4) = efdc_pbc_side_t(NPBN, c_loc(LN_), c_loc(IN_), c_loc(SN_), c_loc(CN_), c_loc(QN_))
  HXBC%mtide = MTIDE
  HXBC%tcp = c_loc(TCPH)
  HXERR = efdc_set_open_bc(HCTX, HXBC)
  call EFDC_HIP_CHECK(HXERR, 'efdc_set_open_bc')
end subroutine EFDC_HIP_OPENBC

SUBROUTINE CALPUV2C
  use GLOBAL
  use efdc_hip
  use efdc_hip_state
  implicit none
  type(efdc_step_t) :: HXSTEP
  type(efdc_cg_stats_t) :: HXST
  integer(c_int) :: HXERR
  integer(c_int32_t) :: HXNCOR
  integer :: L, K, ND, LF, LL, LP, NW, LW, LSX
  real(8) :: TMPX, TMPY

  if (.not. c_associated(HCTX)) call EFDC_HIP_INIT
  if (NBCSOP > 0) call EFDC_HIP_OPENBC
  DELT = DT; DELTI = 1./DELT; ISTL = 2                ! EFDC/calpuv2c.f90:128-137
  if (BSC > 1.E-6) then
    ! external buoyancy integrals at level N and the explicit external
    ! density gradients (EFDC/calpuv2c.f90:161,192-201, IGRIDV = 0)
    call CALEBI
    do L = 2, LA
      LW = LWC(L)
      FPGXE(L) = -SBX(L)*HU(L)*GP*((BI2W(L)+BI2W(LW))*(HP(L)-HP(LW)) + &
                 2.0*HU(L)*(BI1W(L)-BI1W(LW)) + (BEW(L)+BEW(LW))*(BELV(L)-BELV(LW)))
      LSX = LSC(L)
      FPGYE(L) = -SBY(L)*HV(L)*GP*((BI2S(L)+BI2S(LSX))*(HP(L)-HP(LSX)) + &
                 2.0*HV(L)*(BI1S(L)-BI1S(LSX)) + (BES(L)+BES(LSX))*(BELV(L)-BELV(LSX)))
    end do
  end if
  if (ISITB >= 1) then
    ! implicit bottom drag with vegetation (EFDC/calpuv2c.f90:223-233); the
    ! device scales FUHDYE/FVHDXE by them after the explicit terms
    do L = 2, LA
      TMPX = 1.0
      TMPY = 1.0
      if (UHE(L) /= 0.0) TMPX = U(L,KSZU(L))*HU(L)/UHE(L)
      if (VHE(L) /= 0.0) TMPY = V(L,KSZV(L))*HV(L)/VHE(L)
      RCX(L) = 1./(1. + TMPX*RITB*DT*HUI(L)*STBX(L)*SQRT(VU(L)*VU(L) + &
               U(L,KSZU(L))*U(L,KSZU(L))) + DT*FXVEGE(L))
      RCY(L) = 1./(1. + TMPY*RITB*DT*HVI(L)*STBY(L)*SQRT(UV(L)*UV(L) + &
               V(L,KSZV(L))*V(L,KSZV(L))) + DT*FYVEGE(L))
    end do
  else if (ISVEG > 0) then
    ! vegetation drag only (:234-241)
    do L = 2, LA
      RCX(L) = 1./(1. + DT*FXVEGE(L))
      RCY(L) = 1./(1. + DT*FYVEGE(L))
    end do
  end if
  ! inputs produced by the host since the last call
  call PUSH(EFDC_UHDYE, UHDYE); call PUSH(EFDC_VHDXE, VHDXE); call PUSH(EFDC_P, P)
  call PUSH(EFDC_HU, HU); call PUSH(EFDC_HV, HV); call PUSH(EFDC_HUI, HUI)
  call PUSH(EFDC_HVI, HVI); call PUSH(EFDC_HP, HP); call PUSH(EFDC_H1P, H1P)
  call PUSH(EFDC_TSX, TSX); call PUSH(EFDC_TSY, TSY); call PUSH(EFDC_TBX, TBX)
  call PUSH(EFDC_TBY, TBY); call PUSH(EFDC_FCAXE, FCAXE); call PUSH(EFDC_FCAYE, FCAYE)
  call PUSH(EFDC_FPGXE, FPGXE); call PUSH(EFDC_FPGYE, FPGYE); call PUSH(EFDC_FXE, FXE)
  call PUSH(EFDC_FYE, FYE); call PUSH(EFDC_QSUME, QSUME); call PUSH(EFDC_RCX, RCX)
  call PUSH(EFDC_RCY, RCY); call PUSH(EFDC_HPK, HPK)
  if (ISDRY > 0) then
    ! step start with drying on (calpuv2c.f90:169-177,246-253): the mask
    ! of the last step, and the buoyancy switches of the host's momentum
    ! terms reopened like SUB/SVB (done on the device)
    do L = 2, LA
      OLDMASK(L) = LMASKDRY(L)
      SBX(L) = SBXO(L)
      SBY(L) = SBYO(L)
    end do
  end if
  HXSTEP%dt = DT; HXSTEP%n = N; HXSTEP%niter = NITER
  HXERR = efdc_calpuv2c(HCTX, HXSTEP, HXST)
  IDRYTBP = HXST%iters                                ! EFDC/congrad.f90:187
  if (HXERR == EFDC_ERR_CG_MAXITER) then
    ! the reference's failure path (EFDC/congrad.f90:171-184)
    call PULL(EFDC_CCS, CCS); call PULL(EFDC_CCW, CCW); call PULL(EFDC_CCC, CCC)
    call PULL(EFDC_CCE, CCE); call PULL(EFDC_CCN, CCN); call PULL(EFDC_FPTMP, FPTMP)
    open(8, FILE=OUTDIR//'EFDCLOG.OUT', POSITION='APPEND')
    write(6,'(A)') '  MAXIMUM ITERATIONS EXCEEDED IN EXTERNAL SOLUTION'
    write(8,'(A)') '   I    J     CCS     CCW      CCC      CCE      CCN     FPTMP'
    do L = 2, LA
      write(8,'(2I6,6E13.4)') IL(L), JL(L), CCS(L), CCW(L), CCC(L), CCE(L), CCN(L), FPTMP(L)
    end do
    close(8)
    stop
  end if
  if (HXERR /= EFDC_ERR_NEGDEPTH) call EFDC_HIP_CHECK(HXERR, 'efdc_calpuv2c')
  ! outputs read by the host afterwards
  call PULL(EFDC_P, P); call PULL(EFDC_P1, P1); call PULL(EFDC_HP, HP)
  call PULL(EFDC_H1P, H1P); call PULL(EFDC_H2P, H2P); call PULL(EFDC_HU, HU)
  call PULL(EFDC_HV, HV); call PULL(EFDC_HUI, HUI); call PULL(EFDC_HVI, HVI)
  call PULL(EFDC_HPI, HPI); call PULL(EFDC_H1U, H1U); call PULL(EFDC_H1V, H1V)
  call PULL(EFDC_H1UI, H1UI); call PULL(EFDC_H1VI, H1VI); call PULL(EFDC_UHDYE, UHDYE)
  call PULL(EFDC_VHDXE, VHDXE); call PULL(EFDC_UHDY1E, UHDY1E); call PULL(EFDC_VHDX1E, VHDX1E)
  call PULL(EFDC_UHE, UHE); call PULL(EFDC_VHE, VHE); call PULL(EFDC_FUHDYE, FUHDYE)
  call PULL(EFDC_FVHDXE, FVHDXE); call PULL(EFDC_FP, FP); call PULL(EFDC_FP1, FP1)
  call PULL(EFDC_CCC, CCC); call PULL(EFDC_CCCI, CCCI); call PULL(EFDC_FPTMP, FPTMP)
  call PULL(EFDC_H1PK, H1PK)
  if (HXERR == EFDC_ERR_NEGDEPTH) then
    write(6,'(A,F12.4)') ' EFDC_HIP: NEGATIVE DEPTH AFTER THE EXTERNAL SOLUTION AT TIMEDAY ', TIMEDAY
    stop 'NEGDEP'
  end if
  if (ISDRY == 0) then
    LAWET = LA - 1                                    ! EFDC/calpuv2c.f90:1061-1063
    LADRY = 0
  else
    call PULL(EFDC_SUB, SUB); call PULL(EFDC_SVB, SVB)
    if (NDRYSTP < 0) then
      ! drained cells (calpuv2c.f90:838-893,1010-1021): depths of all
      ! levels, the drained rate and volume, the isolation counters
      call PULL(EFDC_HPK, HPK); call PULL(EFDC_HPKI, HPKI)
      call EFDC_HIP_CHECK(efdc_get_dry_waste(HCTX, QDWASTE, VDWASTE, DRYI), 'efdc_get_dry_waste')
      do L = 2, LA
        NATDRY(L) = DRYI(L)
        if (ISBAL == 0) QDWASTE(L) = 0.0
      end do
    end if
    call EFDC_HIP_CHECK(efdc_get_dry_state(HCTX, MASKI, DRYI, HXNCOR), 'efdc_get_dry_state')
    do L = 2, LA
      LMASKDRY(L) = MASKI(L) /= 0
      ISCDRY(L) = DRYI(L)
    end do
    ! the buoyancy switches of the host's momentum terms close and reopen
    ! with the faces (calpuv2c.f90:519-522,589-592,633-665,703-710)
    do L = 2, LA
      if (SUB(L) < 0.5) SBX(L) = 0.
      if (SVB(L) < 0.5) SBY(L) = 0.
    end do
    ! layer switches of the host's internal mode (calpuv2c.f90:982-990)
    do K = 1, KC
      do L = 2, LA
        if (LKSZ(L,K)) cycle
        SUB3D(L,K) = SUB(L)*SUB3DO(L,K)
        SVB3D(L,K) = SVB(L)*SVB3DO(L,K)
      end do
    end do
    ! wet list and newly dry list (calpuv2c.f90:1064-1078), then the
    ! per-domain and whole-domain layer lists (:1096-1115,1170-1185)
    LAWET = 0
    LADRY = 0
    do L = 2, LA
      if (LMASKDRY(L)) then
        LAWET = LAWET + 1
        LWET(LAWET) = L
      else if (OLDMASK(L) .or. N < NTSTBC + 1) then
        LADRY = LADRY + 1
        LDRY(LADRY) = L
      end if
    end do
    LDMWET = INT(FLOAT(LAWET)/FLOAT(NDM)) + 1
    LDMDRY = INT(FLOAT(LADRY)/FLOAT(NDM)) + 1
    do ND = 1, NDM
      LF = (ND - 1)*LDMWET + 1
      LL = MIN(LF + LDMWET - 1, LAWET)
      do K = 1, KC
        NW = 0
        do LP = LF, LL
          L = LWET(LP)
          if (LKSZ(L,K)) cycle
          NW = NW + 1
          LKWET(NW,K,ND) = L
        end do
        LLWET(K,ND) = NW
      end do
      do K = 1, KS
        LLWETZ(K,ND) = LLWET(K,ND)
        LKWETZ(1:LLWET(K,ND),K,ND) = LKWET(1:LLWET(K,ND),K,ND)
      end do
      LLWETZ(KC,ND) = LLWET(KS,ND)
      LKWETZ(1:LLWET(KS,ND),KC,ND) = LKWET(1:LLWET(KS,ND),KS,ND)
    end do
    do K = 1, KC
      NW = 0
      do L = 2, LA
        if (LKSZ(L,K) .or. .not. LMASKDRY(L)) cycle
        NW = NW + 1
        LKWET(NW,K,0) = L
      end do
      LLWET(K,0) = NW
    end do
  end if
contains
  subroutine PUSH(ID, A)
    integer(c_int32_t), intent(in) :: ID
    real(c_double), intent(in) :: A(*)
    call EFDC_HIP_CHECK(efdc_set_array(HCTX, ID, A), 'efdc_set_array')
  end subroutine
  subroutine PULL(ID, A)
    integer(c_int32_t), intent(in) :: ID
    real(c_double), intent(inout) :: A(*)
    call EFDC_HIP_CHECK(efdc_get_array(HCTX, ID, A), 'efdc_get_array')
  end subroutine
END SUBROUTINE CALPUV2C

SUBROUTINE CONGRAD(NOPTIMAL, LDMOPT)
  use GLOBAL
  use efdc_hip
  use efdc_hip_state
  implicit none
  integer, intent(in) :: NOPTIMAL, LDMOPT
  type(efdc_cg_stats_t) :: HXST
  integer(c_int) :: HXERR

  if (.not. c_associated(HCTX)) call EFDC_HIP_INIT
  call EFDC_HIP_CHECK(efdc_set_array(HCTX, EFDC_CCC, CCC), 'CCC')
  call EFDC_HIP_CHECK(efdc_set_array(HCTX, EFDC_CCS, CCS), 'CCS')
  call EFDC_HIP_CHECK(efdc_set_array(HCTX, EFDC_CCW, CCW), 'CCW')
  call EFDC_HIP_CHECK(efdc_set_array(HCTX, EFDC_CCE, CCE), 'CCE')
  call EFDC_HIP_CHECK(efdc_set_array(HCTX, EFDC_CCN, CCN), 'CCN')
  call EFDC_HIP_CHECK(efdc_set_array(HCTX, EFDC_CCCI, CCCI), 'CCCI')
  call EFDC_HIP_CHECK(efdc_set_array(HCTX, EFDC_FPTMP, FPTMP), 'FPTMP')
  call EFDC_HIP_CHECK(efdc_set_array(HCTX, EFDC_P, P), 'P')
  HXERR = efdc_congrad(HCTX, HXST)
  call EFDC_HIP_CHECK(efdc_get_array(HCTX, EFDC_P, P), 'P')
  IDRYTBP = HXST%iters
  if (HXERR == EFDC_ERR_CG_MAXITER) then
    write(6,'(A)') '  MAXIMUM ITERATIONS EXCEEDED IN EXTERNAL SOLUTION'
    stop
  end if
  call EFDC_HIP_CHECK(HXERR, 'efdc_congrad')
END SUBROUTINE CONGRAD

SUBROUTINE CALCONC(ISTL_, IS2TL_)
  use GLOBAL
  use efdc_hip
  use efdc_hip_state
  implicit none
  integer :: ISTL_, IS2TL_
  type(efdc_step_t) :: HXSTEP
  integer :: J, K, L, LP, NS
  real, save :: SEDTIME = 0.0

  if (ISTL_ /= 2 .or. IS2TL_ /= 1) stop 'EFDC_HIP: CALCONC needs ISTL=2, IS2TL=1'
  if (.not. c_associated(HCTX)) call EFDC_HIP_INIT
  DELT = DT
  ! mass balance terms of the sources at the step start (EFDC/calconc.f90:76-80)
  if (ISBAL >= 1) call BAL2T3A
  call EFDC_HIP_TABLES
  call PUSH(EFDC_UHDY2, UHDY2); call PUSH(EFDC_VHDX2, VHDX2); call PUSH(EFDC_W2, W2)
  call PUSH(EFDC_H1PK, H1PK); call PUSH(EFDC_HPK, HPK); call PUSH(EFDC_HPKI, HPKI)
  call PUSH(EFDC_AB, AB); call PUSH(EFDC_HPI, HPI); call PUSH(EFDC_QSUM, QSUM)
  call PUSH(EFDC_U2, U2); call PUSH(EFDC_V2, V2)
  do J = 1, NSLOT
    select case (SLOT_KIND(J))
    case (1)
      call EFDC_HIP_CHECK(efdc_set_constituent(HCTX, J-1, SAL, SAL1), 'SAL')
    case (2)
      call EFDC_HIP_CHECK(efdc_set_constituent(HCTX, J-1, TEM, TEM1), 'TEM')
    case (3)
      call EFDC_HIP_CHECK(efdc_set_constituent(HCTX, J-1, DYE, DYE1), 'DYE')
    case (5)
      call EFDC_HIP_CHECK(efdc_set_constituent(HCTX, J-1, TOX(:,:,SLOT_NS(J)), &
                                              TOX1(:,:,SLOT_NS(J))), 'TOX')
    case (6)
      call EFDC_HIP_CHECK(efdc_set_constituent(HCTX, J-1, SED(:,:,SLOT_NS(J)), &
                                              SED1(:,:,SLOT_NS(J))), 'SED')
    case (7)
      call EFDC_HIP_CHECK(efdc_set_constituent(HCTX, J-1, SND(:,:,SLOT_NS(J)), &
                                              SND1(:,:,SLOT_NS(J))), 'SND')
    end select
  end do
  HXSTEP%dt = DT; HXSTEP%n = N; HXSTEP%niter = NITER
  call EFDC_HIP_CHECK(efdc_calconc(HCTX, HXSTEP), 'efdc_calconc')
  do J = 1, NSLOT
    select case (SLOT_KIND(J))
    case (1)
      call EFDC_HIP_CHECK(efdc_get_constituent(HCTX, J-1, SAL, SAL1), 'SAL')
    case (2)
      call EFDC_HIP_CHECK(efdc_get_constituent(HCTX, J-1, TEM, TEM1), 'TEM')
    case (3)
      call EFDC_HIP_CHECK(efdc_get_constituent(HCTX, J-1, DYE, DYE1), 'DYE')
    case (5)
      call EFDC_HIP_CHECK(efdc_get_constituent(HCTX, J-1, TOX(:,:,SLOT_NS(J)), &
                                              TOX1(:,:,SLOT_NS(J))), 'TOX')
    case (6)
      call EFDC_HIP_CHECK(efdc_get_constituent(HCTX, J-1, SED(:,:,SLOT_NS(J)), &
                                              SED1(:,:,SLOT_NS(J))), 'SED')
    case (7)
      call EFDC_HIP_CHECK(efdc_get_constituent(HCTX, J-1, SND(:,:,SLOT_NS(J)), &
                                              SND1(:,:,SLOT_NS(J))), 'SND')
    end select
  end do
  ! sediment totals after the transport (EFDC/calconc.f90:479-570)
  if (ISTRAN(6) > 0) then
    do K = 1, KC
      do LP = 1, LLWET(K,0)
        L = LKWET(LP,K,0)
        SEDT(L,K) = 0.
        do NS = 1, NSED
          SEDT(L,K) = SEDT(L,K) + SED(L,K,NS)
        end do
      end do
    end do
  end if
  if (ISTRAN(7) > 0) then
    do K = 1, KC
      do LP = 1, LLWET(K,0)
        L = LKWET(LP,K,0)
        SNDT(L,K) = 0.
        do NS = 1, NSND
          SNDT(L,K) = SNDT(L,K) + SND(L,K,NS)
        end do
      end do
    end do
  end if
  ! surface and internal heat sources and sinks (:574-578)
  if (ISTRAN(2) >= 1) call CALHEAT(ISTL_)
  ! sediment and toxics settling, deposition and resuspension stay on the
  ! host with the bed model, on the reference's clock (2TL branch of
  ! EFDC/calconc.f90:621-630)
  if ((ISTRAN(6) >= 1 .or. ISTRAN(7) >= 1) .and. TIMEDAY >= SEDSTART) then
    SEDTIME = SEDTIME + DELT
    if (SEDTIME >= SEDSTEP) then
      DTSED = SEDTIME
      call SSEDTOX
      SEDTIME = 0.0
    end if
  end if
  ! two-level balances: boundary fluxes and bed exchange (:659-664)
  if (ISBAL >= 1) then
    call BAL2T2
    call BAL2T3B(1)
  end if
contains
  subroutine PUSH(ID, A)
    integer(c_int32_t), intent(in) :: ID
    real(c_double), intent(in) :: A(*)
    call EFDC_HIP_CHECK(efdc_set_array(HCTX, ID, A), 'efdc_set_array')
  end subroutine
END SUBROUTINE CALCONC
