! efdc_hip.f90 -- ISO_C_BINDING interface to the C-ABI of include/efdc_hip.h.
!
! The Fortran host (EFDC+ 8.5) calls the gfx950 library through this
! module.  It does not depend on MODULE GLOBAL; the drop-in replacements of
! CALPUV2C, CONGRAD and CALCONC that copy the GLOBAL arrays across the
! boundary are in efdc_hip_dropin.f90.
module efdc_hip
  use iso_c_binding
  implicit none

  ! efdc_status
  integer(c_int), parameter :: EFDC_OK = 0, EFDC_ERR_ARG = 1, EFDC_ERR_HIP = 2, &
      EFDC_ERR_CG_MAXITER = 3, EFDC_ERR_NEGDEPTH = 4, EFDC_ERR_STATE = 5, &
      EFDC_ERR_NOMEM = 6, EFDC_ERR_NODEVICE = 7, EFDC_ERR_IO = 8
  ! efdc_cg_mode, efdc_store
  integer(c_int32_t), parameter :: EFDC_CG_FP64 = 0, EFDC_CG_MIXED = 1
  integer(c_int32_t), parameter :: EFDC_STORE_FP64 = 0, EFDC_STORE_FP32 = 1
  ! efdc_cg_kernel
  integer(c_int32_t), parameter :: EFDC_CG_KERNEL_AUTO = 0, EFDC_CG_KERNEL_MULTI = 1, &
      EFDC_CG_KERNEL_PERSISTENT = 2

  ! efdc_array_id: (LCM) arrays
  integer(c_int32_t), parameter :: EFDC_DXYP = 0, EFDC_DXYIP = 1, EFDC_DXYU = 2, &
      EFDC_DXYV = 3, EFDC_DXIU = 4, EFDC_DYIU = 5, EFDC_DXIV = 6, EFDC_DYIV = 7, &
      EFDC_HRUO = 8, EFDC_HRVO = 9, EFDC_SUB = 10, EFDC_SVB = 11, EFDC_BELV = 12, &
      EFDC_P = 13, EFDC_P1 = 14, EFDC_HP = 15, EFDC_H1P = 16, EFDC_H2P = 17, &
      EFDC_HU = 18, EFDC_HV = 19, EFDC_HUI = 20, EFDC_HVI = 21, EFDC_HPI = 22, &
      EFDC_H1U = 23, EFDC_H1V = 24, EFDC_H1UI = 25, EFDC_H1VI = 26, &
      EFDC_UHDYE = 27, EFDC_VHDXE = 28, EFDC_UHDY1E = 29, EFDC_VHDX1E = 30, &
      EFDC_UHE = 31, EFDC_VHE = 32, EFDC_TSX = 33, EFDC_TSY = 34, EFDC_TBX = 35, &
      EFDC_TBY = 36, EFDC_FCAXE = 37, EFDC_FCAYE = 38, EFDC_FPGXE = 39, &
      EFDC_FPGYE = 40, EFDC_FXE = 41, EFDC_FYE = 42, EFDC_QSUME = 43, &
      EFDC_RCX = 44, EFDC_RCY = 45, EFDC_FUHDYE = 46, EFDC_FVHDXE = 47, &
      EFDC_FP = 48, EFDC_FP1 = 49, EFDC_CC = 50, EFDC_CS = 51, EFDC_CW = 52, &
      EFDC_CE = 53, EFDC_CN = 54, EFDC_CCC = 55, EFDC_CCS = 56, EFDC_CCW = 57, &
      EFDC_CCE = 58, EFDC_CCN = 59, EFDC_CCCI = 60, EFDC_FPTMP = 61, &
      EFDC_SUBO = 62, EFDC_SVBO = 63
  ! (LCM,KCM) arrays
  integer(c_int32_t), parameter :: EFDC_HPK = 64, EFDC_H1PK = 65, EFDC_HPKI = 66, &
      EFDC_UHDY2 = 67, EFDC_VHDX2 = 68, EFDC_CDZKK = 69, EFDC_CDZKMK = 70, &
      EFDC_AB = 71, EFDC_QSUM = 72, EFDC_U2 = 73, EFDC_V2 = 74, EFDC_SUB3D = 75, &
      EFDC_SVB3D = 76, EFDC_DZC = 77
  ! (LCM,0:KCM) arrays
  integer(c_int32_t), parameter :: EFDC_W2 = 96, EFDC_DZIG = 97
  ! internal mode (CALUVW): (LCM), (LCM,KCM) and (LCM,0:KCM) arrays
  integer(c_int32_t), parameter :: EFDC_STBX = 128, EFDC_STBY = 129, EFDC_VU = 130, &
      EFDC_UV = 131, EFDC_V1U = 132, EFDC_U1V = 133, EFDC_DYU = 134, EFDC_DXV = 135
  integer(c_int32_t), parameter :: EFDC_UHDYF = 160, EFDC_VHDXF = 161, EFDC_UHDY = 162, &
      EFDC_VHDX = 163, EFDC_UHDYF1 = 164, EFDC_VHDXF1 = 165, EFDC_UHDY1 = 166, &
      EFDC_VHDX1 = 167, EFDC_U = 168, EFDC_V = 169, EFDC_U1 = 170, EFDC_V1 = 171, &
      EFDC_UHDYF2 = 172, EFDC_VHDXF2 = 173, EFDC_UHDYEK = 174, EFDC_VHDXEK = 175, &
      EFDC_UHDY1EK = 176, EFDC_VHDX1EK = 177, EFDC_DU = 178, EFDC_DV = 179, &
      EFDC_AVUI = 180, EFDC_AVVI = 181
  integer(c_int32_t), parameter :: EFDC_W = 192, EFDC_W1 = 193

  type, bind(C) :: efdc_grid_t
    integer(c_int32_t) :: LA, LCM, KC, KCM
    type(c_ptr) :: LWC, LEC, LSC, LNC, LNWC, LSEC
  end type

  type, bind(C) :: efdc_params_t
    real(c_double) :: G, RITB1, SNLT, HDRY, TCON_TBEGIN
    integer(c_int32_t) :: ITERM
    real(c_double) :: RSQM, cg_rel_tol
    integer(c_int32_t) :: cg_mode, store
    real(c_double) :: mixed_inner_rtol
    integer(c_int32_t) :: mixed_max_outer, cg_kernel, mixed_precond_degree, tran_shape, ISDRY, &
                          NDRYSTP, implicit_drag
  end type

  type, bind(C) :: efdc_pbc_side_t
    integer(c_int32_t) :: n
    type(c_ptr) :: L, ispb, pser, pcb, psb
  end type

  type, bind(C) :: efdc_open_bc_t
    type(efdc_pbc_side_t) :: side(4)     ! W, E, S, N
    integer(c_int32_t) :: mtide
    type(c_ptr) :: tcp
  end type

  type, bind(C) :: efdc_cbc_side_t
    integer(c_int32_t) :: n
    type(c_ptr) :: L, cb, csert, ntscr
  end type

  type, bind(C) :: efdc_conc_bc_t
    type(efdc_cbc_side_t) :: side(4)     ! S, W, E, N
    type(c_ptr) :: wtci
  end type

  type, bind(C) :: efdc_sources_t
    integer(c_int32_t) :: n
    type(c_ptr) :: LQS, qss, qsercell, cqs, csert
  end type

  type, bind(C) :: efdc_constituent_t
    integer(c_int32_t) :: mo, isadac, isfct
    real(c_double) :: rkdye
    integer(c_int32_t) :: is_dye
    real(c_double) :: wset, probdep, wese
  end type

  type, bind(C) :: efdc_step_t
    real(c_double) :: dt
    integer(c_int64_t) :: n, niter
  end type

  type, bind(C) :: efdc_uvw_opts_t
    real(c_double) :: AVCON1
    integer(c_int32_t) :: shear, wind, uv_after
  end type

  type, bind(C) :: efdc_cg_stats_t
    integer(c_int32_t) :: iters, outer
    real(c_double) :: rsq, rsqm_used, ccmnm
  end type

  ! efdc_adj_field
  integer(c_int32_t), parameter :: EFDC_ADJ_MANNING = 0, EFDC_ADJ_HP = 1, &
      EFDC_ADJ_UHDYE = 2, EFDC_ADJ_VHDXE = 3, EFDC_ADJ_P = 4, EFDC_ADJ_HU = 5, &
      EFDC_ADJ_HV = 6, EFDC_ADJ_CON = 16

  type, bind(C) :: efdc_adj_opts_t
    integer(c_int32_t) :: nadj
    integer(c_int32_t) :: slot(16)
    real(c_double) :: cg_rel_tol
    integer(c_int32_t) :: cg_itmax
  end type

  type, bind(C) :: efdc_adj_stats_t
    integer(c_int32_t) :: cg_iters
    real(c_double) :: cg_rsq
  end type

  ! RESTOUT's records (EFDC/restartmod.f90:17-297); host arrays by c_loc
  type, bind(C) :: efdc_restart_t
    integer(c_int64_t) :: n
    real(c_double) :: timeday
    integer(c_int32_t) :: isco(0:8), istran(0:8), isice, slot(3), nsed
    type(c_ptr) :: sed_slot
    integer(c_int32_t) :: kb, kbm, keep_nlo
    type(c_ptr) :: hwq, h2wq, qq, qq1, qql, qql1, dml, ice_thick, ice_temp, temb
    type(c_ptr) :: sedb, sedb1, hbed, hbed1, vdrbed, vdrbed1
  end type

  ! HEADEROUT's grid (EFDC/efdcout.f90:73-113)
  type, bind(C) :: efdc_ee_grid_t
    integer(c_int32_t) :: ic, jc, igridv
    type(c_ptr) :: ksz, rssbce, rssbcw, rssbcs, rssbcn
  end type

  type, bind(C) :: efdc_bal2t_t
    integer(c_int32_t) :: nconst
    integer(c_int64_t) :: nsteps
    real(c_double) :: volbeg, volend, volout, volerr
    real(c_double) :: massbeg(16), massend(16), massout(16), masserr(16)
  end type

  type, bind(C) :: efdc_balance_t
    real(c_double) :: volume
    real(c_double) :: mass(16)
    integer(c_int32_t) :: nconst
  end type

  interface
    integer(c_int) function efdc_abi_version() bind(C, name="efdc_abi_version")
      import :: c_int
    end function

    subroutine efdc_params_default(p) bind(C, name="efdc_params_default")
      import :: efdc_params_t
      type(efdc_params_t), intent(out) :: p
    end subroutine

    integer(c_int) function efdc_ctx_create(grid, params, ctx) bind(C, name="efdc_ctx_create")
      import :: c_int, c_ptr, efdc_grid_t, efdc_params_t
      type(efdc_grid_t), intent(in) :: grid
      type(efdc_params_t), intent(in) :: params
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function efdc_ctx_destroy(ctx) bind(C, name="efdc_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function efdc_last_error(ctx) bind(C, name="efdc_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function efdc_set_params(ctx, params) bind(C, name="efdc_set_params")
      import :: c_int, c_ptr, efdc_params_t
      type(c_ptr), value :: ctx
      type(efdc_params_t), intent(in) :: params
    end function

    integer(c_int) function efdc_set_array(ctx, id, host) bind(C, name="efdc_set_array")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: id
      real(c_double), intent(in) :: host(*)
    end function

    integer(c_int) function efdc_get_array(ctx, id, host) bind(C, name="efdc_get_array")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: id
      real(c_double), intent(inout) :: host(*)
    end function

    integer(c_int) function efdc_set_open_bc(ctx, bc) bind(C, name="efdc_set_open_bc")
      import :: c_int, c_ptr, efdc_open_bc_t
      type(c_ptr), value :: ctx
      type(efdc_open_bc_t), intent(in) :: bc
    end function

    integer(c_int) function efdc_set_conc_bc(ctx, bc) bind(C, name="efdc_set_conc_bc")
      import :: c_int, c_ptr, efdc_conc_bc_t
      type(c_ptr), value :: ctx
      type(efdc_conc_bc_t), intent(in) :: bc
    end function

    integer(c_int) function efdc_set_sources(ctx, src) bind(C, name="efdc_set_sources")
      import :: c_int, c_ptr, efdc_sources_t
      type(c_ptr), value :: ctx
      type(efdc_sources_t), intent(in) :: src
    end function

    integer(c_int) function efdc_add_constituent(ctx, desc, slot) bind(C, name="efdc_add_constituent")
      import :: c_int, c_int32_t, c_ptr, efdc_constituent_t
      type(c_ptr), value :: ctx
      type(efdc_constituent_t), intent(in) :: desc
      integer(c_int32_t), intent(out) :: slot
    end function

    integer(c_int) function efdc_set_constituent(ctx, slot, con, con1) bind(C, name="efdc_set_constituent")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: slot
      real(c_double), intent(in) :: con(*), con1(*)
    end function

    integer(c_int) function efdc_get_constituent(ctx, slot, con, con1) bind(C, name="efdc_get_constituent")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: slot
      real(c_double), intent(inout) :: con(*), con1(*)
    end function

    integer(c_int) function efdc_get_constituent_at(ctx, slot, n, L, K, out) &
        bind(C, name="efdc_get_constituent_at")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: slot, n
      integer(c_int32_t), intent(in) :: L(*), K(*)
      real(c_double), intent(out) :: out(*)
    end function

    integer(c_int) function efdc_get_dry_state(ctx, lmaskdry, iscdry, ncordry) &
        bind(C, name="efdc_get_dry_state")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: lmaskdry(*), iscdry(*)
      integer(c_int32_t), intent(out) :: ncordry
    end function

    integer(c_int) function efdc_set_dry_counters(ctx, vdwaste, natdry) &
        bind(C, name="efdc_set_dry_counters")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: vdwaste(*)
      integer(c_int32_t), intent(in) :: natdry(*)
    end function

    integer(c_int) function efdc_get_dry_waste(ctx, qdwaste, vdwaste, natdry) &
        bind(C, name="efdc_get_dry_waste")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: qdwaste(*), vdwaste(*)
      integer(c_int32_t), intent(out) :: natdry(*)
    end function

    integer(c_int) function efdc_congrad(ctx, stats) bind(C, name="efdc_congrad")
      import :: c_int, c_ptr, efdc_cg_stats_t
      type(c_ptr), value :: ctx
      type(efdc_cg_stats_t), intent(out) :: stats
    end function

    ! CONGRAD on device arrays the caller holds (C_DEVLOC of a device
    ! allocation), with no device allocation of its own
    integer(c_int) function efdc_congrad_noalloc(la, lwc, lec, lsc, lnc, ccc, ccs, ccw, &
        cce, ccn, ccci, fptmp, p, rsqm, iterm, stream, stats) &
        bind(C, name="efdc_congrad_noalloc")
      import :: c_int, c_int32_t, c_ptr, c_double, efdc_cg_stats_t
      integer(c_int32_t), value :: la, iterm
      type(c_ptr), value :: lwc, lec, lsc, lnc, ccc, ccs, ccw, cce, ccn, ccci, fptmp, p
      real(c_double), value :: rsqm
      type(c_ptr), value :: stream
      type(efdc_cg_stats_t), intent(out) :: stats
    end function

    integer(c_int) function efdc_calpuv2c(ctx, step, stats) bind(C, name="efdc_calpuv2c")
      import :: c_int, c_ptr, efdc_step_t, efdc_cg_stats_t
      type(c_ptr), value :: ctx
      type(efdc_step_t), intent(in) :: step
      type(efdc_cg_stats_t), intent(out) :: stats
    end function

    integer(c_int) function efdc_calconc(ctx, step) bind(C, name="efdc_calconc")
      import :: c_int, c_ptr, efdc_step_t
      type(c_ptr), value :: ctx
      type(efdc_step_t), intent(in) :: step
    end function

    integer(c_int) function efdc_caluvw(ctx, step, opts) bind(C, name="efdc_caluvw")
      import :: c_int, c_ptr, efdc_step_t, efdc_uvw_opts_t
      type(c_ptr), value :: ctx
      type(efdc_step_t), intent(in) :: step
      type(efdc_uvw_opts_t), intent(in) :: opts
    end function

    integer(c_int) function efdc_bal2t_begin(ctx) bind(C, name="efdc_bal2t_begin")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function efdc_bal2t_end(ctx, out) bind(C, name="efdc_bal2t_end")
      import :: c_int, c_ptr, efdc_bal2t_t
      type(c_ptr), value :: ctx
      type(efdc_bal2t_t), intent(out) :: out
    end function

    integer(c_int) function efdc_set_cg_replay(ctx, mode) bind(C, name="efdc_set_cg_replay")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode
    end function

    integer(c_int) function efdc_get_cg_replay(ctx, mode) bind(C, name="efdc_get_cg_replay")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: mode
    end function

    integer(c_int) function efdc_set_cg_pipeline(ctx, mode) bind(C, name="efdc_set_cg_pipeline")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode
    end function

    integer(c_int) function efdc_get_cg_pipeline(ctx, mode) bind(C, name="efdc_get_cg_pipeline")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: mode
    end function

    integer(c_int) function efdc_set_cg64_pipeline(ctx, on) bind(C, name="efdc_set_cg64_pipeline")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function

    integer(c_int) function efdc_get_cg64_pipeline(ctx, on) bind(C, name="efdc_get_cg64_pipeline")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: on
    end function

    integer(c_int) function efdc_set_cg_launch(ctx, mode, spin_limit) bind(C, name="efdc_set_cg_launch")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode, spin_limit
    end function

    integer(c_int) function efdc_get_cg_launch(ctx, mode, retries) bind(C, name="efdc_get_cg_launch")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: mode, retries
    end function

    integer(c_int) function efdc_set_segment_tile(ctx, cells) bind(C, name="efdc_set_segment_tile")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: cells
    end function

    integer(c_int) function efdc_get_segment_tile(ctx, cells) bind(C, name="efdc_get_segment_tile")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: cells
    end function

    integer(c_int) function efdc_restart_text(ctx, r, buf, cap, bytes) &
        bind(C, name="efdc_restart_text")
      import :: c_int, c_ptr, c_int64_t, efdc_restart_t
      type(c_ptr), value :: ctx
      type(efdc_restart_t), intent(in) :: r
      type(c_ptr), value :: buf
      integer(c_int64_t), value :: cap
      integer(c_int64_t), intent(out) :: bytes
    end function

    integer(c_int) function efdc_restart_write(ctx, r, path) bind(C, name="efdc_restart_write")
      import :: c_int, c_ptr, c_char, efdc_restart_t
      type(c_ptr), value :: ctx
      type(efdc_restart_t), intent(in) :: r
      character(kind=c_char), dimension(*), intent(in) :: path
    end function

    integer(c_int) function efdc_ee_open(ctx, dir, grid) bind(C, name="efdc_ee_open")
      import :: c_int, c_ptr, c_char, efdc_ee_grid_t
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: dir
      type(efdc_ee_grid_t), intent(in) :: grid
    end function

    integer(c_int) function efdc_ee_append(ctx, dir, n, eetime, delt, lminstep) &
        bind(C, name="efdc_ee_append")
      import :: c_int, c_ptr, c_char, c_int32_t, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: dir
      integer(c_int32_t), value :: n, lminstep
      real(c_double), value :: eetime, delt
    end function

    integer(c_int) function efdc_balance(ctx, out) bind(C, name="efdc_balance")
      import :: c_int, c_ptr, efdc_balance_t
      type(c_ptr), value :: ctx
      type(efdc_balance_t), intent(out) :: out
    end function

    integer(c_int) function efdc_synchronize(ctx) bind(C, name="efdc_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function efdc_set_source_cqs(ctx, cqs) bind(C, name="efdc_set_source_cqs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: cqs(*)
    end function

    ! independent basins in one context
    integer(c_int) function efdc_set_segments(ctx, nseg, first_L) bind(C, name="efdc_set_segments")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nseg
      integer(c_int32_t), intent(in) :: first_L(*)
    end function

    integer(c_int) function efdc_segment_stats(ctx, cap, iters, status) &
        bind(C, name="efdc_segment_stats")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: cap
      integer(c_int32_t), intent(out) :: iters(*), status(*)
    end function

    integer(c_int) function efdc_balance_segments(ctx, cap, volume, mass) &
        bind(C, name="efdc_balance_segments")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: cap
      real(c_double), intent(out) :: volume(*), mass(*)
    end function

    ! snapshots
    integer(c_int) function efdc_state_bytes(ctx, bytes) bind(C, name="efdc_state_bytes")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), intent(out) :: bytes
    end function

    integer(c_int) function efdc_state_alloc(ctx, buf) bind(C, name="efdc_state_alloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: buf
    end function

    integer(c_int) function efdc_state_free(ctx, buf) bind(C, name="efdc_state_free")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, buf
    end function

    integer(c_int) function efdc_state_save(ctx, buf) bind(C, name="efdc_state_save")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, buf
    end function

    integer(c_int) function efdc_state_load(ctx, buf) bind(C, name="efdc_state_load")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, buf
    end function

    ! discrete adjoint
    integer(c_int) function efdc_adj_enable(ctx, opts) bind(C, name="efdc_adj_enable")
      import :: c_int, c_ptr, efdc_adj_opts_t
      type(c_ptr), value :: ctx
      type(efdc_adj_opts_t), intent(in) :: opts
    end function

    integer(c_int) function efdc_adj_disable(ctx) bind(C, name="efdc_adj_disable")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function efdc_adj_reset(ctx) bind(C, name="efdc_adj_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function efdc_adj_seed_con(ctx, slot, n, L, K, val) &
        bind(C, name="efdc_adj_seed_con")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: slot, n
      integer(c_int32_t), intent(in) :: L(*), K(*)
      real(c_double), intent(in) :: val(*)
    end function

    integer(c_int) function efdc_adj_step(ctx, step, stats) bind(C, name="efdc_adj_step")
      import :: c_int, c_ptr, efdc_step_t, efdc_adj_stats_t
      type(c_ptr), value :: ctx
      type(efdc_step_t), intent(in) :: step
      type(efdc_adj_stats_t), intent(out) :: stats
    end function

    integer(c_int) function efdc_adj_source_grad(ctx, out) bind(C, name="efdc_adj_source_grad")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(*)
    end function

    integer(c_int) function efdc_adj_get(ctx, field, out) bind(C, name="efdc_adj_get")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: field
      real(c_double), intent(out) :: out(*)
    end function

    integer(c_int) function efdc_adj_record_bytes(ctx, bytes) &
        bind(C, name="efdc_adj_record_bytes")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), intent(out) :: bytes
    end function

    integer(c_int) function efdc_adj_record_save(ctx, buf) bind(C, name="efdc_adj_record_save")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, buf
    end function

    integer(c_int) function efdc_adj_record_load(ctx, buf) bind(C, name="efdc_adj_record_load")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, buf
    end function

    ! autotuner; the report is an efdc_tune_report_t the caller lays out
    integer(c_int) function efdc_autotune(ctx, on) bind(C, name="efdc_autotune")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function

    integer(c_int) function efdc_tune_report(ctx, out) bind(C, name="efdc_tune_report")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, out
    end function

    ! kernel timers; the report is an efdc_timers_t the caller lays out
    integer(c_int) function efdc_timers_enable(ctx, on) bind(C, name="efdc_timers_enable")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function

    integer(c_int) function efdc_timers_read(ctx, out) bind(C, name="efdc_timers_read")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, out
    end function

    ! device pointer of an array (for hosts that launch their own kernels)
    integer(c_int) function efdc_device_ptr(ctx, id, ptr, pitch) bind(C, name="efdc_device_ptr")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: id
      type(c_ptr), intent(out) :: ptr
      integer(c_int64_t), intent(out) :: pitch
    end function

    ! achievable HBM bandwidth of a plain copy
    integer(c_int) function efdc_stream_copy_gbs(ctx, bytes, reps, gbs) &
        bind(C, name="efdc_stream_copy_gbs")
      import :: c_int, c_ptr, c_size_t, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_size_t), value :: bytes
      integer(c_int32_t), value :: reps
      real(c_double), intent(out) :: gbs
    end function

    ! depth-uniform stand-in of the internal mode (no reference counterpart)
    integer(c_int) function efdc_closure_set_manning(ctx, n_lcm) &
        bind(C, name="efdc_closure_set_manning")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: n_lcm(*)
    end function

    integer(c_int) function efdc_closure_pre(ctx) bind(C, name="efdc_closure_pre")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function efdc_closure_post(ctx) bind(C, name="efdc_closure_post")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface

contains

  ! Text of the last error of a context as a Fortran string.
  function efdc_error_text(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: j
    msg = ' '
    p = efdc_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do j = 1, 256
      if (s(j) == c_null_char) exit
      msg(j:j) = s(j)
    end do
  end function
end module efdc_hip
