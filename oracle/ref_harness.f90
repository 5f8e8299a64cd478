! ref_harness.f90 - driver around the reference's own LEQ_BICGS path.
!
! TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) against
! the reference modules compiled from model/ of the MFIX tree; the binary lives
! in oracle/_ref/ and writes golden vectors for tests/golden/.  It sets the
! serial module state that GRIDMAP_INIT / SET_INCREMENTS would set
! (model/dmp_modules/gridmap_mod.f:259-351,511-556; model/set_increments.f:
! 72-178; model/set_max2.f:53-69) and calls:
!   op 1  LEQ_BICGS        model/leq_bicgs.f:23
!   op 2  LEQ_MATVEC       model/leqsol_mod.f:103
!   op 3  LEQ_MSOLVE       model/leqsol_mod.f:286  (LINE, CMETHOD)
!   op 4  LEQ_MSOLVE1      model/leqsol_mod.f:681  (DIAG)
!   op 5  dot_product_par  model/leqsol_mod.f:1094
!   op 6  dot_product_par2 model/leqsol_mod.f:1210
!   op 7  CALC_RESID_C     model/calc_resid.f:16
!   op 8  CALC_RESID_S     model/calc_resid.f:195
!   op 9  CALC_RESID_PP    model/calc_resid.f:391
!   op 10 INIT_AB_M + CONV_PP_G + SOURCE_PP_G (gas phase, MMAX = 0)
!                          model/init_ab_m.f:20, model/conv_pp_g.f:33,
!                          model/source_pp_g.f:18
! Input file (stream): 12 int32 (op, imax, jmax, kmax, no_k, cyc_x, cyc_y,
! cyc_z, is_serial, minimize_dotproducts, icheck_bicgs, itmax), int32 seed,
! real64 tol, 4+4 chars pc/cmethod, then A(N3*7), B(N3), X(N3) real64; ops
! 7-9 then add int32 debug_resid, int32 phase M, real64 TOL (op 8), real64
! NORM (op 9), FLAG(N3) int32 and EP_s(N3) real64 (phase M when M > 0).
! For these ops the neighbour tables come from the reference's SET_INDEX1A
! (model/set_index1a.f:34) on that FLAG, one class per cell.
! Output file (stream): int32 ier, int32 iter_tot, real64 dot results(2)
! (op 1: dres(2) = wall seconds of the LEQ_BICGS call),
! X(N3), A(N3*7), B(N3), noise(N3).  Ops 7-9: ier = IJK_RESID, iter_tot =
! 0, dres = (NUM, DEN), X(1:2) = (RESID, MAX_RESID).
! Op 10 adds, after X: int32 IJK_P_g (0 = undefined), real64 ODT, FLAG(N3)
! int32 and 16 real64 arrays of N3: ROP_gE, ROP_gN, ROP_gT, AYZ, AXZ, AXY,
! VOL, ROP_g, ROP_go, U_g, V_g, W_g, SUM_R_g, D_E(:,0), D_N(:,0), D_T(:,0);
! it returns the assembled A_M/B_M in A/B and B_MMAX in X.
! Op 11 (scalar transport, phase 0) adds, after X: int32 stage, real64 UR,
! FLAG(N3) int32, 14 real64 arrays of N3 (DIF, U, V, W, FLUX_E, FLUX_N,
! FLUX_T, AYZ, AXZ, AXY, S_P, S_C, EP, PHI), then ODX_E, OX, FX_E, FX_E_bar
! (IMAX2 each, from I = 1), ODY_N, FY_N, FY_N_bar (JMAX2), ODZ_T, FZ_T,
! FZ_T_bar (KMAX2).  Stage 1: INIT_AB_M + CONV_DIF_PHI with DISC = 0
! (model/init_ab_m.f:20, model/conv_dif_phi.f:17); stage 2: SOURCE_PHI
! (model/source_phi.f:25) on the A/B read in; stage 3: UNDER_RELAX_S
! (model/under_relax.f:20) with VAR = PHI on the A/B read in.  FLAG_E/N/T
! follow SET_FLAGS (model/set_flags.f:400-470) without internal surfaces.
! Op 12: INITEXTRACT(whole domain) + MATRIXEXTRACT(A_M, B_M, 0, 'HARNESS')
! (model/dbg_mod.f:82, :228) write Am_HARNESS.csv and Bm_HARNESS.csv in the
! working directory; op 13: WRITE_AB_M(A_M, B_M, 0) (model/write_ab_m.f:20)
! writes to UNIT_LOG, opened as harness_ab_m.log in the working directory.
program ref_harness
  use param, only: dimension_3
  use compar
  use geometry
  use indices
  use leqsol
  use parallel, only: is_serial, use_doloop
  use cutcell, only: re_indexing, cartesian_grid
  use funits, only: dmp_log
  use run, only: debug_resid
  use fldvar, only: rop_s, ro_s
  use param, only: dimension_m
  use mflux, only: rop_ge, rop_gn, rop_gt
  use fldvar, only: rop_g, rop_go, u_g, v_g, w_g
  use rxns, only: sum_r_g
  use pgcor, only: d_e, d_n, d_t
  use run, only: odt, shear
  use physprop, only: mmax, ro_g0
  use bc, only: ijk_p_g
  use param1, only: undefined_i
  use cutcell, only: cut_treatment_at, cut_cell_at
  use run, only: def_cor
  use is, only: is_defined
  use functions, only: wall_at, fluid_at, im_of, jm_of, km_of, ip_of, jp_of, kp_of
  use dbg, only: initExtract, matrixExtract
  use funits, only: unit_log
  implicit none
  integer :: shdr(1)
  double precision :: sur
  double precision, allocatable :: sf(:,:), g1i(:,:), g1j(:,:), g1k(:,:)
  integer :: phdr(1)
  double precision :: podt
  double precision, allocatable :: pf(:,:), bmm(:,:)
  integer :: rhdr(2), mph, ijkr, ijkw, ijke, ijks, ijkn, ijkb, ijkt, imjk, ipjk, ijmk, ijpk, ijkm, ijkp
  double precision :: rtol, rnorm, rnum, rden, rres, rmax
  integer, allocatable :: flagin(:)
  double precision, allocatable :: epsin(:), a3(:,:,:), b2(:,:)
  integer :: hdr(12), seed, op, itmax
  double precision :: tol
  character(len=4) :: pc, cmethod
  character(len=512) :: fin, fout
  double precision, allocatable :: amat(:,:), bvec(:), xvec(:), yvec(:), noise(:)
  integer, allocatable :: seedarr(:), saved(:)
  integer :: nseed, ier, ni, nj, nk, i, j, k, ijk, n3, ncls
  integer :: im, ip, jm, jp, km, kp, key, cls
  integer, allocatable :: classof(:)
  double precision :: dres(2), elapsed
  integer(8) :: t0, t1, rate

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  read(10) seed
  read(10) tol
  read(10) pc, cmethod
  op = hdr(1); itmax = hdr(12)

  ! ---- serial grid state (set_max2.f:53-69, gridmap_mod.f) ----
  imax = hdr(2); jmax = hdr(3); kmax = hdr(4)
  no_k = hdr(5) /= 0
  do_k = .not. no_k
  no_i = .false.; no_j = .false.; do_i = .true.; do_j = .true.
  cyclic_x = hdr(6) /= 0; cyclic_y = hdr(7) /= 0; cyclic_z = hdr(8) /= 0
  ni = imax + 2; nj = jmax + 2
  if (no_k) then
     nk = 1
  else
     nk = kmax + 2
  end if
  myPE = 0; numPEs = 1; nodesi = 1; nodesj = 1; nodesk = 1; root = 0
  dmp_log = .false.
  re_indexing = .false.; cartesian_grid = .false.
  is_serial = hdr(9) /= 0
  use_doloop = .false.
  minimize_dotproducts = hdr(10) /= 0
  icheck_bicgs = max(1, hdr(11))
  nlayers_bicgs = 1
  imin1 = 2; imax1 = imax + 1; imin2 = 1; imax2 = imax + 2; imin3 = 1; imax3 = imax + 2
  jmin1 = 2; jmax1 = jmax + 1; jmin2 = 1; jmax2 = jmax + 2; jmin3 = 1; jmax3 = jmax + 2
  if (no_k) then
     kmin1 = 1; kmax1 = 1; kmin2 = 1; kmax2 = 1; kmin3 = 1; kmax3 = 1
  else
     kmin1 = 2; kmax1 = kmax + 1; kmin2 = 1; kmax2 = kmax + 2; kmin3 = 1; kmax3 = kmax + 2
  end if
  istart1 = imin1; iend1 = imax1; istart2 = imin2; iend2 = imax2; istart3 = imin3; iend3 = imax3
  jstart1 = jmin1; jend1 = jmax1; jstart2 = jmin2; jend2 = jmax2; jstart3 = jmin3; jend3 = jmax3
  kstart1 = kmin1; kend1 = kmax1; kstart2 = kmin2; kend2 = kmax2; kstart3 = kmin3; kend3 = kmax3
  istart = istart2; iend = iend2; jstart = jstart2; jend = jend2; kstart = kstart2; kend = kend2
  c1 = nj
  c2 = nj * ni
  c0 = 1 - jstart3 - c1 * istart3 - c2 * kstart3
  n3 = ni * nj * nk
  dimension_3 = n3
  ijkstart3 = 1; ijkend3 = n3; ijksize3 = n3
  allocate(ijk_array_of(istart3:iend3, jstart3:jend3, kstart3:kend3))
  allocate(funijk_map_c(istart3:iend3, jstart3:jend3, kstart3:kend3))
  allocate(imap_c(imin3:imax3), jmap_c(jmin3:jmax3), kmap_c(kmin3:kmax3))
  ! no cut cells: every DEAD_CELL_AT flag is false (gridmap_mod.f:572-579)
  allocate(dead_cell_at(imin3-1:imax3+1, jmin3-1:jmax3+1, kmin3-1:kmax3+1))
  dead_cell_at = .false.
  do i = imin3, imax3; imap_c(i) = i; end do
  do j = jmin3, jmax3; jmap_c(j) = j; end do
  do k = kmin3, kmax3; kmap_c(k) = k; end do
  if (cyclic_x) then; imap_c(imax2) = imin1; imap_c(imin2) = imax1; end if
  if (cyclic_y) then; jmap_c(jmax2) = jmin1; jmap_c(jmin2) = jmax1; end if
  if (cyclic_z .and. do_k) then; kmap_c(kmax2) = kmin1; kmap_c(kmin2) = kmax1; end if
  do k = kstart3, kend3
     do i = istart3, iend3
        do j = jstart3, jend3
           ijk_array_of(i, j, k) = j + c0 + i * c1 + k * c2
        end do
     end do
  end do
  do k = kstart3, kend3
     do i = istart3, iend3
        do j = jstart3, jend3
           funijk_map_c(i, j, k) = ijk_array_of(imap_c(i), jmap_c(j), kmap_c(k))
        end do
     end do
  end do

  ! ---- neighbour classes (set_increments.f:72-112,173-178) ----
  if (allocated(cell_class)) deallocate(cell_class)
  allocate(cell_class(n3))
  allocate(classof(0:124))
  classof = 0
  ncls = 0
  do k = kstart3, kend3
     do i = istart3, iend3
        do j = jstart3, jend3
           ijk = ijk_array_of(i, j, k)
           im = nbm(i, imin3, imax3, cyclic_x, imap_c)
           ip = nbp(i, imin3, imax3, cyclic_x, imap_c)
           jm = nbm(j, jmin3, jmax3, cyclic_y, jmap_c)
           jp = nbp(j, jmin3, jmax3, cyclic_y, jmap_c)
           if (do_k) then
              km = nbm(k, kmin3, kmax3, cyclic_z, kmap_c)
              kp = nbp(k, kmin3, kmax3, cyclic_z, kmap_c)
           else
              km = k; kp = k
           end if
           key = edge(i, imin3, imax3) + 5 * edge(j, jmin3, jmax3) + 25 * edge(k, kmin3, kmax3)
           if (classof(key) == 0) then
              ncls = ncls + 1
              classof(key) = ncls
              cls = ncls
              increment_for_mp(1, cls) = ijk_array_of(im, j, k) - ijk
              increment_for_mp(2, cls) = ijk_array_of(ip, j, k) - ijk
              increment_for_mp(3, cls) = ijk_array_of(i, jm, k) - ijk
              increment_for_mp(4, cls) = ijk_array_of(i, jp, k) - ijk
              increment_for_mp(5, cls) = ijk_array_of(i, j, km) - ijk
              increment_for_mp(6, cls) = ijk_array_of(i, j, kp) - ijk
           end if
           cell_class(ijk) = classof(key)
        end do
     end do
  end do
  use_corecell_loop = (ni >= 5 .and. nj >= 5 .and. (nk >= 5 .or. no_k))

  allocate(amat(n3, -3:3), bvec(n3), xvec(n3), yvec(n3), noise(n3))
  read(10) amat
  read(10) bvec
  read(10) xvec
  if (op == 10) then
     read(10) phdr
     read(10) podt
     allocate(flagin(n3), pf(n3, 16))
     read(10) flagin
     read(10) pf
  end if
  if (op == 11) then
     read(10) shdr
     read(10) sur
     allocate(flagin(n3), sf(n3, 14), g1i(ni, 4), g1j(nj, 3), g1k(nk, 3))
     read(10) flagin
     read(10) sf
     read(10) g1i
     read(10) g1j
     read(10) g1k
  end if
  if (op >= 7 .and. op <= 9) then
     read(10) rhdr
     read(10) rtol, rnorm
     allocate(flagin(n3), epsin(n3))
     read(10) flagin
     read(10) epsin
  end if
  close(10)

  if (op >= 7 .and. op <= 11) then
     if (op == 10 .or. op == 11) then
        rhdr = 0
        rtol = 0d0; rnorm = 1d0
        allocate(epsin(n3))
        epsin = 0d0
     end if
     ! ---- cell flags, index maps and SET_INCREMENTS' tables (set_increments.f:
     ! 52-112 IM1..KP1, :113-126 I_OF.., :128-178 classes via SET_INDEX1A)
     debug_resid = rhdr(1) /= 0
     mph = rhdr(2)
     if (.not. allocated(ijksize3_all)) allocate(ijksize3_all(0:0))
     ijksize3_all(0) = n3
     ijkmax2 = n3
     dimension_m = max(1, mph)
     if (allocated(flag)) deallocate(flag)
     allocate(flag(n3))
     flag = flagin
     allocate(i_of(n3), j_of(n3), k_of(n3))
     allocate(im1(istart3:iend3), ip1(istart3:iend3), jm1(jstart3:jend3), jp1(jstart3:jend3))
     allocate(km1(kstart3:kend3), kp1(kstart3:kend3))
     do i = istart3, iend3
        im1(i) = nbm(i, imin3, imax3, cyclic_x, imap_c); ip1(i) = nbp(i, imin3, imax3, cyclic_x, imap_c)
     end do
     do j = jstart3, jend3
        jm1(j) = nbm(j, jmin3, jmax3, cyclic_y, jmap_c); jp1(j) = nbp(j, jmin3, jmax3, cyclic_y, jmap_c)
     end do
     do k = kstart3, kend3
        if (do_k) then
           km1(k) = nbm(k, kmin3, kmax3, cyclic_z, kmap_c); kp1(k) = nbp(k, kmin3, kmax3, cyclic_z, kmap_c)
        else
           km1(k) = k; kp1(k) = k
        end if
     end do
     do k = kstart3, kend3
        do i = istart3, iend3
           do j = jstart3, jend3
              ijk = ijk_array_of(i, j, k)
              i_of(ijk) = i; j_of(ijk) = j; k_of(ijk) = k
           end do
        end do
     end do
     do k = kstart3, kend3
        do i = istart3, iend3
           do j = jstart3, jend3
              ijk = ijk_array_of(i, j, k)
              call set_index1a(i, j, k, ijk, imjk, ipjk, ijmk, ijpk, ijkm, ijkp, ijkw, ijke, ijks, ijkn, &
                               ijkb, ijkt)
              cell_class(ijk) = ijk
              increment_for_mp(1, ijk) = imjk - ijk; increment_for_mp(2, ijk) = ipjk - ijk
              increment_for_mp(3, ijk) = ijmk - ijk; increment_for_mp(4, ijk) = ijpk - ijk
              increment_for_mp(5, ijk) = ijkm - ijk; increment_for_mp(6, ijk) = ijkp - ijk
              increment_for_nb(1, ijk) = ijke - ijk; increment_for_nb(2, ijk) = ijkw - ijk
              increment_for_nb(3, ijk) = ijks - ijk; increment_for_nb(4, ijk) = ijkn - ijk
              increment_for_nb(5, ijk) = ijkb - ijk; increment_for_nb(6, ijk) = ijkt - ijk
           end do
        end do
     end do
     allocate(rop_s(n3, dimension_m), ro_s(n3, dimension_m))
     ro_s = 1d0
     rop_s = 0d0
     if (mph > 0) rop_s(:, mph) = epsin
     allocate(a3(n3, -3:3, 0:dimension_m), b2(n3, 0:dimension_m))
     a3 = 0d0; b2 = 0d0
     a3(:, :, mph) = amat
     b2(:, mph) = bvec
  end if

  iter_tot = 0
  ier = 0
  dres = 0d0
  noise = 0d0
  select case (op)
  case (1)
     call random_seed(size=nseed)
     allocate(seedarr(nseed), saved(nseed))
     do i = 1, nseed
        seedarr(i) = seed + 7919 * i
     end do
     call random_seed(put=seedarr)
     call random_seed(get=saved)
     call system_clock(t0, rate)
     call leq_bicgs('HARNESS', 1, xvec, amat, bvec, cmethod, tol, pc, itmax, ier)
     call system_clock(t1)
     dres(2) = dble(t1 - t0) / dble(rate)
     call random_seed(put=saved)
     call random_number(noise)
     yvec = xvec
  case (2)
     call leq_matvec('HARNESS', xvec, amat, yvec)
  case (3)
     call leq_msolve('HARNESS', xvec, amat, yvec, cmethod)
  case (4)
     call leq_msolve1('HARNESS', xvec, amat, yvec, cmethod)
  case (5)
     dres(1) = dot_product_par(xvec, bvec)
     yvec = xvec
  case (6)
     dres = dot_product_par2(xvec, bvec, xvec, xvec)
     yvec = xvec
  case (10)
     allocate(rop_ge(n3), rop_gn(n3), rop_gt(n3), ayz(n3), axz(n3), axy(n3), vol(n3))
     allocate(rop_g(n3), rop_go(n3), u_g(n3), v_g(n3), w_g(n3), sum_r_g(n3))
     allocate(d_e(n3, 0:dimension_m), d_n(n3, 0:dimension_m), d_t(n3, 0:dimension_m))
     allocate(bmm(n3, 0:dimension_m))
     rop_ge = pf(:, 1); rop_gn = pf(:, 2); rop_gt = pf(:, 3)
     ayz = pf(:, 4); axz = pf(:, 5); axy = pf(:, 6); vol = pf(:, 7)
     rop_g = pf(:, 8); rop_go = pf(:, 9); u_g = pf(:, 10); v_g = pf(:, 11); w_g = pf(:, 12)
     sum_r_g = pf(:, 13)
     d_e = 0d0; d_n = 0d0; d_t = 0d0
     d_e(:, 0) = pf(:, 14); d_n(:, 0) = pf(:, 15); d_t(:, 0) = pf(:, 16)
     odt = podt
     shear = .false.
     mmax = 0
     ro_g0 = 1d0
     if (phdr(1) > 0) then
        ijk_p_g = phdr(1)
     else
        ijk_p_g = undefined_i
     end if
     bmm = 0d0
     call init_ab_m(a3, b2, n3, 0)
     call conv_pp_g(a3, b2)
     call source_pp_g(a3, b2, bmm)
     amat = a3(:, :, 0)
     bvec = b2(:, 0)
     yvec = bmm(:, 0)
  case (11)
     allocate(ayz(n3), axz(n3), axy(n3), flag_e(n3), flag_n(n3), flag_t(n3))
     allocate(cut_treatment_at(n3), cut_cell_at(n3))
     allocate(odx_e(0:ni), ox(0:ni), fx_e(0:ni), fx_e_bar(0:ni))
     allocate(ody_n(0:nj), fy_n(0:nj), fy_n_bar(0:nj), odz_t(0:nk), fz_t(0:nk), fz_t_bar(0:nk))
     ayz = sf(:, 8); axz = sf(:, 9); axy = sf(:, 10)
     odx_e = 0d0; ox = 0d0; fx_e = 0d0; fx_e_bar = 0d0
     ody_n = 0d0; fy_n = 0d0; fy_n_bar = 0d0; odz_t = 0d0; fz_t = 0d0; fz_t_bar = 0d0
     odx_e(1:ni) = g1i(:, 1); ox(1:ni) = g1i(:, 2); fx_e(1:ni) = g1i(:, 3); fx_e_bar(1:ni) = g1i(:, 4)
     ody_n(1:nj) = g1j(:, 1); fy_n(1:nj) = g1j(:, 2); fy_n_bar(1:nj) = g1j(:, 3)
     odz_t(1:nk) = g1k(:, 1); fz_t(1:nk) = g1k(:, 2); fz_t_bar(1:nk) = g1k(:, 3)
     cut_treatment_at = .false.; cut_cell_at = .false.
     def_cor = .false.
     is_defined = .false.
     ! FLAG_E/N/T as SET_FLAGS leaves them with no internal surface
     flag_e = undefined_i; flag_n = undefined_i; flag_t = undefined_i
     do ijk = 1, n3
        i = i_of(ijk); j = j_of(ijk); k = k_of(ijk)
        if (wall_at(ijk)) then
           flag_e(ijk) = 0; flag_n(ijk) = 0; flag_t(ijk) = 0
           flag_e(im_of(ijk)) = 0; flag_n(jm_of(ijk)) = 0; flag_t(km_of(ijk)) = 0
           if (flag(ijk) == 106 .or. flag(ijk) == 107) then
              if (i == imax2 .and. j /= 1 .and. j /= jmax2) then
                 if (no_k) then
                    if (.not. wall_at(im_of(ijk))) flag_e(im_of(ijk)) = 2000
                 else if (k /= 1 .and. k /= kmax2) then
                    if (.not. wall_at(im_of(ijk))) flag_e(im_of(ijk)) = 2000
                 end if
              end if
              if (j == jmax2 .and. i /= 1 .and. i /= imax2) then
                 if (no_k) then
                    if (.not. wall_at(jm_of(ijk))) flag_n(jm_of(ijk)) = 2000
                 else if (k /= 1 .and. k /= kmax2) then
                    if (.not. wall_at(jm_of(ijk))) flag_n(jm_of(ijk)) = 2000
                 end if
              end if
              if (k == kmax2 .and. j /= 1 .and. j /= jmax2 .and. i /= 1 .and. i /= imax2) then
                 if (.not. wall_at(km_of(ijk))) flag_t(km_of(ijk)) = 2000
              end if
           end if
        else if (fluid_at(ijk)) then
           if (.not. wall_at(im_of(ijk)) .and. flag_e(im_of(ijk)) == undefined_i) &
              flag_e(im_of(ijk)) = 2000 + flag(im_of(ijk))
           if (.not. wall_at(jm_of(ijk)) .and. flag_n(jm_of(ijk)) == undefined_i) &
              flag_n(jm_of(ijk)) = 2000 + flag(jm_of(ijk))
           if (.not. wall_at(km_of(ijk)) .and. flag_t(km_of(ijk)) == undefined_i) &
              flag_t(km_of(ijk)) = 2000 + flag(km_of(ijk))
           if (.not. wall_at(ip_of(ijk)) .and. flag_e(ijk) == undefined_i) &
              flag_e(ijk) = 2000 + flag(ip_of(ijk))
           if (.not. wall_at(jp_of(ijk)) .and. flag_n(ijk) == undefined_i) &
              flag_n(ijk) = 2000 + flag(jp_of(ijk))
           if (.not. wall_at(kp_of(ijk)) .and. flag_t(ijk) == undefined_i) &
              flag_t(ijk) = 2000 + flag(kp_of(ijk))
        end if
     end do
     if (shdr(1) == 1) then
        call init_ab_m(a3, b2, n3, 0)
        call conv_dif_phi(sf(:, 14), sf(:, 1), 0, sf(:, 2), sf(:, 3), sf(:, 4), sf(:, 5), sf(:, 6), &
                          sf(:, 7), 0, a3, b2)
     else if (shdr(1) == 2) then
        call source_phi(sf(:, 11), sf(:, 12), sf(:, 13), sf(:, 14), 0, a3, b2)
     else
        call under_relax_s(sf(:, 14), a3, b2, 0, sur)
     end if
     amat = a3(:, :, 0)
     bvec = b2(:, 0)
     yvec = 0d0
  case (12, 13)
     dimension_m = 1
     ijkmax3 = n3
     allocate(a3(n3, -3:3, 0:1), b2(n3, 0:1))
     a3 = 0d0; b2 = 0d0
     a3(:, :, 0) = amat
     b2(:, 0) = bvec
     if (op == 12) then
        ! the default region, given explicitly: initExtract's MERGE of an
        ! absent optional argument (dbg_mod.f:108-113) dereferences it
        call initExtract(imin3, imax3, jmin3, jmax3, kmin3, kmax3)
        call matrixExtract(a3, b2, 0, 'HARNESS')
     else
        dmp_log = .true.
        open(unit_log, file='harness_ab_m.log', status='replace')
        call write_ab_m(a3, b2, 0)
        close(unit_log)
        dmp_log = .false.
     end if
     yvec = 0d0
  case (7, 8, 9)
     rres = 0d0; rmax = 0d0; ijkr = 0
     if (op == 7) then
        call calc_resid_c(xvec, a3, b2, mph, rnum, rden, rres, rmax, ijkr)
     else if (op == 8) then
        call calc_resid_s(xvec, a3, b2, mph, rnum, rden, rres, rmax, ijkr, rtol)
     else
        call calc_resid_pp(b2, rnorm, rnum, rden, rres, rmax, ijkr)
     end if
     ier = ijkr
     dres(1) = rnum; dres(2) = rden
     yvec = 0d0
     yvec(1) = rres; yvec(2) = rmax
  end select

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) ier, iter_tot(1)
  write(11) dres
  write(11) yvec
  write(11) amat
  write(11) bvec
  write(11) noise
  close(11)

contains

  integer function edge(i, lo, hi)
    integer, intent(in) :: i, lo, hi
    if (i == lo) then
       edge = 1
    else if (i == hi) then
       edge = 2
    else if (i == lo + 1) then
       edge = 3
    else if (i == hi - 1) then
       edge = 4
    else
       edge = 0
    end if
  end function edge

  ! IM1/IP1 of set_increments.f:72-112 for one axis (NODES=1)
  integer function nbm(i, lo, hi, cyc, map)
    integer, intent(in) :: i, lo, hi
    logical, intent(in) :: cyc
    integer, intent(in) :: map(lo:hi)
    logical :: shift
    shift = .not. (i == lo .or. i == hi)
    if (cyc .and. shift) then
       nbm = map(map(i) - 1)
    else
       nbm = max(lo, i - 1)
    end if
  end function nbm

  integer function nbp(i, lo, hi, cyc, map)
    integer, intent(in) :: i, lo, hi
    logical, intent(in) :: cyc
    integer, intent(in) :: map(lo:hi)
    logical :: shift
    shift = .not. (i == lo .or. i == hi)
    if (cyc .and. shift) then
       nbp = map(map(i) + 1)
    else
       nbp = min(hi, i + 1)
    end if
  end function nbp

end program ref_harness
