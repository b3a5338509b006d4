! two_e_slater_hip: the reference's two_e_slater program flow on the MI355X
! engine.  Reads the same input deck (./j_in, <MOLECULE>.xyz, basis file;
! src/two_e_slater.F90:101-154, src/read_geometry.F90, src/read_basis.F90),
! builds the same basis arrays (src/build_gaussian_expansion_of_orbitals.F90,
! src/build_c_g_gauss_STO.F90), runs the analytic Gaussian-fit ERIs with the
! Schwarz driver, the ZVMC blocks and the PSD cleanup through the C-ABI, and
! writes ./bielec_ao in the reference's format (src/two_e_slater.F90:605-613).
!
! Environment:
!   ZVMC_FIT_TABLES   path of sto_fit_tables.txt (default: data/ next to the
!                     executable)
!   ZVMC_PRECISION    fp64_parity | fp64 (default) | fp32   (estimator)
!   ZVMC_SVD_CLEAN    0 to skip the cleanup (default: run it, as the reference)
!   ZVMC_ONE_E        1 to also write the one-electron integrals over the
!                     fits as the reference's one_e_slater does
!                     (overlap_ao_ex, nuclear_ao_ex, kinetic_ao_ex;
!                     src/one_e_slater.F90:44-46,123-136), with the deck's
!                     ng0 (one_e_slater always uses 20)
!   ZVMC_EXACT_MONOCENTER  1 for the reference's do_exact_monocenter branch
!                     (src/two_e_slater.F90:74,318-365: exact one-centre
!                     integrals, standard error 0); default off, as there
program two_e_slater_hip
  use iso_c_binding
  use zvmc
  implicit none

  integer, parameter :: nbl = 10, ng_max = 30
  character(2), parameter :: elements(36) = (/ 'H ', 'He', 'Li', 'Be', 'B ', 'C ', 'N ', 'O ', &
      'F ', 'Ne', 'Na', 'Mg', 'Al', 'Si', 'P ', 'S ', 'Cl', 'Ar', 'K ', 'Ca', 'Sc', 'Ti', 'V ',  &
      'Cr', 'Mn', 'Fe', 'Co', 'Ni', 'Cu', 'Zn', 'Ga', 'Ge', 'As', 'Se', 'Br', 'Kr' /)
  double precision, parameter :: ang_to_bohr = 1.d0/0.52917721092d0

  character(80) :: line, molecule
  character(128) :: fbasis, ftab, exe, env
  integer :: num_simulation, ng0, i_add_2p, i_add_3d, i_add_large_g, npts
  double precision :: g_thr, a_zv
  ! geometry / basis file
  integer :: natom, nline, ios, i, j, k, l, m, a, b, t, n, ldg, comp, ncomp, ngi
  character(2), allocatable :: atom(:)
  double precision, allocatable :: xyz(:,:)
  character(2) :: bel(2000)
  character(8) :: borb(2000)
  double precision :: bzeta(2000)
  ! fit tables
  double precision :: c_fit(ng_max, ng_max), g_fit(ng_max, ng_max), cf, gf
  integer :: tng, ti
  ! basis arrays (reference layouts)
  double precision, allocatable :: center(:,:), zeta(:), c_gauss(:,:,:), g_gauss(:,:)
  integer, allocatable :: npower(:,:), n_sto(:), i_type(:), n_gauss(:)
  integer :: comps(3, 15)
  ! quartets and results
  integer(c_int64_t) :: nint, kcp, nscr
  integer(c_int), allocatable :: is(:), js(:), ks(:), ls(:)
  double precision, allocatable :: ijkl_gaus(:), precond(:), mean(:), err(:), exact(:)
  double precision, allocatable :: s1e(:,:), v1e(:,:), k1e(:,:), qnuc(:)
  integer(c_int64_t) :: nmono
  double precision :: error_max, errmoy, dmoy, enucl, rij
  type(c_ptr) :: ctx
  integer(c_int) :: rc, prec, nrem, rank
  integer(c_int64_t) :: seed

  ! ---------------------------------------------------------------- input deck
  open(unit=5, file='j_in', status='old')
  read(5,'(a80)') line
  read(5,*) num_simulation
  read(5,'(a80)') line
  read(5,*) molecule
  read(5,'(a80)') line
  read(5,*) fbasis
  read(5,'(a80)') line
  read(5,*) ng0, i_add_2p, i_add_3d, g_thr, i_add_large_g
  read(5,'(a80)') line
  read(5,*) npts
  if (mod(npts, 40) /= 0) stop 'npts_two_elec must be a multiple of nw'
  read(5,'(a80)') line
  read(5,'(a80)') line
  read(5,*) a_zv
  close(5)

  open(unit=2, file=trim(molecule)//'.xyz', status='old')
  read(2,*) natom
  read(2,*)
  allocate(atom(natom), xyz(3,natom))
  do i = 1, natom
    read(2,*) atom(i), xyz(1,i), xyz(2,i), xyz(3,i)
    xyz(:,i) = ang_to_bohr*xyz(:,i)
  end do
  close(2)
  enucl = 0.d0
  do i = 1, natom
    do j = i+1, natom
      rij = dsqrt((xyz(1,i)-xyz(1,j))**2 + (xyz(2,i)-xyz(2,j))**2 + (xyz(3,i)-xyz(3,j))**2)
      enucl = enucl + dble(znum(atom(i)))*dble(znum(atom(j)))/rij
    end do
  end do
  print *, 'ENUCL=', enucl

  nline = 0
  open(unit=7, file=trim(fbasis), status='old')
  do
    read(7,*,iostat=ios) bel(nline+1), borb(nline+1), bzeta(nline+1)
    if (ios /= 0) exit
    nline = nline + 1
  end do
  close(7)

  ! ---------------------------------------------------------------- fit tables
  call get_environment_variable('ZVMC_FIT_TABLES', ftab)
  if (len_trim(ftab) == 0) then
    call get_command_argument(0, exe)
    t = index(exe, '/', back=.true.)
    ftab = exe(1:t)//'data/sto_fit_tables.txt'
  end if
  c_fit = 0.d0
  g_fit = 0.d0
  open(unit=8, file=trim(ftab), status='old')
  do
    read(8,'(a80)',iostat=ios) line
    if (ios /= 0) exit
    if (line(1:1) == '#') cycle
    read(line,*) tng, ti, cf, gf
    c_fit(ti, tng) = cf
    g_fit(ti, tng) = gf
  end do
  close(8)

  ! ---------------------------------------------------------------- basis
  n = 0
  do i = 1, natom
    do l = 1, nline
      if (znum(bel(l)) /= znum(atom(i))) cycle
      call shell(borb(l), ncomp, comps, m, t, a)
      n = n + ncomp
    end do
  end do
  ldg = ng_max
  allocate(center(3,n), zeta(n), c_gauss(ldg,n,2), g_gauss(ldg,n))
  allocate(npower(3,n), n_sto(n), i_type(n), n_gauss(n))
  c_gauss = 0.d0
  g_gauss = 0.d0
  k = 0
  do i = 1, natom
    do l = 1, nline
      if (znum(bel(l)) /= znum(atom(i))) cycle
      call shell(borb(l), ncomp, comps, m, t, a)
      do comp = 1, ncomp
        k = k + 1
        center(:,k) = xyz(:,i)
        npower(:,k) = comps(:,comp)
        n_sto(k) = m
        i_type(k) = t
        zeta(k) = bzeta(l)
        ngi = ng0
        if (a == 2) ngi = ng0 + i_add_2p
        if (a == 3) ngi = ng0 + i_add_3d
        if (bzeta(l) > g_thr) ngi = ngi + i_add_large_g
        if (ngi > ng_max) stop 'no Gaussian fit with that many functions'
        n_gauss(k) = ngi
        do b = 1, ngi
          g_gauss(b,k) = bzeta(l)*bzeta(l)*g_fit(b,ngi)
          if (t == 1) then
            c_gauss(b,k,1) = 1.d0*c_fit(b,ngi)
          else
            c_gauss(b,k,2) = 1.d0*2.d0*c_fit(b,ngi)*g_fit(b,ngi)*bzeta(l)
          end if
        end do
      end do
    end do
  end do
  print *, 'Number of basis functions ', n

  ! ---------------------------------------------------------------- device
  call zv_check(ctx, zv_create(0_c_int, ctx), 'zv_create')
  ! the reference prints the cleanup's progress lines unconditionally
  call zv_check(ctx, zv_set_verbose(ctx, 1_c_int), 'zv_set_verbose')
  call zv_check(ctx, zv_set_basis(ctx, n, center, npower, n_sto, i_type, zeta, n_gauss, ldg, &
      c_gauss, g_gauss, a_zv), 'zv_set_basis')
  call get_environment_variable('ZVMC_ONE_E', env)
  if (trim(env) == '1') then
    allocate(s1e(n,n), v1e(n,n), k1e(n,n), qnuc(natom))
    do i = 1, natom
      qnuc(i) = dble(znum(atom(i)))
    end do
    call zv_check(ctx, zv_one_e(ctx, int(natom,c_int), qnuc, xyz, s1e, v1e, k1e), 'zv_one_e')
    open(unit=104, file='overlap_ao_ex')
    open(unit=105, file='nuclear_ao_ex')
    open(unit=106, file='kinetic_ao_ex')
    do i = 1, n
      do k = 1, n
        write(104,'(2(I5,X),D22.15)') i, k, s1e(i,k)
        write(105,'(2(I5,X),D22.15)') i, k, v1e(i,k)
        write(106,'(2(I5,X),D22.15)') i, k, k1e(i,k)
      end do
    end do
    close(104)
    close(105)
    close(106)
    deallocate(s1e, v1e, k1e, qnuc)
  end if

  nint = zv_nint(n)
  allocate(is(nint), js(nint), ks(nint), ls(nint))
  allocate(ijkl_gaus(nint), precond(n*n), mean(nint), err(nint))
  call zv_check(ctx, zv_build_mapping(n, is, js, ks, ls), 'zv_build_mapping')

  call zv_check(ctx, zv_gauss_eri(ctx, ijkl_gaus, precond, nscr), 'zv_gauss_eri')
  print *, 'IJKL REMOVED DUE TO CAUCHY-SCHWARZ=', nscr, ' OVER', nint, 'IJKL'

  prec = ZV_PREC_FP64
  call get_environment_variable('ZVMC_PRECISION', env)
  if (trim(env) == 'fp64_parity') prec = ZV_PREC_FP64_PARITY
  if (trim(env) == 'fp32') prec = ZV_PREC_FP32
  if (trim(env) == 'bf16') prec = ZV_PREC_BF16
  seed = 1 + (num_simulation-1)*10000     ! src/two_e_slater.F90:115
  print *, ' Number of Monte Carlo steps= ', nbl, ' X ', npts
  call zv_check(ctx, zv_mc_blocks(ctx, nbl, int(npts,c_int64_t), ZV_SAMPLES_PHILOX, seed, &
      0_c_int64_t, prec, c_null_ptr, c_null_ptr, c_null_ptr, mean, err, c_null_ptr), &
      'zv_mc_blocks')

  ! exact one-centre quartets: value exact, error 0, and no Gaussian-fit part
  ! in the final sum (src/two_e_slater.F90:359-361, 489-493, 558-563)
  call get_environment_variable('ZVMC_EXACT_MONOCENTER', env)
  if (trim(env) == '1') then
    allocate(exact(nint))
    exact(:) = -huge(1.d0)
    call zv_check(ctx, zv_exact_monocenter(ctx, exact, err, nmono), 'zv_exact_monocenter')
    print *, 'One-center integrals computed exactly: ', nmono
    do kcp = 1, nint
      if (exact(kcp) /= -huge(1.d0)) then
        mean(kcp) = exact(kcp)
        ijkl_gaus(kcp) = 0.d0
      end if
    end do
    deallocate(exact)
  end if

  ! report as src/two_e_slater.F90:524-553
  error_max = -1.d0
  errmoy = 0.d0
  dmoy = 0.d0
  do kcp = 1, nint
    if (err(kcp) > error_max) error_max = err(kcp)
    if (dabs(mean(kcp)) > 1.d-6 .and. err(kcp) /= 0.d0) then
      errmoy = errmoy + err(kcp)
      dmoy = dmoy + 1.d0
    end if
  end do
  write(*,*) 'Error mean ijkl=', errmoy/dmoy, ' error_max ', error_max

  mean(:) = mean(:) + ijkl_gaus(:)
  call get_environment_variable('ZVMC_SVD_CLEAN', env)
  if (trim(env) /= '0') then
    print *, 'Cleaning ERI matrix'
    rank = min(512, n*n)
    call zv_check(ctx, zv_svd_clean(ctx, rank, 12_c_int, seed, mean, nrem), 'zv_svd_clean')
    print *, 'Removed ', nrem, ' components'
  end if

  open(unit=104, file='bielec_ao')
  do kcp = 1, nint
    write(104,'(4(I5,X),2D22.15)') is(kcp), js(kcp), ks(kcp), ls(kcp), mean(kcp), err(kcp)
  end do
  close(104)
  rc = zv_destroy(ctx)
  print *, 'ZVMC DONE'

contains

  integer function znum(lab)
    character(*), intent(in) :: lab
    integer :: q
    do q = 1, 36
      if (trim(adjustl(lab)) == trim(elements(q))) then
        znum = q
        return
      end if
    end do
    print *, ':'//trim(lab)//':'
    stop 'Atom not defined'
  end function

  ! Cartesian components, n_STO, i_type and ng class (1 s, 2 p, 3 d/f/g) of a
  ! shell name, src/build_gaussian_expansion_of_orbitals.F90:20-466.
  subroutine shell(orb, nc, cp, nsto, ityp, kind)
    character(*), intent(in) :: orb
    integer, intent(out) :: nc, cp(3,15), nsto, ityp, kind
    integer :: q
    integer, parameter :: pp(3,3) = reshape((/1,0,0, 0,1,0, 0,0,1/), (/3,3/))
    integer, parameter :: dd(3,6) = reshape((/2,0,0, 1,1,0, 1,0,1, 0,2,0, 0,1,1, 0,0,2/), (/3,6/))
    integer, parameter :: ff(3,10) = reshape((/3,0,0, 2,1,0, 2,0,1, 1,2,0, 1,1,1, 1,0,2, &
        0,3,0, 0,2,1, 0,1,2, 0,0,3/), (/3,10/))
    integer, parameter :: gg(3,15) = reshape((/4,0,0, 3,1,0, 2,2,0, 1,3,0, 0,4,0, 3,0,1, &
        2,1,1, 1,2,1, 0,3,1, 2,0,2, 1,1,2, 0,2,2, 1,0,3, 0,1,3, 0,0,4/), (/3,15/))
    cp = 0
    select case (trim(orb))
    case ('1S')
      nc = 1; nsto = 0; ityp = 1; kind = 1
    case ('2S')
      nc = 1; nsto = 1; ityp = 2; kind = 1
    case ('3S')
      nc = 1; nsto = 2; ityp = 2; kind = 1
    case ('2P')
      nc = 3; nsto = 0; ityp = 1; kind = 2
      do q = 1, 3
        cp(:,q) = pp(:,q)
      end do
    case ('3P')
      nc = 3; nsto = 1; ityp = 2; kind = 2
      do q = 1, 3
        cp(:,q) = pp(:,q)
      end do
    case ('3D')
      nc = 6; nsto = 0; ityp = 1; kind = 3
      do q = 1, 6
        cp(:,q) = dd(:,q)
      end do
    case ('4F')
      nc = 10; nsto = 0; ityp = 1; kind = 3
      do q = 1, 10
        cp(:,q) = ff(:,q)
      end do
    case ('5G')
      nc = 15; nsto = 0; ityp = 1; kind = 3
      do q = 1, 15
        cp(:,q) = gg(:,q)
      end do
    case default
      print *, 'orbital ', trim(orb)
      stop 'orbital type not supported'
    end select
  end subroutine

end program two_e_slater_hip
