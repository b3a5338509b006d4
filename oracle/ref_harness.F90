! C-callable harness around the compiled reference (scemama/slater_stoch).
!
! TEST INFRASTRUCTURE ONLY.  oracle/Makefile.ref compiles the reference's
! own Fortran sources where they lie (src/*.F90 under the mounted reference
! tree) together with this file into oracle/_ref/libref.so.  The harness is
! new code: it sets the reference's module state through the reference's
! own input routines and exposes single routines (gauss_ijkl,
! compute_densities, compute_jacobian_and_pi0, rint, dierfc, ...) with
! bind(C) entry points, so tests can produce golden vectors from the
! reference itself.  The estimator of src/two_e_slater.F90:386-496 is inline
! in the reference main program, so ref_mc_fixed below drives the reference
! subroutines with a caller-supplied sample list and applies the estimator
! formula (FP32 casts included) in this file.
module ref_harness
  use iso_c_binding
  use common_data
  implicit none
  double precision, allocatable :: h_ut1(:,:,:), h_ut2(:,:,:)
  double precision, allocatable :: h_rho(:,:,:), h_poly(:,:,:), h_rhog(:,:,:)
  double precision, allocatable :: h_jac1(:,:,:), h_jac2(:,:,:)
  double precision, allocatable :: h_gc(:,:,:)
contains

  subroutine c2f(cstr, fstr)
    character(kind=c_char), intent(in) :: cstr(*)
    character(len=*), intent(out) :: fstr
    integer :: t
    fstr = ' '
    do t = 1, len(fstr)
      if (cstr(t) == c_null_char) exit
      fstr(t:t) = cstr(t)
    end do
  end subroutine

  subroutine free_state()
    if (allocated(ATOM)) deallocate(ATOM)
    if (allocated(centers_nuclei)) deallocate(centers_nuclei)
    if (allocated(charge)) deallocate(charge)
    if (allocated(n_b)) deallocate(n_b)
    if (allocated(n_cont_b)) deallocate(n_cont_b)
    if (allocated(gamma_b)) deallocate(gamma_b)
    if (allocated(coef_b)) deallocate(coef_b)
    if (allocated(orb_b)) deallocate(orb_b)
    if (allocated(nucleus_number)) deallocate(nucleus_number)
    if (allocated(n_STO)) deallocate(n_STO)
    if (allocated(npower)) deallocate(npower)
    if (allocated(i_type)) deallocate(i_type)
    if (allocated(n_contract)) deallocate(n_contract)
    if (allocated(n_gauss)) deallocate(n_gauss)
    if (allocated(ng)) deallocate(ng)
    if (allocated(center)) deallocate(center)
    if (allocated(orb_name)) deallocate(orb_name)
    if (allocated(orb_name_full)) deallocate(orb_name_full)
    if (allocated(c_gauss)) deallocate(c_gauss)
    if (allocated(g_gauss)) deallocate(g_gauss)
    if (allocated(g_slater)) deallocate(g_slater)
    if (allocated(g_contract)) deallocate(g_contract)
    if (allocated(c_contract)) deallocate(c_contract)
    if (allocated(is)) deallocate(is, js, ks, ls)
    if (allocated(h_ut1)) deallocate(h_ut1, h_ut2, h_rho, h_poly, h_rhog, h_jac1, h_jac2, h_gc)
  end subroutine

  ! Reads <mol>.xyz and the basis file with the reference's routines.
  integer(c_int) function ref_setup(mol, basis, ng0_, add2p, add3d, gthr, addlg, azv) &
      bind(C, name="ref_setup")
    character(kind=c_char), intent(in) :: mol(*), basis(*)
    integer(c_int), value :: ng0_, add2p, add3d, addlg
    real(c_double), value :: gthr, azv
    character(80) :: molecule
    character(128) :: fbasis
    integer :: i, k, ll
    call free_state()
    call c2f(mol, molecule)
    call c2f(basis, fbasis)
    ng0 = ng0_
    i_add_2p = add2p
    i_add_3d = add3d
    g_thr = gthr
    i_add_large_g = addlg
    a_ZV = azv
    call read_geometry(molecule)
    call read_basis(fbasis)
    call build_gaussian_expansion_of_orbitals()
    allocate(h_ut1(nw,3,nbasis*nbasis), h_ut2(nw,3,nbasis*nbasis))
    allocate(h_rho(nw,nbasis*nbasis,2), h_poly(nw,nbasis*nbasis,2), h_rhog(nw,nbasis*nbasis,2))
    allocate(h_jac1(nw,nbasis,nbasis), h_jac2(nw,nbasis,nbasis), h_gc(3,nbasis,nbasis))
    do k = 1, nbasis
      do i = 1, nbasis
        do ll = 1, 3
          h_gc(ll,i,k) = (g_slater(i)*center(ll,i)+g_slater(k)*center(ll,k))/(g_slater(i)+g_slater(k))
        end do
      end do
    end do
    ref_setup = nbasis
  end function

  real(c_double) function ref_enucl() bind(C, name="ref_enucl")
    ref_enucl = enucl
  end function

  subroutine ref_get_basis(ldg, cen, npw, nsto, ityp, zeta, ngs, nuc, c1, c2, gg) &
      bind(C, name="ref_get_basis")
    integer(c_int), value :: ldg
    real(c_double), intent(out) :: cen(3,*), zeta(*), c1(ldg,*), c2(ldg,*), gg(ldg,*)
    integer(c_int), intent(out) :: npw(3,*), nsto(*), ityp(*), ngs(*), nuc(*)
    integer :: i, m
    do i = 1, nbasis
      cen(1:3,i) = center(1:3,i)
      npw(1:3,i) = npower(1:3,i)
      nsto(i) = n_STO(i)
      ityp(i) = i_type(i)
      zeta(i) = g_slater(i)
      ngs(i) = n_gauss(i)
      nuc(i) = nucleus_number(i)
      do m = 1, ldg
        c1(m,i) = 0.d0
        c2(m,i) = 0.d0
        gg(m,i) = 0.d0
      end do
      do m = 1, min(ldg, n_gauss(i))
        c1(m,i) = c_gauss(m,i,1)
        c2(m,i) = c_gauss(m,i,2)
        gg(m,i) = g_gauss(m,i)
      end do
    end do
  end subroutine

  real(c_double) function ref_gauss_ijkl(i, k, j, l) bind(C, name="ref_gauss_ijkl")
    integer(c_int), value :: i, k, j, l
    double precision, external :: gauss_ijkl
    ref_gauss_ijkl = gauss_ijkl(i, k, j, l)
  end function

  ! Exact one-centre Slater integral (ik|jl) of the loaded basis, 1-based
  ! indices (src/compute_int_slater.F90, the do_exact_monocenter branch of
  ! src/two_e_slater.F90:350-365).
  real(c_double) function ref_int_slater(i, k, j, l) bind(C, name="ref_int_slater")
    integer(c_int), value :: i, k, j, l
    real(c_double) :: v
    call compute_int_slater(int(i), int(k), int(j), int(l), v)
    ref_int_slater = v
  end function

  ! Overlap, nuclear attraction and kinetic integrals of the loaded basis
  ! over the Gaussian fits, S(i,k), V(i,k), K(i,k) column-major n x n
  ! (src/one_elect.F90, the loop of src/one_e_slater.F90:87-95).
  subroutine ref_one_e(s_out, v_out, k_out) bind(C, name="ref_one_e")
    real(c_double), intent(out) :: s_out(nbasis, nbasis), v_out(nbasis, nbasis), &
                                   k_out(nbasis, nbasis)
    integer :: i, k
    double precision :: sij, vij, kij
    do i = 1, nbasis
      do k = 1, nbasis
        call one_elect(i, k, sij, vij, kij)
        s_out(i, k) = sij
        v_out(i, k) = vij
        k_out(i, k) = kij
      end do
    end do
  end subroutine

  real(c_double) function ref_rint(n, rho) bind(C, name="ref_rint")
    integer(c_int), value :: n
    real(c_double), value :: rho
    double precision, external :: rint
    ref_rint = rint(n, rho)
  end function

  real(c_double) function ref_dierfc(y) bind(C, name="ref_dierfc")
    real(c_double), value :: y
    double precision, external :: dierfc
    ref_dierfc = dierfc(y)
  end function

  integer(c_int) function ref_compare_nuclei(ni, nj, nk, nl) bind(C, name="ref_compare_nuclei")
    integer(c_int), value :: ni, nj, nk, nl
    integer :: nd
    call compare_nuclei(ni, nj, nk, nl, nd)
    ref_compare_nuclei = nd
  end function

  integer(c_int64_t) function ref_build_mapping(ois, ojs, oks, ols) bind(C, name="ref_build_mapping")
    integer(c_int), intent(out) :: ois(*), ojs(*), oks(*), ols(*)
    integer*8 :: nn, q
    if (allocated(is)) deallocate(is, js, ks, ls)
    call build_mapping_ijkl(nn)
    do q = 1, nn
      ois(q) = is(q)
      ojs(q) = js(q)
      oks(q) = ks(q)
      ols(q) = ls(q)
    end do
    ref_build_mapping = nn
  end function

  ! Fills the per-pair arrays for one walker block of nw samples, as the
  ! reference driver does before its estimator loop.
  subroutine fill_block(r1, r2, pi0)
    double precision, intent(in) :: r1(nw,3), r2(nw,3)
    double precision, intent(out) :: pi0(nw)
    integer :: i, k, kw, ik
    double precision :: factor, m1, m2
    do k = 1, nbasis
      do i = k, nbasis
        ik = (k-1)*nbasis + i
        factor = 0.5d0 * a_ZV / (g_slater(i)+g_slater(k))
        do kw = 1, nw
          m1 = dsqrt(r1(kw,1)*r1(kw,1)+r1(kw,2)*r1(kw,2)+r1(kw,3)*r1(kw,3))
          m2 = dsqrt(r2(kw,1)*r2(kw,1)+r2(kw,2)*r2(kw,2)+r2(kw,3)*r2(kw,3))
          h_ut1(kw,1:3,ik) = factor*r1(kw,1:3) * m1 + h_gc(1:3,i,k)
          h_ut2(kw,1:3,ik) = factor*r2(kw,1:3) * m2 + h_gc(1:3,i,k)
        end do
        call compute_densities(i, k, h_ut1, h_ut2, h_rho, h_rhog, h_poly)
      end do
    end do
    call compute_jacobian_and_pi0(r1, r2, h_jac1, h_jac2, pi0, nbasis)
  end subroutine

  ! Pair data for ns samples (ns multiple of nw); samples as r(3,ns).
  subroutine ref_pair_samples(i, k, ns, r1, r2, rho1, rho2, rhog1, rhog2, jac1, jac2, pi0) &
      bind(C, name="ref_pair_samples")
    integer(c_int), value :: i, k, ns
    real(c_double), intent(in) :: r1(3,*), r2(3,*)
    real(c_double), intent(out) :: rho1(*), rho2(*), rhog1(*), rhog2(*), jac1(*), jac2(*), pi0(*)
    double precision :: a(nw,3), b(nw,3), p0(nw)
    integer :: blk, kw, ik, m
    ik = (k-1)*nbasis + i
    do blk = 0, ns/nw - 1
      do kw = 1, nw
        a(kw,1:3) = r1(1:3, blk*nw+kw)
        b(kw,1:3) = r2(1:3, blk*nw+kw)
      end do
      call fill_block(a, b, p0)
      do kw = 1, nw
        m = blk*nw + kw
        rho1(m) = h_rho(kw,ik,1)
        rho2(m) = h_rho(kw,ik,2)
        rhog1(m) = h_rhog(kw,ik,1)
        rhog2(m) = h_rhog(kw,ik,2)
        jac1(m) = h_jac1(kw,i,k)
        jac2(m) = h_jac2(kw,i,k)
        pi0(m) = p0(kw)
      end do
    end do
  end subroutine

  ! nbl blocks of npts samples each from the caller's list (r(3,nbl*npts),
  ! pi0 as given); returns the block mean and the reference's stderr.
  subroutine ref_mc_fixed(nbl, npts, r1, r2, pi0in, nq, qi, qj, qk, ql, mean, err) &
      bind(C, name="ref_mc_fixed")
    integer(c_int), value :: nbl, npts
    integer(c_int64_t), value :: nq
    real(c_double), intent(in) :: r1(3,*), r2(3,*), pi0in(*)
    integer(c_int), intent(in) :: qi(*), qj(*), qk(*), ql(*)
    real(c_double), intent(out) :: mean(*), err(*)
    double precision, allocatable :: es(:), eg(:), s1(:), s2(:)
    double precision :: a(nw,3), b(nw,3), p0(nw), pin(nw), w(nw), wg(nw), rinv(nw)
    double precision :: d(3), r12, fac, x, scal, jac
    integer :: blk, kk, kw, i, j, k, l, ik, jl
    integer*8 :: q, m0
    allocate(es(nq), eg(nq), s1(nq), s2(nq))
    s1 = 0.d0
    s2 = 0.d0
    do blk = 1, nbl
      es = 0.d0
      eg = 0.d0
      do kk = 1, npts/nw
        m0 = int(blk-1,8)*npts + int(kk-1,8)*nw
        do kw = 1, nw
          a(kw,1:3) = r1(1:3, m0+kw)
          b(kw,1:3) = r2(1:3, m0+kw)
          pin(kw) = pi0in(m0+kw)
        end do
        call fill_block(a, b, p0)
        do q = 1, nq
          i = qi(q); j = qj(q); k = qk(q); l = ql(q)
          ik = (k-1)*nbasis + i
          jl = (l-1)*nbasis + j
          do kw = 1, nw
            d(1:3) = h_ut1(kw,1:3,ik) - h_ut2(kw,1:3,jl)
            r12 = d(1)*d(1) + d(2)*d(2) + d(3)*d(3)
            jac = h_jac1(kw,i,k)*h_jac2(kw,j,l)
            fac = real(real(jac,4)/real(pin(kw),4),8)
            w(kw) = fac*h_rho(kw,ik,1)*h_rho(kw,jl,2)
            wg(kw) = fac*h_rhog(kw,ik,1)*h_rhog(kw,jl,2)
            rinv(kw) = real(1./sqrt(real(r12,4)),8)
          end do
          es(q) = es(q) + sum(w(1:nw)*rinv(1:nw))
          eg(q) = eg(q) + sum(wg(1:nw)*rinv(1:nw))
        end do
      end do
      scal = 1.d0/dble(npts)
      do q = 1, nq
        x = es(q)*scal - eg(q)*scal
        s1(q) = s1(q) + x
        s2(q) = s2(q) + x*x
      end do
    end do
    do q = 1, nq
      mean(q) = s1(q)/nbl
      err(q) = dsqrt(dabs(s2(q)/nbl - mean(q)*mean(q)))/dsqrt(dble(nbl))
    end do
    deallocate(es, eg, s1, s2)
  end subroutine

end module ref_harness
