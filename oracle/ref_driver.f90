! Test-infrastructure driver for the reference's own FP64 Fortran core.
!
! This file is part of the oracle: it is compiled together with the
! reference modules (oracle/ref.mk) into oracle/_ref/libsfref.so and is
! called only by tests, golden-vector scripts and the CPU-baseline leg of
! bench.py.  It exposes the reference routines through bind(C) entry points
! (column-major arrays, plain pointers) and adds the time-stepping driver that
! the reference leaves to its callers:
!   * Euler exactly as src/specfabpy/integrator.py:73-77 and
!     demo/fabric-evolution/run-fabric-evolution-ddrx.f90:136-137
!     (steady operators built once per parcel, M_DDRX rebuilt from the
!     current state, dense matmul);
!   * classic 4-stage RK4 with the same per-stage operator rebuild.
! One truncation L is active per process (the reference keeps module globals,
! src/header.f90:16, src/dynamics.f90:17-18).
module sfref_driver

    use iso_c_binding
    use header
    use moments
    use dynamics
    use homogenizations
    use enhancementfactors
    use frames
    use elasticities
    use wavepropagation
    use rotation
    use idealstate
    use deformationmodes
    use reducedform

    implicit none

    integer, private :: Lcur = -1, ncur = 0

contains

    subroutine sfref_init(L) bind(C, name="sfref_init")
        integer(c_int), value :: L
        call initmoments(L)
        call initdynamics(L)
        call inithomogenizations(L)
        call initreduced(L)
        Lcur = L
        ncur = (L+1)*(L+2)/2
    end subroutine

    function sfref_nlm_len() bind(C, name="sfref_nlm_len") result(n)
        integer(c_int) :: n
        n = ncur
    end function

    ! ---- operators (column-major n x n outputs) ----

    subroutine sfref_M_LROT(eps, omg, iota, zeta, M) bind(C, name="sfref_M_LROT")
        real(c_double), intent(in) :: eps(3,3), omg(3,3)
        real(c_double), value :: iota, zeta
        complex(c_double_complex), intent(out) :: M(ncur,ncur)
        M = M_LROT(eps, omg, iota, zeta)
    end subroutine

    subroutine sfref_M_DDRX(nlm, tau, M) bind(C, name="sfref_M_DDRX")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(in) :: tau(3,3)
        complex(c_double_complex), intent(out) :: M(ncur,ncur)
        M = M_DDRX(nlm, tau)
    end subroutine

    subroutine sfref_M_DDRX_src(tau, M) bind(C, name="sfref_M_DDRX_src")
        real(c_double), intent(in) :: tau(3,3)
        complex(c_double_complex), intent(out) :: M(ncur,ncur)
        M = M_DDRX_src(tau)
    end subroutine

    subroutine sfref_M_REG(D, M) bind(C, name="sfref_M_REG")
        real(c_double), intent(in) :: D(3,3)
        real(c_double), intent(out) :: M(ncur,ncur)
        M = M_REG(D)
    end subroutine

    subroutine sfref_M_CDRX(M) bind(C, name="sfref_M_CDRX")
        real(c_double), intent(out) :: M(ncur,ncur)
        M = M_CDRX()
    end subroutine

    function sfref_ev_D2(nlm, tau) bind(C, name="sfref_ev_D2") result(D)
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(in) :: tau(3,3)
        real(c_double) :: D
        D = ev_D2(nlm, tau)
    end function

    ! ---- moments ----

    subroutine sfref_a2(nlm, a2out) bind(C, name="sfref_a2")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(out) :: a2out(3,3)
        a2out = a2(nlm)
    end subroutine

    subroutine sfref_a4(nlm, a4out) bind(C, name="sfref_a4")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(out) :: a4out(3,3,3,3)
        a4out = a4(nlm)
    end subroutine

    subroutine sfref_ev_ck_Mandel(nlm, a2v, a4v) bind(C, name="sfref_ev_ck_Mandel")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(out) :: a2v(6), a4v(6,6)
        call f_ev_ck_Mandel(nlm, a2v, a4v)
    end subroutine

    subroutine sfref_a2_to_nlm(a2in, nlm) bind(C, name="sfref_a2_to_nlm")
        real(c_double), intent(in) :: a2in(3,3)
        complex(c_double_complex), intent(out) :: nlm(6)
        nlm = a2_to_nlm(a2in)
    end subroutine

    subroutine sfref_a4_to_nlm(a4in, nlm) bind(C, name="sfref_a4_to_nlm")
        real(c_double), intent(in) :: a4in(3,3,3,3)
        complex(c_double_complex), intent(out) :: nlm(15)
        nlm = a4_to_nlm(a4in)
    end subroutine

    ! ---- rate factor, bounds, kinematics ----

    function sfref_Gamma0(eps, T, A, Q) bind(C, name="sfref_Gamma0") result(G)
        real(c_double), intent(in) :: eps(3,3)
        real(c_double), value :: T, A, Q
        real(c_double) :: G
        G = Gamma0(eps, T, A, Q)
    end function

    subroutine sfref_apply_bounds(nlm, out) bind(C, name="sfref_apply_bounds")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        complex(c_double_complex), intent(out) :: out(ncur)
        out = apply_bounds(nlm)
    end subroutine

    subroutine sfref_pureshear_F(ax, q, tau, t, F) bind(C, name="sfref_pureshear_F")
        integer(c_int), value :: ax
        real(c_double), value :: q, tau, t
        real(c_double), intent(out) :: F(3,3)
        F = pureshear_F(ax, q, tau, t)
    end subroutine

    subroutine sfref_pureshear_ugrad(ax, q, tau, ugrad) bind(C, name="sfref_pureshear_ugrad")
        integer(c_int), value :: ax
        real(c_double), value :: q, tau
        real(c_double), intent(out) :: ugrad(3,3)
        ugrad = pureshear_ugrad(ax, q, tau)
    end subroutine

    function sfref_pureshear_strainii_to_t(s, tau) bind(C, name="sfref_pureshear_strainii_to_t") result(t)
        real(c_double), value :: s, tau
        real(c_double) :: t
        t = pureshear_strainii_to_t(s, tau)
    end function

    subroutine sfref_simpleshear_F(plane, tau, t, F) bind(C, name="sfref_simpleshear_F")
        integer(c_int), value :: plane
        real(c_double), value :: tau, t
        real(c_double), intent(out) :: F(3,3)
        F = simpleshear_F(plane, tau, t)
    end subroutine

    subroutine sfref_simpleshear_ugrad(plane, tau, ugrad) bind(C, name="sfref_simpleshear_ugrad")
        integer(c_int), value :: plane
        real(c_double), value :: tau
        real(c_double), intent(out) :: ugrad(3,3)
        ugrad = simpleshear_ugrad(plane, tau)
    end subroutine

    function sfref_simpleshear_gamma_to_t(gam, tau) bind(C, name="sfref_simpleshear_gamma_to_t") result(t)
        real(c_double), value :: gam, tau
        real(c_double) :: t
        t = simpleshear_gamma_to_t(gam, tau)
    end function

    ! ---- homogenization ----

    subroutine sfref_eig(nlm, ei, lami) bind(C, name="sfref_eig")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(out) :: ei(3,3), lami(3)
        call eig(nlm, ei, lami)
    end subroutine

    ! Evw_tranisotropic(v, w, tau, nlm, Eij_grain, alpha, n_grain) (src/enhancementfactors.f90:47-69)
    subroutine sfref_Evw_tranisotropic(nlm, v, w, tau, Eg, alpha, ng, E) bind(C, name="sfref_Evw_tranisotropic")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(in) :: v(3), w(3), tau(3,3), Eg(2)
        real(c_double), value :: alpha
        integer(c_int), value :: ng
        real(c_double), intent(out) :: E
        E = Evw_tranisotropic(v, w, tau, nlm, Eg, alpha, ng)
    end subroutine

    ! Evw_orthotropic(v, w, tau, q1, q2, q3, Eij_grain, alpha, n_grain) (src/enhancementfactors.f90:134-160)
    subroutine sfref_Evw_orthotropic(q1, q2, q3, v, w, tau, Eg, alpha, ng, E) bind(C, name="sfref_Evw_orthotropic")
        complex(c_double_complex), intent(in) :: q1(ncur), q2(ncur), q3(ncur)
        real(c_double), intent(in) :: v(3), w(3), tau(3,3), Eg(6)
        real(c_double), value :: alpha
        integer(c_int), value :: ng
        real(c_double), intent(out) :: E
        E = Evw_orthotropic(v, w, tau, q1, q2, q3, Eg, alpha, ng)
    end subroutine

    ! eig3(M, e1,e2,e3, lami) (src/frames.f90:62-80) -> ei(3,3) rows e1, e2, e3
    subroutine sfref_eig3(M, ei, lami) bind(C, name="sfref_eig3")
        real(c_double), intent(in) :: M(3,3)
        real(c_double), intent(out) :: ei(3,3), lami(3)
        real(dp) :: e1(3), e2(3), e3(3)
        call eig3(M, e1, e2, e3, lami)
        ei(1,:) = e1
        ei(2,:) = e2
        ei(3,:) = e3
    end subroutine

    ! eigframe(M, plane, ei, lami) (src/frames.f90:24-60); plane 0 = 'ij', 1 = 'xy', 2 = 'xz'
    subroutine sfref_eigframe(M, iplane, ei, lami) bind(C, name="sfref_eigframe")
        real(c_double), intent(in) :: M(3,3)
        integer(c_int), value :: iplane
        real(c_double), intent(out) :: ei(3,3), lami(3)
        character(len=2) :: plane
        plane = 'ij'
        if (iplane == 1) plane = 'xy'
        if (iplane == 2) plane = 'xz'
        call eigframe(M, plane, ei, lami)
    end subroutine

    ! nlm_LROT(nlm0, dt, Nt, D, W, iota) (src/dynamics.f90:99-110) -> nlm(Nt, ncur)
    subroutine sfref_nlm_LROT(nlm0, dt, Nt, D, W, iota, nlm) bind(C, name="sfref_nlm_LROT")
        complex(c_double_complex), intent(in) :: nlm0(ncur)
        real(c_double), value :: dt, iota
        integer(c_int), value :: Nt
        real(c_double), intent(in) :: D(Nt,3,3), W(Nt,3,3)
        complex(c_double_complex), intent(out) :: nlm(Nt, ncur)
        nlm = nlm_LROT(nlm0, dt, Nt, D, W, iota)
    end subroutine

    subroutine sfref_Eij_tranisotropic(nlm, e1, e2, e3, Eij_grain, alpha, n_grain, Eij) &
            bind(C, name="sfref_Eij_tranisotropic")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(in) :: e1(3), e2(3), e3(3), Eij_grain(2)
        real(c_double), value :: alpha
        integer(c_int), value :: n_grain
        real(c_double), intent(out) :: Eij(6)
        Eij = Eij_tranisotropic(nlm, e1, e2, e3, Eij_grain, alpha, n_grain)
    end subroutine

    ! Eigenframe + Eij for P parcels (nlm(ncur,P)), the reference call pattern
    ! of src/specfabpy/common.py:13-37 followed by Eij_tranisotropic.
    subroutine sfref_Eij_eigframe_batch(P, nlm, Eij_grain, alpha, n_grain, Eij, lami) &
            bind(C, name="sfref_Eij_eigframe_batch")
        integer(c_int), value :: P, n_grain
        complex(c_double_complex), intent(in) :: nlm(ncur, P)
        real(c_double), intent(in) :: Eij_grain(2)
        real(c_double), value :: alpha
        real(c_double), intent(out) :: Eij(6, P), lami(3, P)
        real(c_double) :: ei(3,3)
        integer :: p_
        do p_ = 1, P
            call eig(nlm(:,p_), ei, lami(:,p_))
            Eij(:,p_) = Eij_tranisotropic(nlm(:,p_), ei(1,:), ei(2,:), ei(3,:), Eij_grain, alpha, n_grain)
        end do
    end subroutine

    subroutine sfref_Cij_to_Lame_tranisotropic(Cij, Lame) bind(C, name="sfref_Cij_to_Lame_tranisotropic")
        real(c_double), intent(in) :: Cij(5)
        real(c_double), intent(out) :: Lame(5)
        call Cij_to_Lame_tranisotropic(Cij, Lame)
    end subroutine

    subroutine sfref_Vi_elastic_tranisotropic(nlm, alpha, Lame, rho, ndir, theta, phi, Vi) &
            bind(C, name="sfref_Vi_elastic_tranisotropic")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), value :: alpha, rho
        real(c_double), intent(in) :: Lame(5)
        integer(c_int), value :: ndir
        real(c_double), intent(in) :: theta(ndir), phi(ndir)
        real(c_double), intent(out) :: Vi(3, ndir)
        Vi = Vi_elastic_tranisotropic(nlm, alpha, Lame, rho, theta, phi)
    end subroutine

    subroutine sfref_Qnorm_tranisotropic(nlm, alpha, Lame, Q) bind(C, name="sfref_Qnorm_tranisotropic")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), value :: alpha
        real(c_double), intent(in) :: Lame(5)
        real(c_double), intent(out) :: Q(3,3)
        Q = Qnorm_tranisotropic(nlm, alpha, Lame)
    end subroutine

    subroutine sfref_Vi_electromagnetic_tranisotropic(nlm, eps_m, eps_t, mu, ndir, theta, phi, Vi) &
            bind(C, name="sfref_Vi_electromagnetic_tranisotropic")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), value :: eps_m, eps_t, mu
        integer(c_int), value :: ndir
        real(c_double), intent(in) :: theta(ndir), phi(ndir)
        real(c_double), intent(out) :: Vi(2, ndir)
        Vi = Vi_electromagnetic_tranisotropic(nlm, eps_m, eps_t, mu, theta, phi)
    end subroutine

    subroutine sfref_rotate_nlm4(nlm4, theta, phi, out) bind(C, name="sfref_rotate_nlm4")
        complex(c_double_complex), intent(in) :: nlm4(15)
        real(c_double), value :: theta, phi
        complex(c_double_complex), intent(out) :: out(15)
        out = rotate_nlm4(nlm4, theta, phi)
    end subroutine

    ! rotate_nlm(nlm, theta, phi) of a state of the active truncation (L <= 12)
    subroutine sfref_rotate_nlm(nlm, theta, phi, out) bind(C, name="sfref_rotate_nlm")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), value :: theta, phi
        complex(c_double_complex), intent(out) :: out(ncur)
        out = rotate_nlm(nlm, theta, phi)
    end subroutine

    ! a4_eigentensors(nlm, Q1..Q6, eigvals): Q(3,3,6) column-major, eigvals(6)
    subroutine sfref_a4_eigentensors(nlm, Q, ev) bind(C, name="sfref_a4_eigentensors")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(out) :: Q(3,3,6), ev(6)
        call a4_eigentensors(nlm, Q(:,:,1), Q(:,:,2), Q(:,:,3), Q(:,:,4), Q(:,:,5), Q(:,:,6), ev)
    end subroutine

    ! nlm_isvalid(nhat20, nhat40) of k (nhat20, nhat40) pairs -> isvalid(k) as 0/1
    subroutine sfref_nlm_isvalid(k, n20, n40, ok) bind(C, name="sfref_nlm_isvalid")
        integer(c_int), value :: k
        complex(c_double_complex), intent(in) :: n20(k), n40(k)
        integer(c_int), intent(out) :: ok(k)
        logical :: v(k)
        v = nlm_isvalid(n20, n40)
        ok = merge(1, 0, v)
    end subroutine

    ! ---- n' = 3 structure tensors, deformability, CAFFE ----

    ! f_ev_ck(nlm, 'f', ...): a2(3,3), a4(3,3,3,3), a6(3^6), a8(3^8) (column-major)
    subroutine sfref_ev_ck(nlm, a2, a4, a6, a8) bind(C, name="sfref_ev_ck")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(out) :: a2(3,3), a4(3,3,3,3), a6(3,3,3,3,3,3), a8(3,3,3,3,3,3,3,3)
        call f_ev_ck(nlm, 'f', a2, a4, a6, a8)
    end subroutine

    ! ev_D(nlm, tau, n_RSS)
    subroutine sfref_ev_D(nlm, tau, n, D) bind(C, name="sfref_ev_D")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(in) :: tau(3,3)
        integer(c_int), value :: n
        real(c_double), intent(out) :: D
        D = ev_D(nlm, tau, n)
    end subroutine

    ! E_CAFFE(nlm, eps, Emin, Emax, n_grain)
    subroutine sfref_E_CAFFE(nlm, eps, Emin, Emax, ng, E) bind(C, name="sfref_E_CAFFE")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        real(c_double), intent(in) :: eps(3,3)
        real(c_double), value :: Emin, Emax
        integer(c_int), value :: ng
        real(c_double), intent(out) :: E
        E = E_CAFFE(nlm, eps, Emin, Emax, ng)
    end subroutine

    ! ---- orthotropic grains ----

    ! joint structure tensors of two distributions (L >= 4 states of length ncur)
    subroutine sfref_ortho_moments(b, n, a4j, a4jc, a2o, a4o) bind(C, name="sfref_ortho_moments")
        complex(c_double_complex), intent(in) :: b(ncur), n(ncur)
        real(c_double), intent(out) :: a4j(3,3,3,3), a4jc(3,3,3,3), a2o(3,3), a4o(3,3,3,3)
        a4j = a4_joint(b, n)
        a4jc = a4_jointcross(b, n)
        a2o = a2_orth(b, n)
        a4o = a4_orth(b, n)
    end subroutine

    ! Eij_orthotropic(q1, q2, q3, e1, e2, e3, Eij_grain(6), alpha, n_grain)
    subroutine sfref_Eij_orthotropic(q1, q2, q3, e1, e2, e3, Eg, alpha, ng, Eij) bind(C, name="sfref_Eij_orthotropic")
        complex(c_double_complex), intent(in) :: q1(ncur), q2(ncur), q3(ncur)
        real(c_double), intent(in) :: e1(3), e2(3), e3(3), Eg(6)
        real(c_double), value :: alpha
        integer(c_int), value :: ng
        real(c_double), intent(out) :: Eij(6)
        Eij = Eij_orthotropic(q1, q2, q3, e1, e2, e3, Eg, alpha, ng)
    end subroutine

    ! Eij_orthotropic_discrete(bi(N,3), ni(N,3), vi(N,3), e1, e2, e3, Eij_grain(6), alpha, n_grain)
    subroutine sfref_Eij_orthotropic_discrete(N, bi, ni, vi, e1, e2, e3, Eg, alpha, ng, Eij) &
            bind(C, name="sfref_Eij_orthotropic_discrete")
        integer(c_int), value :: N
        real(c_double), intent(in) :: bi(N,3), ni(N,3), vi(N,3), e1(3), e2(3), e3(3), Eg(6)
        real(c_double), value :: alpha
        integer(c_int), value :: ng
        real(c_double), intent(out) :: Eij(6)
        Eij = Eij_orthotropic_discrete(bi, ni, vi, e1, e2, e3, Eg, alpha, ng)
    end subroutine

    ! Vi_elastic_orthotropic(q1, q2, q3, alpha, Lame(9), rho, theta, phi) -> Vi(3, ndir)
    subroutine sfref_Vi_elastic_orthotropic(q1, q2, q3, alpha, Lame, rho, ndir, theta, phi, Vi) &
            bind(C, name="sfref_Vi_elastic_orthotropic")
        complex(c_double_complex), intent(in) :: q1(ncur), q2(ncur), q3(ncur)
        real(c_double), value :: alpha, rho
        real(c_double), intent(in) :: Lame(9)
        integer(c_int), value :: ndir
        real(c_double), intent(in) :: theta(ndir), phi(ndir)
        real(c_double), intent(out) :: Vi(3, ndir)
        Vi = Vi_elastic_orthotropic(q1, q2, q3, alpha, Lame, rho, theta, phi)
    end subroutine

    ! Vi_elastic_orthotropic__discrete(mi(3,3,N), alpha, Lame(9), rho, theta, phi) -> Vi(3, ndir)
    subroutine sfref_Vi_elastic_orthotropic_discrete(N, mi, alpha, Lame, rho, ndir, theta, phi, Vi) &
            bind(C, name="sfref_Vi_elastic_orthotropic_discrete")
        integer(c_int), value :: N, ndir
        real(c_double), intent(in) :: mi(3,3,N), Lame(9), theta(ndir), phi(ndir)
        real(c_double), value :: alpha, rho
        real(c_double), intent(out) :: Vi(3, ndir)
        Vi = Vi_elastic_orthotropic__discrete(mi, alpha, Lame, rho, theta, phi)
    end subroutine

    ! ---- discrete grains ----

    ! ri_LROT(ri0, dt, Nt, D, W, iota): ri0(N,3), D/W(Nt,3,3) -> ri(Nt,N,3)
    subroutine sfref_ri_LROT(N, Nt, ri0, dt, D, W, iota, ri) bind(C, name="sfref_ri_LROT")
        integer(c_int), value :: N, Nt
        real(c_double), intent(in) :: ri0(N,3), D(Nt,3,3), W(Nt,3,3)
        real(c_double), value :: dt, iota
        real(c_double), intent(out) :: ri(Nt,N,3)
        ri = ri_LROT(ri0, dt, Nt, D, W, iota)
    end subroutine

    ! ri_to_nlm(ri, wi, L): ri(N,3), wi(N) -> nlm((L+1)(L+2)/2)
    subroutine sfref_ri_to_nlm(N, L, ri, wi, nlm) bind(C, name="sfref_ri_to_nlm")
        integer(c_int), value :: N, L
        real(c_double), intent(in) :: ri(N,3), wi(N)
        complex(c_double_complex), intent(out) :: nlm((L+1)*(L+2)/2)
        nlm = ri_to_nlm(ri, wi, L)
    end subroutine

    ! ---- time stepping ----
    !
    ! procs bitmask: 1 = LROT, 2 = DDRX, 4 = CDRX, 8 = REG.  scheme: 0 = Euler,
    ! 1 = RK4.  ugrad(3,3,P) column-major velocity gradients; D = sym, W = skew,
    ! stress tau := D (coaxial, src/specfabpy/integrator.py:38-39).
    ! nlm(ncur,P) is advanced in place by nsteps steps of size dt.
    subroutine sfref_evolve(P, nsteps, scheme, procs, dt, iota, zeta, ugrad, gam0, lam, nlm) &
            bind(C, name="sfref_evolve")
        integer(c_int), value :: P, nsteps, scheme, procs
        real(c_double), value :: dt, iota, zeta
        real(c_double), intent(in) :: ugrad(3,3,P), gam0(P), lam(P)
        complex(c_double_complex), intent(inout) :: nlm(ncur, P)
        complex(c_double_complex), allocatable :: Mlin(:,:), M(:,:)
        complex(c_double_complex), allocatable :: x0(:), x(:), k1(:), k2(:), k3(:), k4(:)
        real(c_double) :: D(3,3), W(3,3)
        integer :: p_, t_

        allocate(Mlin(ncur,ncur), M(ncur,ncur))
        allocate(x0(ncur), x(ncur), k1(ncur), k2(ncur), k3(ncur), k4(ncur))

        do p_ = 1, P
            D = (ugrad(:,:,p_) + transpose(ugrad(:,:,p_)))/2
            W = (ugrad(:,:,p_) - transpose(ugrad(:,:,p_)))/2
            Mlin = (0.0d0, 0.0d0)
            if (iand(procs, 1) /= 0) Mlin = Mlin + M_LROT(D, W, iota, zeta)
            if (iand(procs, 4) /= 0) Mlin = Mlin + lam(p_)*M_CDRX()
            if (iand(procs, 8) /= 0) Mlin = Mlin + M_REG(D)
            x0 = nlm(:, p_)
            do t_ = 1, nsteps
                if (scheme == 0) then
                    call rate(x0, k1)
                    x0 = x0 + dt*k1
                else
                    call rate(x0, k1)
                    x = x0 + (dt/2)*k1
                    call rate(x, k2)
                    x = x0 + (dt/2)*k2
                    call rate(x, k3)
                    x = x0 + dt*k3
                    call rate(x, k4)
                    x0 = x0 + (dt/6)*(k1 + 2*k2 + 2*k3 + k4)
                end if
            end do
            nlm(:, p_) = x0
        end do

    contains

        subroutine rate(xin, kout)
            complex(c_double_complex), intent(in) :: xin(ncur)
            complex(c_double_complex), intent(out) :: kout(ncur)
            if (iand(procs, 2) /= 0) then
                M = Mlin + gam0(p_)*M_DDRX(xin, D)
                kout = matmul(M, xin)
            else
                kout = matmul(Mlin, xin)
            end if
        end subroutine

    end subroutine

    ! ---- reduced form (src/reducedform.f90:76-165) ----

    function sfref_rnlm_len() bind(C, name="sfref_rnlm_len") result(r)
        integer(c_int) :: r
        r = rnlm_len
    end function

    subroutine sfref_reduce_M(M, Mrr, Mri, Mir, Mii) bind(C, name="sfref_reduce_M")
        complex(c_double_complex), intent(in) :: M(ncur,ncur)
        real(c_double), intent(out), dimension(rnlm_len,rnlm_len) :: Mrr, Mri, Mir, Mii
        call reduce_M(M, Mrr, Mri, Mir, Mii)
    end subroutine

    subroutine sfref_nlm_to_rnlm(nlm, rnlm) bind(C, name="sfref_nlm_to_rnlm")
        complex(c_double_complex), intent(in) :: nlm(ncur)
        complex(c_double_complex), intent(out) :: rnlm(rnlm_len)
        rnlm = nlm_to_rnlm(nlm)
    end subroutine

    subroutine sfref_rnlm_to_nlm(rnlm, nlm) bind(C, name="sfref_rnlm_to_nlm")
        complex(c_double_complex), intent(in) :: rnlm(rnlm_len)
        complex(c_double_complex), intent(out) :: nlm(ncur)
        nlm = rnlm_to_nlm(rnlm)
    end subroutine

end module sfref_driver
