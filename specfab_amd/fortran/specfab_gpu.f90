! specfab_gpu: Fortran face of libspecfab_gpu (include/specfab_gpu.h) through
! ISO_C_BINDING, mirroring the names of the reference's `use specfab` module
! API for the fabric-evolution path (src/specfab.f90, src/dynamics.f90,
! src/moments.f90, src/frames.f90, src/enhancementfactors.f90) so that a
! Fortran host code can switch by changing its `use` line:
!
!     use specfab_gpu
!     call initspecfab(12)                  ! as src/specfab.f90:37
!     M = M_LROT(eps, omg, 1d0, 0d0)        ! complex(dp) (nlm_len, nlm_len)
!     M = M + Gam0*M_DDRX(nlm, tau)
!
! plus the batched entry point sfg_evolve_parcels for many parcels at once.
! The C-ABI is row-major for 3x3 tensors and n x n operators, so the wrappers
! transpose (a Fortran array A(3,3) passed as-is is A^T in C order).
module specfab_gpu

    use iso_c_binding
    implicit none

    integer, parameter, private :: dp = 8
    type(c_ptr), private, save :: ctx = c_null_ptr
    integer, save :: Lcap = -1
    integer, save :: nlm_len = 0

    integer(c_int), parameter :: SFG_LROT = 1, SFG_DDRX = 2, SFG_CDRX = 4, SFG_REG = 8
    integer(c_int), parameter :: SFG_EULER = 0, SFG_RK4 = 1, SFG_FP64 = 0, SFG_FP32C = 1

    type, bind(C) :: sfg_evolve_desc
        integer(c_int) :: L
        integer(c_int64_t) :: P
        integer(c_int) :: procs, scheme, precision, nsteps
        real(c_double) :: dt, iota, zeta
        type(c_ptr) :: ugrad_dev, tau_dev, gam0_dev, lam_dev, steps_dev, x_dev, status_dev
        real(c_double) :: reg_nu, reg_expo
    end type

    interface
        integer(c_int) function sfg_ctx_create(device, pctx) bind(C, name="sfg_ctx_create")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr) :: pctx
        end function
        integer(c_int) function sfg_nlm_len(L) bind(C, name="sfg_nlm_len")
            import :: c_int
            integer(c_int), value :: L
        end function
        integer(c_int) function c_M_LROT(c, L, eps, omg, iota, zeta, M) bind(C, name="sfg_M_LROT")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L
            real(c_double) :: eps(9), omg(9)
            real(c_double), value :: iota, zeta
            complex(c_double_complex) :: M(*)
        end function
        integer(c_int) function c_M_DDRX(c, L, nlm, tau, M) bind(C, name="sfg_M_DDRX")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L
            complex(c_double_complex) :: nlm(*), M(*)
            real(c_double) :: tau(9)
        end function
        integer(c_int) function c_M_DDRX_src(c, L, tau, M) bind(C, name="sfg_M_DDRX_src")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L
            real(c_double) :: tau(9)
            complex(c_double_complex) :: M(*)
        end function
        integer(c_int) function c_Eij_batch(c, L, P, x, frame, frames, Eg, alpha, ng, Eij, lami, ei, st, stream) &
                bind(C, name="sfg_Eij_tranisotropic_batch")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: c, x, frames, Eij, lami, ei, st, stream
            integer(c_int), value :: L, frame, ng
            integer(c_int64_t), value :: P
            real(c_double) :: Eg(2)
            real(c_double), value :: alpha
        end function
        integer(c_int) function c_Vi_el_batch(c, L, P, x, alpha, Lame, rho, ndir, th, ph, Vi, st, stream) &
                bind(C, name="sfg_Vi_elastic_tranisotropic_batch")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: c, x, Vi, st, stream
            integer(c_int), value :: L, ndir
            integer(c_int64_t), value :: P
            real(c_double), value :: alpha, rho
            real(c_double) :: Lame(5), th(*), ph(*)
        end function
        integer(c_int) function c_M_REG(c, L, D, M) bind(C, name="sfg_M_REG")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int), value :: L
            real(c_double) :: D(9), M(*)
        end function
        integer(c_int) function c_M_CDRX(c, L, M) bind(C, name="sfg_M_CDRX")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int), value :: L
            real(c_double) :: M(*)
        end function
        integer(c_int) function c_a2(c, L, nlm, a2) bind(C, name="sfg_a2")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L
            complex(c_double_complex) :: nlm(*)
            real(c_double) :: a2(9)
        end function
        integer(c_int) function c_eig(c, L, nlm, ei, lami) bind(C, name="sfg_eig")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L
            complex(c_double_complex) :: nlm(*)
            real(c_double) :: ei(9), lami(3)
        end function
        integer(c_int) function c_Eij(c, L, nlm, e1, e2, e3, Eg, alpha, ng, Eij) bind(C, name="sfg_Eij_tranisotropic")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L, ng
            complex(c_double_complex) :: nlm(*)
            real(c_double) :: e1(3), e2(3), e3(3), Eg(2), Eij(6)
            real(c_double), value :: alpha
        end function
        integer(c_int) function c_a4(c, L, nlm, a4) bind(C, name="sfg_a4")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L
            complex(c_double_complex) :: nlm(*)
            real(c_double) :: a4(81)
        end function
        integer(c_int) function c_Vi_el(c, L, nlm, alpha, Lame, rho, ndir, th, ph, Vi) &
                bind(C, name="sfg_Vi_elastic_tranisotropic")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L, ndir
            complex(c_double_complex) :: nlm(*)
            real(c_double), value :: alpha, rho
            real(c_double) :: Lame(5), th(*), ph(*), Vi(*)
        end function
        integer(c_int) function c_Vi_em(c, L, nlm, eps_m, eps_t, mu, ndir, th, ph, Vi) &
                bind(C, name="sfg_Vi_electromagnetic_tranisotropic")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L, ndir
            complex(c_double_complex) :: nlm(*)
            real(c_double), value :: eps_m, eps_t, mu
            real(c_double) :: th(*), ph(*), Vi(*)
        end function
        integer(c_int) function c_a2_to_nlm(c, a2, nlm) bind(C, name="sfg_a2_to_nlm")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            real(c_double) :: a2(9)
            complex(c_double_complex) :: nlm(6)
        end function
        integer(c_int) function c_a4_to_nlm(c, a4, nlm) bind(C, name="sfg_a4_to_nlm")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            real(c_double) :: a4(81)
            complex(c_double_complex) :: nlm(15)
        end function
        integer(c_int) function c_Gamma0(c, eps, T, A, Q, G) bind(C, name="sfg_Gamma0")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double) :: eps(9), G
            real(c_double), value :: T, A, Q
        end function
        integer(c_int) function c_apply_bounds(c, L, nlm, out) bind(C, name="sfg_apply_bounds")
            import :: c_int, c_ptr, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L
            complex(c_double_complex) :: nlm(*), out(*)
        end function
        integer(c_int) function c_evolve(c, desc, stream) bind(C, name="sfg_evolve")
            import :: c_int, c_ptr, sfg_evolve_desc
            type(c_ptr), value :: c
            type(sfg_evolve_desc) :: desc
            type(c_ptr), value :: stream
        end function
        integer(c_int) function c_nlm_to_state(c, L, P, nlm, x, stream) bind(C, name="sfg_nlm_to_state")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: c, nlm, x, stream
            integer(c_int), value :: L
            integer(c_int64_t), value :: P
        end function
        integer(c_int) function c_state_to_nlm(c, L, P, x, nlm, stream) bind(C, name="sfg_state_to_nlm")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: c, x, nlm, stream
            integer(c_int), value :: L
            integer(c_int64_t), value :: P
        end function
        ! HIP runtime (device buffers for the batched path)
        integer(c_int) function c_ri_LROT(c, N, Nt, ri0, dt, D, W, iota, ri) bind(C, name="sfg_ri_LROT")
            import :: c_ptr, c_int, c_int64_t, c_double
            type(c_ptr), value :: c
            integer(c_int64_t), value :: N
            integer(c_int), value :: Nt
            real(c_double), intent(in) :: ri0(*), D(*), W(*)
            real(c_double), value :: dt, iota
            real(c_double), intent(out) :: ri(*)
        end function
        integer(c_int) function c_ri_to_nlm(c, L, N, ri, wi, nlm) bind(C, name="sfg_ri_to_nlm")
            import :: c_ptr, c_int, c_int64_t, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L
            integer(c_int64_t), value :: N
            real(c_double), intent(in) :: ri(*), wi(*)
            complex(c_double_complex), intent(out) :: nlm(*)
        end function
        integer(c_int) function c_Eij_ortho(c, L, q1, q2, q3, e1, e2, e3, Eg, alpha, ng, Eij) &
                bind(C, name="sfg_Eij_orthotropic")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L, ng
            complex(c_double_complex) :: q1(*), q2(*), q3(*)
            real(c_double) :: e1(3), e2(3), e3(3), Eg(6), Eij(6)
            real(c_double), value :: alpha
        end function
        integer(c_int) function c_Vi_ortho(c, L, q1, q2, q3, alpha, Lame, rho, ndir, th, ph, Vi) &
                bind(C, name="sfg_Vi_elastic_orthotropic")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L, ndir
            complex(c_double_complex) :: q1(*), q2(*), q3(*)
            real(c_double), value :: alpha, rho
            real(c_double) :: Lame(9), th(*), ph(*), Vi(*)
        end function
        integer(c_int) function c_Eij_ortho_disc(c, ngr, b, nx, v, e1, e2, e3, Eg, alpha, ng, Eij) &
                bind(C, name="sfg_Eij_orthotropic_discrete")
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int64_t), value :: ngr
            integer(c_int), value :: ng
            real(c_double) :: b(*), nx(*), v(*), e1(3), e2(3), e3(3), Eg(6), Eij(6)
            real(c_double), value :: alpha
        end function
        integer(c_int) function c_Vi_ortho_disc(c, ngr, b, nx, v, alpha, Lame, rho, ndir, th, ph, Vi) &
                bind(C, name="sfg_Vi_elastic_orthotropic_discrete")
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int64_t), value :: ngr
            integer(c_int), value :: ndir
            real(c_double) :: b(*), nx(*), v(*), Lame(9), th(*), ph(*), Vi(*)
            real(c_double), value :: alpha, rho
        end function
        integer(c_int) function c_Evw(c, L, nlm, v, w, tau, Eg, alpha, ng, E) bind(C, name="sfg_Evw_tranisotropic")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L, ng
            complex(c_double_complex) :: nlm(*)
            real(c_double) :: v(3), w(3), tau(9), Eg(2), E
            real(c_double), value :: alpha
        end function
        integer(c_int) function c_Evw_ortho(c, L, q1, q2, q3, v, w, tau, Eg, alpha, ng, E) &
                bind(C, name="sfg_Evw_orthotropic")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L, ng
            complex(c_double_complex) :: q1(*), q2(*), q3(*)
            real(c_double) :: v(3), w(3), tau(9), Eg(6), E
            real(c_double), value :: alpha
        end function
        integer(c_int) function c_eig3(c, M, ei, lami) bind(C, name="sfg_eig3")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double) :: M(9), ei(9), lami(3)
        end function
        integer(c_int) function c_Qnorm(c, L, nlm, alpha, Lame, Q) bind(C, name="sfg_Qnorm_tranisotropic")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L
            complex(c_double_complex) :: nlm(*)
            real(c_double), value :: alpha
            real(c_double) :: Lame(5), Q(9)
        end function
        integer(c_int) function c_moments68(c, L, nlm, m6, m8) bind(C, name="sfg_moments68")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: c
            integer(c_int), value :: L
            complex(c_double_complex) :: nlm(*)
            real(c_double) :: m6(28), m8(45)
        end function
        integer(c_int) function c_Cij_to_Lame_ortho(Cij, Lame) bind(C, name="sfg_Cij_to_Lame_orthotropic")
            import :: c_int, c_double
            real(c_double) :: Cij(9), Lame(9)
        end function
        integer(c_int) function hipMalloc(p, bytes) bind(C, name="hipMalloc")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr) :: p
            integer(c_size_t), value :: bytes
        end function
        integer(c_int) function hipFree(p) bind(C, name="hipFree")
            import :: c_int, c_ptr
            type(c_ptr), value :: p
        end function
        integer(c_int) function hipMemcpy(dst, src, bytes, kind) bind(C, name="hipMemcpy")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
            integer(c_int), value :: kind
        end function
        integer(c_int) function hipDeviceSynchronize() bind(C, name="hipDeviceSynchronize")
            import :: c_int
        end function
    end interface

contains

    subroutine check(rc, what)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: what
        if (rc /= 0) then
            print *, 'specfab_gpu error in ', what, ': ', rc
            error stop 'specfab_gpu: libspecfab_gpu call failed'
        end if
    end subroutine

    ! initspecfab(L): src/specfab.f90:37-53 (context on HIP device 0)
    subroutine initspecfab(L)
        integer, intent(in) :: L
        if (.not. c_associated(ctx)) call check(sfg_ctx_create(0_c_int, ctx), 'sfg_ctx_create')
        Lcap = L
        nlm_len = sfg_nlm_len(int(L, c_int))
    end subroutine

    ! M_LROT(eps, omg, iota, zeta): src/dynamics.f90:52-97
    function M_LROT(eps, omg, iota, zeta) result(M)
        real(dp), intent(in) :: eps(3,3), omg(3,3), iota, zeta
        complex(dp) :: M(nlm_len, nlm_len), Mc(nlm_len, nlm_len)
        real(c_double) :: e(9), o(9)
        e = reshape(transpose(eps), [9])
        o = reshape(transpose(omg), [9])
        call check(c_M_LROT(ctx, int(Lcap, c_int), e, o, iota, zeta, Mc), 'M_LROT')
        M = transpose(Mc)
    end function

    ! M_DDRX(nlm, tau): src/dynamics.f90:251-275
    function M_DDRX(nlm, tau) result(M)
        complex(dp), intent(in) :: nlm(:)
        real(dp), intent(in) :: tau(3,3)
        complex(dp) :: M(nlm_len, nlm_len), Mc(nlm_len, nlm_len), z(nlm_len)
        real(c_double) :: t(9)
        z = nlm(1:nlm_len)
        t = reshape(transpose(tau), [9])
        call check(c_M_DDRX(ctx, int(Lcap, c_int), z, t, Mc), 'M_DDRX')
        M = transpose(Mc)
    end function

    ! M_DDRX_src(tau): src/dynamics.f90:277-298
    function M_DDRX_src(tau) result(M)
        real(dp), intent(in) :: tau(3,3)
        complex(dp) :: M(nlm_len, nlm_len), Mc(nlm_len, nlm_len)
        real(c_double) :: t(9)
        t = reshape(transpose(tau), [9])
        call check(c_M_DDRX_src(ctx, int(Lcap, c_int), t, Mc), 'M_DDRX_src')
        M = transpose(Mc)
    end function

    ! M_REG(D): src/dynamics.f90:494-518
    function M_REG(D) result(M)
        real(dp), intent(in) :: D(3,3)
        real(dp) :: M(nlm_len, nlm_len), Mc(nlm_len, nlm_len)
        real(c_double) :: t(9)
        t = reshape(transpose(D), [9])
        call check(c_M_REG(ctx, int(Lcap, c_int), t, Mc), 'M_REG')
        M = transpose(Mc)
    end function

    ! M_CDRX(): src/dynamics.f90:474-492
    function M_CDRX() result(M)
        real(dp) :: M(nlm_len, nlm_len), Mc(nlm_len, nlm_len)
        call check(c_M_CDRX(ctx, int(Lcap, c_int), Mc), 'M_CDRX')
        M = transpose(Mc)
    end function

    ! a2(nlm): src/moments.f90:37-44
    function a2(nlm) result(A)
        complex(dp), intent(in) :: nlm(:)
        real(dp) :: A(3,3)
        real(c_double) :: v(9)
        complex(dp) :: z(nlm_len)
        z = nlm(1:nlm_len)
        call check(c_a2(ctx, int(Lcap, c_int), z, v), 'a2')
        A = transpose(reshape(v, [3,3]))
    end function

    ! eig(nlm, ei, lami): src/frames.f90:14-22 (ei(i,:) = i-th eigenvector)
    subroutine eig(nlm, ei, lami)
        complex(dp), intent(in) :: nlm(:)
        real(dp), intent(out) :: ei(3,3), lami(3)
        real(c_double) :: v(9)
        complex(dp) :: z(nlm_len)
        z = nlm(1:nlm_len)
        call check(c_eig(ctx, int(Lcap, c_int), z, v, lami), 'eig')
        ei = transpose(reshape(v, [3,3]))
    end subroutine

    ! Eij_tranisotropic(nlm, e1,e2,e3, Eij_grain, alpha, n_grain): src/enhancementfactors.f90:23-45
    function Eij_tranisotropic(nlm, e1, e2, e3, Eij_grain, alpha, n_grain) result(Eij)
        complex(dp), intent(in) :: nlm(:)
        real(dp), intent(in) :: e1(3), e2(3), e3(3), Eij_grain(2), alpha
        integer, intent(in) :: n_grain
        real(dp) :: Eij(6), a1(3), a2_(3), a3(3), Eg(2)
        complex(dp) :: z(nlm_len)
        z = nlm(1:nlm_len)
        a1 = e1; a2_ = e2; a3 = e3; Eg = Eij_grain
        call check(c_Eij(ctx, int(Lcap, c_int), z, a1, a2_, a3, Eg, alpha, int(n_grain, c_int), Eij), 'Eij')
    end function

    ! a4(nlm): src/moments.f90:46-55 (the full tensor, REAL(4) body literals)
    function a4(nlm) result(A)
        complex(dp), intent(in) :: nlm(:)
        real(dp) :: A(3,3,3,3)
        real(c_double) :: v(81)
        complex(dp) :: z(nlm_len)
        z = nlm(1:nlm_len)
        call check(c_a4(ctx, int(Lcap, c_int), z, v), 'a4')
        A = reshape(v, [3,3,3,3], order=[4,3,2,1])   ! C row-major -> Fortran
    end function

    ! Vi_elastic_tranisotropic(nlm, alpha, Lame, rho, theta, phi): src/wavepropagation.f90:121-144
    function Vi_elastic_tranisotropic(nlm, alpha, Lame, rho, theta, phi) result(Vi)
        complex(dp), intent(in) :: nlm(:)
        real(dp), intent(in) :: alpha, Lame(5), rho, theta(:), phi(:)
        real(dp) :: Vi(3, size(theta)), lm(5), th(size(theta)), ph(size(theta))
        complex(dp) :: z(nlm_len)
        z = nlm(1:nlm_len); lm = Lame; th = theta; ph = phi
        call check(c_Vi_el(ctx, int(Lcap, c_int), z, alpha, lm, rho, int(size(theta), c_int), th, ph, Vi), &
                   'Vi_elastic_tranisotropic')
    end function

    ! Eij_orthotropic(nlm_r1, nlm_r2, nlm_r3, e1,e2,e3, Eij_grain, alpha, n_grain):
    ! src/enhancementfactors.f90:134-188 (nlm_r3 = 0: v = b x n)
    function Eij_orthotropic(nlm_r1, nlm_r2, nlm_r3, e1, e2, e3, Eij_grain, alpha, n_grain) result(Eij)
        complex(dp), intent(in) :: nlm_r1(:), nlm_r2(:), nlm_r3(:)
        real(dp), intent(in) :: e1(3), e2(3), e3(3), Eij_grain(6), alpha
        integer, intent(in) :: n_grain
        real(dp) :: Eij(6), a1(3), a2_(3), a3(3), Eg(6)
        complex(dp) :: z1(nlm_len), z2(nlm_len), z3(nlm_len)
        z1 = nlm_r1(1:nlm_len); z2 = nlm_r2(1:nlm_len); z3 = nlm_r3(1:nlm_len)
        a1 = e1; a2_ = e2; a3 = e3; Eg = Eij_grain
        call check(c_Eij_ortho(ctx, int(Lcap, c_int), z1, z2, z3, a1, a2_, a3, Eg, alpha, int(n_grain, c_int), Eij), &
                   'Eij_orthotropic')
    end function

    ! Vi_elastic_orthotropic(qlm_1, qlm_2, qlm_3, alpha, Lame_grain, rho, theta, phi):
    ! src/wavepropagation.f90:41-68 (Voigt), Vi(3, ndir) = (qS1, qS2, qP)
    function Vi_elastic_orthotropic(qlm_1, qlm_2, qlm_3, alpha, Lame_grain, rho, theta, phi) result(Vi)
        complex(dp), intent(in) :: qlm_1(:), qlm_2(:), qlm_3(:)
        real(dp), intent(in) :: alpha, Lame_grain(9), rho, theta(:), phi(:)
        real(dp) :: Vi(3, size(theta)), lm(9), th(size(theta)), ph(size(theta))
        complex(dp) :: z1(nlm_len), z2(nlm_len), z3(nlm_len)
        z1 = qlm_1(1:nlm_len); z2 = qlm_2(1:nlm_len); z3 = qlm_3(1:nlm_len)
        lm = Lame_grain; th = theta; ph = phi
        call check(c_Vi_ortho(ctx, int(Lcap, c_int), z1, z2, z3, alpha, lm, rho, int(size(theta), c_int), th, ph, Vi), &
                   'Vi_elastic_orthotropic')
    end function

    ! Eij_orthotropic_discrete(bi, ni, vi, e1,e2,e3, Eij_grain, alpha, n_grain), bi(N,3):
    ! src/enhancementfactors.f90:191-227
    function Eij_orthotropic_discrete(bi, ni, vi, e1, e2, e3, Eij_grain, alpha, n_grain) result(Eij)
        real(dp), intent(in) :: bi(:,:), ni(:,:), vi(:,:), e1(3), e2(3), e3(3), Eij_grain(6), alpha
        integer, intent(in) :: n_grain
        real(dp) :: Eij(6), a1(3), a2_(3), a3(3), Eg(6)
        real(c_double) :: b(3*size(bi,1)), nx(3*size(bi,1)), v(3*size(bi,1))
        b = reshape(transpose(bi), [3*size(bi,1)])   ! (N,3) column-major -> row-major
        nx = reshape(transpose(ni), [3*size(bi,1)])
        v = reshape(transpose(vi), [3*size(bi,1)])
        a1 = e1; a2_ = e2; a3 = e3; Eg = Eij_grain
        call check(c_Eij_ortho_disc(ctx, int(size(bi,1), c_int64_t), b, nx, v, a1, a2_, a3, Eg, alpha, &
                                    int(n_grain, c_int), Eij), 'Eij_orthotropic_discrete')
    end function

    ! Vi_elastic_orthotropic__discrete(mi, alpha, Lame_grain, rho, theta, phi), mi(3,3,N)
    ! = (axis, xyz, grain): src/wavepropagation.f90:70-96
    function Vi_elastic_orthotropic__discrete(mi, alpha, Lame_grain, rho, theta, phi) result(Vi)
        real(dp), intent(in) :: mi(:,:,:), alpha, Lame_grain(9), rho, theta(:), phi(:)
        real(dp) :: Vi(3, size(theta)), lm(9), th(size(theta)), ph(size(theta))
        real(c_double) :: b(3*size(mi,3)), nx(3*size(mi,3)), v(3*size(mi,3))
        b = reshape(mi(1,:,:), [3*size(mi,3)])        ! grain g's xyz at 3(g-1)+1..3
        nx = reshape(mi(2,:,:), [3*size(mi,3)])
        v = reshape(mi(3,:,:), [3*size(mi,3)])
        lm = Lame_grain; th = theta; ph = phi
        call check(c_Vi_ortho_disc(ctx, int(size(mi,3), c_int64_t), b, nx, v, alpha, lm, rho, int(size(theta), c_int), &
                                   th, ph, Vi), 'Vi_elastic_orthotropic__discrete')
    end function

    ! Evw_tranisotropic(v, w, tau, nlm, Eij_grain, alpha, n_grain): src/enhancementfactors.f90:47-69
    ! (the module function's argument order; tau symmetric)
    function Evw_tranisotropic(v, w, tau, nlm, Eij_grain, alpha, n_grain) result(Evw)
        real(dp), intent(in) :: v(3), w(3), tau(3,3), Eij_grain(2), alpha
        complex(dp), intent(in) :: nlm(:)
        integer, intent(in) :: n_grain
        real(dp) :: Evw, a(3), b(3), t(9), Eg(2)
        complex(dp) :: z(nlm_len)
        z = nlm(1:nlm_len); a = v; b = w; Eg = Eij_grain
        t = reshape(transpose(tau), [9])
        call check(c_Evw(ctx, int(Lcap, c_int), z, a, b, t, Eg, alpha, int(n_grain, c_int), Evw), 'Evw_tranisotropic')
    end function

    ! Evw_orthotropic(v, w, tau, nlm_r1, nlm_r2, nlm_r3, Eij_grain, alpha, n_grain):
    ! src/enhancementfactors.f90:134-160
    function Evw_orthotropic(v, w, tau, nlm_r1, nlm_r2, nlm_r3, Eij_grain, alpha, n_grain) result(Evw)
        real(dp), intent(in) :: v(3), w(3), tau(3,3), Eij_grain(6), alpha
        complex(dp), intent(in) :: nlm_r1(:), nlm_r2(:), nlm_r3(:)
        integer, intent(in) :: n_grain
        real(dp) :: Evw, a(3), b(3), t(9), Eg(6)
        complex(dp) :: z1(nlm_len), z2(nlm_len), z3(nlm_len)
        z1 = nlm_r1(1:nlm_len); z2 = nlm_r2(1:nlm_len); z3 = nlm_r3(1:nlm_len)
        a = v; b = w; Eg = Eij_grain
        t = reshape(transpose(tau), [9])
        call check(c_Evw_ortho(ctx, int(Lcap, c_int), z1, z2, z3, a, b, t, Eg, alpha, int(n_grain, c_int), Evw), &
                   'Evw_orthotropic')
    end function

    ! eig3(M, e1, e2, e3, lami): src/frames.f90:62-80 (largest eigenvalue first)
    subroutine eig3(M, e1, e2, e3, lami)
        real(dp), intent(in) :: M(3,3)
        real(dp), intent(out) :: e1(3), e2(3), e3(3), lami(3)
        real(c_double) :: a(9), v(9)
        a = reshape(transpose(M), [9])
        call check(c_eig3(ctx, a, v, lami), 'eig3')
        e1 = v(1:3); e2 = v(4:6); e3 = v(7:9)
    end subroutine

    ! Qnorm_tranisotropic(nlm, alpha, Lame_grain): src/wavepropagation.f90:146-211
    function Qnorm_tranisotropic(nlm, alpha, Lame_grain) result(Q)
        complex(dp), intent(in) :: nlm(:)
        real(dp), intent(in) :: alpha, Lame_grain(5)
        real(dp) :: Q(3,3), lm(5)
        real(c_double) :: v(9)
        complex(dp) :: z(nlm_len)
        z = nlm(1:nlm_len); lm = Lame_grain
        call check(c_Qnorm(ctx, int(Lcap, c_int), z, alpha, lm, v), 'Qnorm_tranisotropic')
        Q = transpose(reshape(v, [3,3]))
    end function

    ! a6(nlm): src/moments.f90:57-66 (exact sixth moments; L >= 6)
    function a6(nlm) result(A)
        complex(dp), intent(in) :: nlm(:)
        real(dp) :: A(3,3,3,3,3,3)
        real(c_double) :: m6(28), m8(45)
        complex(dp) :: z(nlm_len)
        integer :: i1, i2, i3, i4, i5, i6, nx, ny, k, ia
        z = nlm(1:nlm_len)
        call check(c_moments68(ctx, int(Lcap, c_int), z, m6, m8), 'a6')
        ! monomial <x^a y^b z^c> at position sum_{a' > a} (7 - a') + (6 - a - b) (order a = 6..0, b = 6 - a..0)
        do i1 = 1, 3; do i2 = 1, 3; do i3 = 1, 3; do i4 = 1, 3; do i5 = 1, 3; do i6 = 1, 3
            nx = count([i1, i2, i3, i4, i5, i6] == 1)
            ny = count([i1, i2, i3, i4, i5, i6] == 2)
            k = 0
            do ia = nx + 1, 6
                k = k + (7 - ia)
            end do
            A(i1, i2, i3, i4, i5, i6) = m6(k + (6 - nx - ny) + 1)
        end do; end do; end do; end do; end do; end do
    end function

    ! Cij_to_Lame_orthotropic(Cij): src/elasticities.f90:122-139
    function Cij_to_Lame_orthotropic(Cij) result(Lame)
        real(dp), intent(in) :: Cij(9)
        real(dp) :: Lame(9), c(9)
        c = Cij
        call check(c_Cij_to_Lame_ortho(c, Lame), 'Cij_to_Lame_orthotropic')
    end function

    ! Vi_electromagnetic_tranisotropic(nlm, eps_m, eps_t, mu, theta, phi): src/wavepropagation.f90:217-261
    function Vi_electromagnetic_tranisotropic(nlm, eps_m, eps_t, mu, theta, phi) result(Vi)
        complex(dp), intent(in) :: nlm(:)
        real(dp), intent(in) :: eps_m, eps_t, mu, theta(:), phi(:)
        real(dp) :: Vi(2, size(theta)), th(size(theta)), ph(size(theta))
        complex(dp) :: z(nlm_len)
        z = nlm(1:nlm_len); th = theta; ph = phi
        call check(c_Vi_em(ctx, int(Lcap, c_int), z, eps_m, eps_t, mu, int(size(theta), c_int), th, ph, Vi), &
                   'Vi_electromagnetic_tranisotropic')
    end function

    ! a2_to_nlm(a2), a4_to_nlm(a4): src/moments.f90:68-84
    function a2_to_nlm(a2in) result(nlm)
        real(dp), intent(in) :: a2in(3,3)
        complex(dp) :: nlm(6)
        real(c_double) :: t(9)
        t = reshape(transpose(a2in), [9])
        call check(c_a2_to_nlm(ctx, t, nlm), 'a2_to_nlm')
    end function

    function a4_to_nlm(a4in) result(nlm)
        real(dp), intent(in) :: a4in(3,3,3,3)
        complex(dp) :: nlm(15)
        real(c_double) :: t(81)
        t = reshape(reshape(a4in, [3,3,3,3], order=[4,3,2,1]), [81])   ! Fortran -> C row-major
        call check(c_a4_to_nlm(ctx, t, nlm), 'a4_to_nlm')
    end function

    ! ri_LROT(ri0, dt, Nt, D, W, iota): src/dynamics.f90:125-137 -- ri0(N,3),
    ! D(Nt,3,3) and W(Nt,3,3) as the reference takes them, ri(Nt,N,3)
    function ri_LROT(ri0, dt, Nt, D, W, iota) result(ri)
        real(dp), intent(in) :: ri0(:,:), dt, D(:,:,:), W(:,:,:), iota
        integer, intent(in) :: Nt
        real(dp) :: ri(Nt, size(ri0, 1), 3)
        real(c_double), allocatable :: r0(:,:), Dc(:,:,:), Wc(:,:,:), out(:,:,:)
        integer :: N, t, i, j
        N = size(ri0, 1)
        allocate(r0(3, N), Dc(3, 3, Nt), Wc(3, 3, Nt), out(3, N, Nt))
        r0 = transpose(ri0)                       ! C row-major (N, 3)
        Dc = 0; Wc = 0
        do t = 1, min(Nt, size(D, 1)); do i = 1, 3; do j = 1, 3
            Dc(j, i, t) = D(t, i, j)               ! C row-major (Nt, 3, 3)
            Wc(j, i, t) = W(t, i, j)
        end do; end do; end do
        call check(c_ri_LROT(ctx, int(N, c_int64_t), int(Nt, c_int), r0, dt, Dc, Wc, iota, out), 'ri_LROT')
        do t = 1, Nt; do i = 1, N; do j = 1, 3
            ri(t, i, j) = out(j, i, t)
        end do; end do; end do
    end function

    ! ri_to_nlm(ri, wi, L): src/moments.f90:98-131 -- ri(N,3), wi(N)
    function ri_to_nlm(ri, wi, L) result(nlm)
        real(dp), intent(in) :: ri(:,:), wi(:)
        integer, intent(in) :: L
        complex(dp) :: nlm((L+1)*(L+2)/2)
        real(c_double), allocatable :: r(:,:), w(:)
        allocate(r(3, size(ri, 1)), w(size(wi)))
        r = transpose(ri)
        w = wi
        call check(c_ri_to_nlm(ctx, int(L, c_int), int(size(ri, 1), c_int64_t), r, w, nlm), 'ri_to_nlm')
    end function

    ! Gamma0(eps, T, A, Q): src/dynamics.f90:451-472
    function Gamma0(eps, T, A, Q) result(G)
        real(dp), intent(in) :: eps(3,3), T, A, Q
        real(dp) :: G
        real(c_double) :: e(9)
        e = reshape(transpose(eps), [9])
        call check(c_Gamma0(ctx, e, T, A, Q, G), 'Gamma0')
    end function

    ! apply_bounds(nlm): src/dynamics.f90:530-557
    function apply_bounds(nlm) result(nlm_bounded)
        complex(dp), intent(in) :: nlm(:)
        complex(dp) :: nlm_bounded(nlm_len), z(nlm_len)
        z = nlm(1:nlm_len)
        call check(c_apply_bounds(ctx, int(Lcap, c_int), z, nlm_bounded), 'apply_bounds')
    end function

    ! Device state (n, P) of host coefficient vectors nlm(:, p), p = 1..P
    function states_on_device(nlm) result(dx)
        complex(dp), intent(in), target :: nlm(:,:)
        type(c_ptr) :: dx, dnlm
        integer(c_int64_t) :: P
        integer :: n
        n = size(nlm, 1)
        P = size(nlm, 2)
        call check(hipMalloc(dx, int(n*P*8, c_size_t)), 'hipMalloc')
        call check(hipMalloc(dnlm, int(2*n*P*8, c_size_t)), 'hipMalloc')
        call check(hipMemcpy(dnlm, c_loc(nlm), int(2*n*P*8, c_size_t), 1_c_int), 'hipMemcpy')
        call check(c_nlm_to_state(ctx, int(Lcap, c_int), P, dnlm, dx, c_null_ptr), 'nlm_to_state')
        call check(hipDeviceSynchronize(), 'hipDeviceSynchronize')
        call check(hipFree(dnlm), 'hipFree')
    end function

    ! Eij_tranisotropic of many parcels (the loop of specfabpy's
    ! Eij_tranisotropic_arr, src/specfabpy.f90:474-486) in one batched call:
    ! Eij(:, p) of nlm(:, p) in the a2 eigenframe of each parcel (frame = 0,
    ! as eig gives it) or the Cartesian frame (frame = 2)
    function Eij_tranisotropic_batch(nlm, Eij_grain, alpha, n_grain, frame) result(Eij)
        complex(dp), intent(in) :: nlm(:,:)
        real(dp), intent(in) :: Eij_grain(2), alpha
        integer, intent(in) :: n_grain, frame
        real(dp) :: Eij(6, size(nlm, 2))
        real(dp), allocatable, target :: buf(:,:)
        real(c_double) :: Eg(2)
        type(c_ptr) :: dx, dE
        integer(c_int64_t) :: P
        P = size(nlm, 2)
        Eg = Eij_grain
        dx = states_on_device(nlm)
        call check(hipMalloc(dE, int(6*P*8, c_size_t)), 'hipMalloc')
        call check(c_Eij_batch(ctx, int(Lcap, c_int), P, dx, int(frame, c_int), c_null_ptr, Eg, alpha, &
                               int(n_grain, c_int), dE, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), &
                   'Eij_tranisotropic_batch')
        call check(hipDeviceSynchronize(), 'hipDeviceSynchronize')
        allocate(buf(6, P))
        call check(hipMemcpy(c_loc(buf), dE, int(6*P*8, c_size_t), 2_c_int), 'hipMemcpy')
        Eij = buf
        call check(hipFree(dE), 'hipFree'); call check(hipFree(dx), 'hipFree')
    end function

    ! Vi_elastic_tranisotropic of many parcels: Vi(:, d, p) = (qS1, qS2, qP)
    ! of nlm(:, p) along (theta(d), phi(d))
    function Vi_elastic_tranisotropic_batch(nlm, alpha, Lame, rho, theta, phi) result(Vi)
        complex(dp), intent(in) :: nlm(:,:)
        real(dp), intent(in) :: alpha, Lame(5), rho, theta(:), phi(:)
        real(dp) :: Vi(3, size(theta), size(nlm, 2))
        real(dp), allocatable, target :: buf(:,:,:)
        real(dp) :: lm(5), th(size(theta)), ph(size(theta))
        type(c_ptr) :: dx, dV
        integer(c_int64_t) :: P
        integer(c_size_t) :: nb
        P = size(nlm, 2)
        lm = Lame; th = theta; ph = phi
        nb = int(3*size(theta)*P*8, c_size_t)
        dx = states_on_device(nlm)
        call check(hipMalloc(dV, nb), 'hipMalloc')
        call check(c_Vi_el_batch(ctx, int(Lcap, c_int), P, dx, alpha, lm, rho, int(size(theta), c_int), th, ph, &
                                 dV, c_null_ptr, c_null_ptr), 'Vi_elastic_tranisotropic_batch')
        call check(hipDeviceSynchronize(), 'hipDeviceSynchronize')
        allocate(buf(3, size(theta), P))
        call check(hipMemcpy(c_loc(buf), dV, nb, 2_c_int), 'hipMemcpy')
        Vi = buf
        call check(hipFree(dV), 'hipFree'); call check(hipFree(dx), 'hipFree')
    end function

    ! Batched evolution of P parcels on the GPU (host arrays in, host arrays out):
    ! nlm(nlm_len, P) advanced in place by nsteps steps of size dt; ugrad(3,3,P)
    ! velocity gradients, gam0(P) DDRX and lam(P) CDRX rates; procs a bitmask
    ! of SFG_LROT/SFG_DDRX/SFG_CDRX/SFG_REG, scheme SFG_EULER or SFG_RK4.
    subroutine sfg_evolve_parcels(nlm, ugrad, gam0, lam, nsteps, dt, procs, scheme, iota, zeta)
        complex(dp), intent(inout), target :: nlm(:,:)
        real(dp), intent(in) :: ugrad(:,:,:), dt, iota, zeta
        real(dp), intent(in), target :: gam0(:), lam(:)
        integer, intent(in) :: nsteps, procs, scheme
        integer(c_int64_t) :: P
        real(c_double), allocatable, target :: G(:), nlmT(:,:)
        type(c_ptr) :: dG, dg0, dlam, dx, dnlm, dst
        type(sfg_evolve_desc) :: d
        integer :: p_, i, j, n
        integer(c_size_t) :: nb
        n = size(nlm, 1)
        P = size(nlm, 2)
        allocate(G(9*P))
        do p_ = 1, int(P)
            do i = 1, 3
                do j = 1, 3
                    G(9*(p_-1) + 3*(i-1) + j) = ugrad(i, j, p_)   ! row-major per parcel
                end do
            end do
        end do
        call check(hipMalloc(dG, int(9*P*8, c_size_t)), 'hipMalloc')
        call check(hipMalloc(dg0, int(P*8, c_size_t)), 'hipMalloc')
        call check(hipMalloc(dlam, int(P*8, c_size_t)), 'hipMalloc')
        call check(hipMalloc(dx, int(n*P*8, c_size_t)), 'hipMalloc')
        call check(hipMalloc(dnlm, int(2*n*P*8, c_size_t)), 'hipMalloc')
        call check(hipMalloc(dst, int(P*4, c_size_t)), 'hipMalloc')
        call check(hipMemcpy(dG, c_loc(G), int(9*P*8, c_size_t), 1_c_int), 'hipMemcpy')
        call check(hipMemcpy(dg0, c_loc(gam0), int(P*8, c_size_t), 1_c_int), 'hipMemcpy')
        call check(hipMemcpy(dlam, c_loc(lam), int(P*8, c_size_t), 1_c_int), 'hipMemcpy')
        ! nlm(:, p) is parcel p's coefficient vector: the (P, n) parcel-major C layout
        nb = int(2*n*P*8, c_size_t)
        call check(hipMemcpy(dnlm, c_loc(nlm), nb, 1_c_int), 'hipMemcpy')
        call check(c_nlm_to_state(ctx, int(Lcap, c_int), P, dnlm, dx, c_null_ptr), 'nlm_to_state')
        d%L = Lcap; d%P = P; d%procs = procs; d%scheme = scheme; d%precision = SFG_FP64
        d%nsteps = nsteps; d%dt = dt; d%iota = iota; d%zeta = zeta
        d%ugrad_dev = dG; d%tau_dev = c_null_ptr; d%gam0_dev = dg0; d%lam_dev = dlam
        d%steps_dev = c_null_ptr; d%x_dev = dx; d%status_dev = c_null_ptr
        d%reg_nu = 0; d%reg_expo = 0
        call check(c_evolve(ctx, d, c_null_ptr), 'sfg_evolve')
        call check(c_state_to_nlm(ctx, int(Lcap, c_int), P, dx, dnlm, c_null_ptr), 'state_to_nlm')
        call check(hipDeviceSynchronize(), 'hipDeviceSynchronize')
        call check(hipMemcpy(c_loc(nlm), dnlm, nb, 2_c_int), 'hipMemcpy')
        call check(hipFree(dG), 'hipFree'); call check(hipFree(dg0), 'hipFree'); call check(hipFree(dlam), 'hipFree')
        call check(hipFree(dx), 'hipFree'); call check(hipFree(dnlm), 'hipFree'); call check(hipFree(dst), 'hipFree')
    end subroutine

end module specfab_gpu
