! Fortran host program for the homogenization / initial-state faces of the
! specfab_gpu module: reads inputs from the text file named on the command
! line, calls a4, Vi_elastic_tranisotropic, Vi_electromagnetic_tranisotropic,
! the batched Eij_tranisotropic_batch / Vi_elastic_tranisotropic_batch,
! M_DDRX_src, a2_to_nlm, a4_to_nlm, Gamma0 and apply_bounds, and prints
! tagged results (tests/test_gpu_fortran.py compares them with the
! reference fixtures).
program homog_faces
    use specfab_gpu
    implicit none
    integer, parameter :: dp = 8
    integer :: ns, nd, k, i, j, m, r, l
    complex(dp), allocatable :: nlm(:,:)
    real(dp), allocatable :: th(:), ph(:), Vi(:,:), Ve(:,:), Eb(:,:), Vb(:,:,:)
    complex(dp) :: Md(45,45)
    real(dp) :: tau(3,3), Eg(2)
    real(dp) :: lame(5), rho, em, et, mu, re(45), im(45), a4t(3,3,3,3), a2t(3,3), eps(3,3), T, A, QQ, G
    complex(dp) :: z6(6), z15(15), zb(45)
    character(len=512) :: path

    call get_command_argument(1, path)
    open(10, file=trim(path), status='old', action='read')
    call initspecfab(8)
    read(10, *) ns, nd
    allocate(nlm(45, ns), th(nd), ph(nd), Vi(3, nd), Ve(2, nd))
    do k = 1, ns
        read(10, *) re
        read(10, *) im
        nlm(:, k) = cmplx(re, im, kind=dp)
    end do
    read(10, *) th
    read(10, *) ph
    read(10, *) lame, rho, em, et, mu
    do k = 1, ns
        a4t = a4(nlm(:, k))
        do i = 1, 3; do j = 1, 3; do m = 1, 3; do r = 1, 3
            print '(a,5i3,es25.16)', 'a4 ', k, i, j, m, r, a4t(i, j, m, r)
        end do; end do; end do; end do
        Vi = Vi_elastic_tranisotropic(nlm(:, k), 0.5d0, lame, rho, th, ph)
        Ve = Vi_electromagnetic_tranisotropic(nlm(:, k), em, et, mu, th, ph)
        do j = 1, nd
            print '(a,2i3,3es25.16)', 'vi ', k, j, Vi(:, j)
            print '(a,2i3,2es25.16)', 've ', k, j, Ve(:, j)
        end do
    end do
    ! batched faces: Eij in the eigenframe and the Cartesian frame, Vi (Hill)
    Eg = [1d0, 1d3]
    Eb = Eij_tranisotropic_batch(nlm, Eg, 0.0125d0, 1, 0)
    do k = 1, ns
        print '(a,i3,6es25.16)', 'ebe ', k, Eb(:, k)
    end do
    Eb = Eij_tranisotropic_batch(nlm, Eg, 0.0125d0, 1, 2)
    do k = 1, ns
        print '(a,i3,6es25.16)', 'ebx ', k, Eb(:, k)
    end do
    Vb = Vi_elastic_tranisotropic_batch(nlm, 0.5d0, lame, rho, th, ph)
    do k = 1, ns
        do j = 1, nd
            print '(a,2i3,3es25.16)', 'vb ', k, j, Vb(:, j, k)
        end do
    end do
    ! M_DDRX_src(tau)
    read(10, *) tau
    Md = M_DDRX_src(transpose(tau))
    do i = 1, 45
        do j = 1, 45
            print '(a,2i4,2es25.16)', 'mds ', i, j, Md(i, j)
        end do
    end do
    ! structure tensors -> nlm
    read(10, *) a2t
    z6 = a2_to_nlm(transpose(a2t))
    do i = 1, 6
        print '(a,i3,2es25.16)', 'a2n ', i, z6(i)
    end do
    read(10, *) a4t
    z15 = a4_to_nlm(a4t)
    do i = 1, 15
        print '(a,i3,2es25.16)', 'a4n ', i, z15(i)
    end do
    ! Gamma0
    read(10, *) eps
    read(10, *) T, A, QQ
    G = Gamma0(transpose(eps), T, A, QQ)
    print '(a,es25.16)', 'g0 ', G
    ! apply_bounds
    read(10, *) re
    read(10, *) im
    zb = apply_bounds(cmplx(re, im, kind=dp))
    do l = 1, 45
        print '(a,i3,2es25.16)', 'bnd ', l, zb(l)
    end do
    close(10)
end program
