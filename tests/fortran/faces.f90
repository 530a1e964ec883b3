! Fortran host program for the Evw, eig3, Qnorm_tranisotropic and a6 faces
! of the specfab_gpu module: reads states and stress states from the text
! file named on the command line and prints tagged results
! (tests/test_gpu_fortran.py compares them with the reference fixture).
program faces
    use specfab_gpu
    implicit none
    integer, parameter :: dp = 8
    integer :: K, S, i, j, nl
    real(dp) :: row(90), v(3), w(3), tau(3,3), M(3,3), e1(3), e2(3), e3(3), lami(3), Lame(5), Qn(3,3), Ego(6)
    real(dp) :: T6(3,3,3,3,3,3)
    complex(dp), allocatable :: x(:), qd(:,:)
    character(len=512) :: path

    call get_command_argument(1, path)
    open(10, file=trim(path), status='old', action='read')
    call initspecfab(8)
    nl = 45
    allocate(x(nl), qd(nl, 3))
    read(10, *) K, S
    read(10, *) Lame
    read(10, *) Ego
    do i = 1, K
        read(10, *) row
        x = cmplx(row(1:89:2), row(2:90:2), kind=dp)
        Qn = Qnorm_tranisotropic(x, 0.5d0, Lame)
        print '(a,i4,9es25.16)', 'Q ', i, Qn(1,:), Qn(2,:), Qn(3,:)
        T6 = a6(x)
        print '(a,i4,4es25.16)', 'A6 ', i, T6(1,1,1,1,1,1), T6(1,1,2,2,3,3), T6(1,2,3,3,3,3), T6(3,3,3,3,3,3)
        do j = 1, S
            read(10, *) row(1:15)
            v = row(1:3); w = row(4:6)
            tau = transpose(reshape(row(7:15), [3,3]))
            print '(a,i4,i4,es25.16)', 'EVW ', i, j, Evw_tranisotropic(v, w, tau, x, [1.0d0, 1.0d3], 0.0125d0, 1)
        end do
    end do
    do j = 1, 3
        read(10, *) row
        qd(:, j) = cmplx(row(1:89:2), row(2:90:2), kind=dp)
    end do
    do j = 1, S
        read(10, *) row(1:15)
        v = row(1:3); w = row(4:6)
        tau = transpose(reshape(row(7:15), [3,3]))
        print '(a,i4,es25.16)', 'EVWO ', j, Evw_orthotropic(v, w, tau, qd(:,1), qd(:,2), qd(:,3), Ego, 0.0d0, 1)
    end do
    read(10, *) row(1:9)
    M = transpose(reshape(row(1:9), [3,3]))
    call eig3(M, e1, e2, e3, lami)
    print '(a,12es25.16)', 'EIG ', lami, e1, e2, e3
end program
