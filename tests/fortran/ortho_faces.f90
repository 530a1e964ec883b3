! Fortran host program for the orthotropic-grain faces of the specfab_gpu
! module: reads slip-system-axis distributions, grain axes and parameters
! from the text file named on the command line, calls Eij_orthotropic,
! Vi_elastic_orthotropic, their discrete-grain forms and
! Cij_to_Lame_orthotropic with the reference's argument shapes, and prints
! tagged results (tests/test_gpu_fortran.py compares them with the reference
! fixtures).
program ortho_faces
    use specfab_gpu
    implicit none
    integer, parameter :: dp = 8
    integer :: K, nd, N, i, j, g
    real(dp) :: Eg(6), Lame(9), rho, row(30), Cij(9), e(3,3), Eij(6)
    real(dp), allocatable :: th(:), ph(:), bi(:,:), ni(:,:), vi(:,:), mi(:,:,:), V(:,:)
    complex(dp) :: q(15, 3)
    character(len=512) :: path

    call get_command_argument(1, path)
    open(10, file=trim(path), status='old', action='read')
    call initspecfab(4)
    read(10, *) K, nd
    allocate(th(nd), ph(nd), V(3, nd))
    read(10, *) Eg
    read(10, *) Lame, rho
    read(10, *) th
    read(10, *) ph
    e = 0
    do i = 1, 3
        e(i, i) = 1
    end do
    do i = 1, K
        do j = 1, 3
            read(10, *) row
            q(:, j) = cmplx(row(1:29:2), row(2:30:2), kind=dp)
        end do
        Eij = Eij_orthotropic(q(:,1), q(:,2), q(:,3), e(:,1), e(:,2), e(:,3), Eg, 0.0d0, 1)
        print '(a,i4,6es25.16)', 'E ', i, Eij
        V = Vi_elastic_orthotropic(q(:,1), q(:,2), q(:,3), 1.0d0, Lame, rho, th, ph)
        do j = 1, nd
            print '(a,i4,i4,3es25.16)', 'V ', i, j, V(:, j)
        end do
    end do
    read(10, *) N
    allocate(bi(N, 3), ni(N, 3), vi(N, 3), mi(3, 3, N))
    do g = 1, N
        read(10, *) row(1:9)
        bi(g, :) = row(1:3)
        ni(g, :) = row(4:6)
        vi(g, :) = row(7:9)
        mi(1, :, g) = row(1:3)
        mi(2, :, g) = row(4:6)
        mi(3, :, g) = row(7:9)
    end do
    Eij = Eij_orthotropic_discrete(bi, ni, vi, e(:,1), e(:,2), e(:,3), Eg, 0.0d0, 1)
    print '(a,6es25.16)', 'ED ', Eij
    V = Vi_elastic_orthotropic__discrete(mi, 1.0d0, Lame, rho, th, ph)
    do j = 1, nd
        print '(a,i4,3es25.16)', 'VD ', j, V(:, j)
    end do
    read(10, *) Cij
    print '(a,9es25.16)', 'LAME ', Cij_to_Lame_orthotropic(Cij)
end program
