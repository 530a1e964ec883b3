! Fortran host program for the discrete-grain faces of the specfab_gpu
! module: reads grain axes and kinematics from the text file named on the
! command line, calls ri_LROT and ri_to_nlm, and prints tagged results
! (tests/test_gpu_fortran.py compares them with the reference fixture).
program discrete_faces
    use specfab_gpu
    implicit none
    integer, parameter :: dp = 8
    integer :: N, Nt, M, L, t, i, j
    real(dp) :: dt, iota, row(9)
    real(dp), allocatable :: ri0(:,:), D(:,:,:), W(:,:,:), ri(:,:,:), rs(:,:), ws(:)
    complex(dp), allocatable :: nlm(:)
    character(len=512) :: path

    call get_command_argument(1, path)
    open(10, file=trim(path), status='old', action='read')
    call initspecfab(8)
    read(10, *) N, Nt, dt, iota
    allocate(ri0(N, 3), D(Nt, 3, 3), W(Nt, 3, 3))
    do i = 1, N
        read(10, *) ri0(i, :)
    end do
    do t = 1, Nt
        read(10, *) row
        do i = 1, 3; do j = 1, 3
            D(t, i, j) = row(3*(i-1) + j)
        end do; end do
    end do
    do t = 1, Nt
        read(10, *) row
        do i = 1, 3; do j = 1, 3
            W(t, i, j) = row(3*(i-1) + j)
        end do; end do
    end do
    ri = ri_LROT(ri0, dt, Nt, D, W, iota)
    do i = 1, N
        print '(a,i6,3es25.16)', 'ri ', i, ri(Nt, i, :)
    end do
    read(10, *) M, L
    allocate(rs(M, 3), ws(M))
    do i = 1, M
        read(10, *) rs(i, :)
    end do
    read(10, *) ws
    nlm = ri_to_nlm(rs, ws, L)
    do i = 1, size(nlm)
        print '(a,i4,2es25.16)', 'nlm ', i, real(nlm(i)), aimag(nlm(i))
    end do
end program
