! Fortran host code driving libspecfab_gpu through the specfab_gpu module
! (ISO_C_BINDING), in the pattern of the reference's Fortran demo
! (demo/fabric-evolution/run-fabric-evolution-ddrx.f90:91-137):
!   M built from the module operators, nlm = nlm + dt*matmul(M, nlm).
! Runs the one-parcel L=8 uniaxial-compression case (LROT + REG, Euler dt=1,
! 1000 steps) twice -- with the single-parcel operators on the host loop and
! with the batched GPU stepper -- and prints the final state and a2.
program c1_parcel
    use specfab_gpu
    implicit none
    integer, parameter :: dp = 8
    complex(dp), allocatable :: nlm(:), M(:,:), nlmb(:,:)
    real(dp) :: D(3,3), W(3,3), ug(3,3,1), g0(1), lam(1), A(3,3)
    integer :: t, i
    call initspecfab(8)
    allocate(nlm(nlm_len), M(nlm_len, nlm_len), nlmb(nlm_len, 1))
    D = 0.0d0
    D(1,1) = 0.5d-3; D(2,2) = 0.5d-3; D(3,3) = -1.0d-3
    W = 0.0d0
    nlm = (0.0d0, 0.0d0)
    nlm(1) = 1/sqrt(4*3.141592653589793d0)
    nlmb(:,1) = nlm
    M = M_LROT(D, W, 1.0d0, 0.0d0) + M_REG(D)
    do t = 1, 1000
        nlm = nlm + 1.0d0*matmul(M, nlm)
    end do
    ug(:,:,1) = D
    g0 = 0.0d0; lam = 0.0d0
    call sfg_evolve_parcels(nlmb, ug, g0, lam, 1000, 1.0d0, SFG_LROT + SFG_REG, SFG_EULER, 1.0d0, 0.0d0)
    A = a2(nlm)
    do i = 1, nlm_len
        print '(A,I3,2ES26.17,2ES26.17)', 'nlm ', i, real(nlm(i)), aimag(nlm(i)), real(nlmb(i,1)), aimag(nlmb(i,1))
    end do
    print '(A,3ES26.17)', 'a2diag ', A(1,1), A(2,2), A(3,3)
end program
