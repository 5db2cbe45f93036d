! ref_shim.f90 -- C-callable entry points onto the reference's own Fortran
! routines, compiled unchanged from the reference tree by oracle/Makefile into
! oracle/_ref/libfml_ref.so.  Test infrastructure only: it replaces the F2PY
! wrapper layer (which needs a Fortran compiler numpy.distutils does not find
! here), so the interface blocks below restate the reference's dummy-argument
! declarations and nothing else.  Every array is passed in Fortran order, as
! F2PY would pass it.
module fml_ref_shim
  use iso_c_binding
  implicit none

  interface
    ! fml/representations/frepresentations.f90:23-43
    subroutine fgenerate_coulomb_matrix(nuclear_charges, coordinates, natoms, nmax, cm)
      double precision, dimension(:), intent(in) :: nuclear_charges
      double precision, dimension(:,:), intent(in) :: coordinates
      integer :: natoms
      integer, intent(in) :: nmax
      double precision, dimension((nmax + 1) * nmax / 2), intent(out) :: cm
    end subroutine
    ! fml/kernels/fkernels.f90:317-327
    subroutine fgaussian_kernel(a, na, b, nb, k, sigma)
      double precision, dimension(:,:), intent(in) :: a
      double precision, dimension(:,:), intent(in) :: b
      integer, intent(in) :: na, nb
      double precision, dimension(:,:), intent(inout) :: k
      double precision, intent(in) :: sigma
    end subroutine
    ! fml/kernels/fkernels.f90:351-361
    subroutine flaplacian_kernel(a, na, b, nb, k, sigma)
      double precision, dimension(:,:), intent(in) :: a
      double precision, dimension(:,:), intent(in) :: b
      integer, intent(in) :: na, nb
      double precision, dimension(:,:), intent(inout) :: k
      double precision, intent(in) :: sigma
    end subroutine
    ! fml/kernels/farad_kernels.f90:439-478
    subroutine fget_kernels_arad(q1, q2, z1, z2, n1, n2, sigmas, nm1, nm2, nsigmas, &
        & width, cut_distance, r_width, c_width, kernels)
      double precision, dimension(:,:,:,:), intent(in) :: q1
      double precision, dimension(:,:,:,:), intent(in) :: q2
      double precision, dimension(:,:,:), intent(in) :: z1
      double precision, dimension(:,:,:), intent(in) :: z2
      integer, dimension(:), intent(in) :: n1
      integer, dimension(:), intent(in) :: n2
      double precision, dimension(:), intent(in) :: sigmas
      integer, intent(in) :: nm1
      integer, intent(in) :: nm2
      integer, intent(in) :: nsigmas
      double precision, intent(in) :: width
      double precision, intent(in) :: cut_distance
      double precision, intent(in) :: r_width
      double precision, intent(in) :: c_width
      double precision, dimension(nsigmas,nm1,nm2), intent(out) :: kernels
    end subroutine
    ! fml/kernels/farad_kernels.f90:619-654
    subroutine fget_symmetric_kernels_arad(q1, z1, n1, sigmas, nm1, nsigmas, &
        & width, cut_distance, r_width, c_width, kernels)
      double precision, dimension(:,:,:,:), intent(in) :: q1
      double precision, dimension(:,:,:), intent(in) :: z1
      integer, dimension(:), intent(in) :: n1
      double precision, dimension(:), intent(in) :: sigmas
      integer, intent(in) :: nm1
      integer, intent(in) :: nsigmas
      double precision, intent(in) :: width
      double precision, intent(in) :: cut_distance
      double precision, intent(in) :: r_width
      double precision, intent(in) :: c_width
      double precision, dimension(nsigmas,nm1,nm1), intent(out) :: kernels
    end subroutine
    ! fml/kernels/farad_kernels.f90:95-125
    subroutine fget_kernel_arad(q, z, n, sigma, nm, kernel)
      double precision, dimension(:,:,:,:), intent(in) :: q
      double precision, dimension(:,:,:), intent(in) :: z
      integer, dimension(:), intent(in) :: n
      double precision, intent(in) :: sigma
      integer :: nm
      double precision, dimension(nm,nm), intent(out) :: kernel
    end subroutine
    ! fml/kernels/farad_kernels.f90:757-796
    subroutine fget_atomic_kernels_arad(q1, q2, z1, z2, n1, n2, sigmas, na1, na2, nsigmas, &
        & width, cut_distance, r_width, c_width, kernels)
      double precision, dimension(:,:,:), intent(in) :: q1
      double precision, dimension(:,:,:), intent(in) :: q2
      double precision, dimension(:,:), intent(in) :: z1
      double precision, dimension(:,:), intent(in) :: z2
      double precision, dimension(:), intent(in) :: sigmas
      integer, dimension(:), intent(in) :: n1
      integer, dimension(:), intent(in) :: n2
      integer, intent(in) :: na1
      integer, intent(in) :: na2
      integer, intent(in) :: nsigmas
      double precision, intent(in) :: width
      double precision, intent(in) :: cut_distance
      double precision, intent(in) :: r_width
      double precision, intent(in) :: c_width
      double precision, dimension(nsigmas,na1,na2), intent(out) :: kernels
    end subroutine
    ! fml/kernels/farad_kernels.f90:902-932
    subroutine fget_atomic_distance_arad(q1, q2, z1, z2, n1, n2, na1, na2, &
        & width, cut_distance, r_width, c_width, kernel)
      double precision, dimension(:,:,:), intent(in) :: q1
      double precision, dimension(:,:,:), intent(in) :: q2
      double precision, dimension(:,:), intent(in) :: z1
      double precision, dimension(:,:), intent(in) :: z2
      integer, dimension(:), intent(in) :: n1
      integer, dimension(:), intent(in) :: n2
      integer, intent(in) :: na1
      integer, intent(in) :: na2
      double precision, intent(in) :: width
      double precision, intent(in) :: cut_distance
      double precision, intent(in) :: r_width
      double precision, intent(in) :: c_width
      double precision, dimension(na1,na2), intent(out) :: kernel
    end subroutine
    ! fml/math/farad.f90:452-485
    subroutine atomic_arad_l2_distance_all(q1, q2, z1, z2, n1, n2, nm1, nm2, width, &
        & cut_distance, r_width, c_width, amax, dmatrix)
      double precision, dimension(:,:,:,:), intent(in) :: q1
      double precision, dimension(:,:,:,:), intent(in) :: q2
      double precision, dimension(:,:,:), intent(in) :: z1
      double precision, dimension(:,:,:), intent(in) :: z2
      integer, dimension(:), intent(in) :: n1
      integer, dimension(:), intent(in) :: n2
      integer, intent(in) :: nm1
      integer, intent(in) :: nm2
      integer, intent(in) :: amax
      double precision, intent(in) :: width
      double precision, intent(in) :: cut_distance
      double precision, intent(in) :: r_width
      double precision, intent(in) :: c_width
      double precision, dimension(nm1,nm2,amax,amax), intent(out) :: dmatrix
    end subroutine
    ! fml/math/fcho_solve.f90:23-29
    subroutine fcho_solve(A, y, x)
      double precision, dimension(:,:), intent(in) :: A
      double precision, dimension(:), intent(in) :: y
      double precision, dimension(:), intent(inout) :: x
    end subroutine
    ! fml/math/fcho_solve.f90:49-53
    subroutine fcho_invert(A)
      double precision, dimension(:,:), intent(inout) :: A
    end subroutine
    ! fml/kernels/farad_kernels.f90:231-266
    subroutine fget_alpha_arad(q, z, n, y, sigma, lambda, alpha)
      double precision, dimension(:,:,:,:), intent(in) :: q
      double precision, dimension(:,:,:), intent(in) :: z
      integer, dimension(:), intent(in) :: n
      double precision, dimension(:), intent(in) :: y
      double precision, intent(in) :: sigma
      double precision, intent(in) :: lambda
      double precision, dimension(:), intent(inout) :: alpha
    end subroutine
    ! fml/representations/frepresentations.f90:112-131
    subroutine fgenerate_unsorted_coulomb_matrix(nuclear_charges, coordinates, natoms, nmax, cm)
      double precision, dimension(:), intent(in) :: nuclear_charges
      double precision, dimension(:,:), intent(in) :: coordinates
      integer, intent(in) :: nmax
      double precision, dimension((nmax + 1) * nmax / 2), intent(out) :: cm
      integer :: natoms
    end subroutine
    ! fml/representations/frepresentations.f90:204-214
    subroutine fgenerate_local_coulomb_matrix(nuclear_charges, coordinates, natoms, nmax, cm)
      double precision, dimension(:), intent(in) :: nuclear_charges
      double precision, dimension(:,:), intent(in) :: coordinates
      integer, intent(in) :: natoms
      integer, intent(in) :: nmax
      double precision, dimension(natoms, (nmax + 1) * nmax / 2), intent(out) :: cm
    end subroutine
    ! fml/representations/frepresentations.f90:297-307
    subroutine fgenerate_atomic_coulomb_matrix(nuclear_charges, coordinates, natoms, nmax, cm)
      double precision, dimension(:), intent(in) :: nuclear_charges
      double precision, dimension(:,:), intent(in) :: coordinates
      integer, intent(in) :: natoms
      integer, intent(in) :: nmax
      double precision, dimension(natoms, (nmax + 1) * nmax / 2), intent(out) :: cm
    end subroutine
    ! fml/kernels/fkernels.f90:23-51
    subroutine fget_vector_kernels_laplacian(q1, q2, n1, n2, sigmas, nm1, nm2, nsigmas, kernels)
      double precision, dimension(:,:,:), intent(in) :: q1
      double precision, dimension(:,:,:), intent(in) :: q2
      integer, dimension(:), intent(in) :: n1
      integer, dimension(:), intent(in) :: n2
      double precision, dimension(:), intent(in) :: sigmas
      integer, intent(in) :: nm1
      integer, intent(in) :: nm2
      integer, intent(in) :: nsigmas
      double precision, dimension(nsigmas,nm1,nm2), intent(out) :: kernels
    end subroutine
    ! fml/kernels/fkernels.f90:93-122
    subroutine fget_vector_kernels_gaussian(q1, q2, n1, n2, sigmas, nm1, nm2, nsigmas, kernels)
      double precision, dimension(:,:,:), intent(in) :: q1
      double precision, dimension(:,:,:), intent(in) :: q2
      integer, dimension(:), intent(in) :: n1
      integer, dimension(:), intent(in) :: n2
      double precision, dimension(:), intent(in) :: sigmas
      integer, intent(in) :: nm1
      integer, intent(in) :: nm2
      integer, intent(in) :: nsigmas
      double precision, dimension(nsigmas,nm1,nm2), intent(out) :: kernels
    end subroutine
    ! fml/kernels/fkernels.f90:380-391
    subroutine fget_alpha(q, n, y, sigma, lambda, alpha)
      double precision, dimension(:,:), intent(in) :: q
      integer, dimension(:), intent(in) :: n
      double precision, dimension(:), intent(in) :: y
      double precision, intent(in) :: lambda
      double precision, intent(in) :: sigma
      double precision, dimension(:), intent(inout) :: alpha
    end subroutine
    ! fml/kernels/fkernels.f90:480-490
    subroutine fget_prediction(Q, Q2, N2, alpha, sigma, Y)
      double precision, dimension(:,:), intent(in) :: Q
      double precision, dimension(:,:), intent(in) :: Q2
      integer, dimension(:), intent(in) :: N2
      double precision, dimension(:), intent(in) :: alpha
      double precision, intent(in) :: sigma
      double precision, dimension(:), intent(inout) :: Y
    end subroutine
    ! fml/kernels/fkernels.f90:541-551
    subroutine fget_alpha_from_distance(d, n, y, dgamma, lambda, alpha)
      double precision, dimension(:,:), intent(in) :: d
      integer, dimension(:), intent(in) :: n
      double precision, dimension(:), intent(in) :: y
      double precision, intent(in) :: dgamma
      double precision, intent(in) :: lambda
      double precision, dimension(:), intent(inout) :: alpha
    end subroutine
    ! fml/kernels/fkernels.f90:645-654
    subroutine fget_prediction_from_distance(D2, N2, alpha, dgamma, Y)
      double precision, dimension(:,:), intent(in) :: D2
      integer, dimension(:), intent(in) :: N2
      double precision, dimension(:), intent(in) :: alpha
      double precision, intent(in) :: dgamma
      double precision, dimension(:), intent(inout) :: Y
    end subroutine
    ! fml/kernels/faras_kernels.f90:150-175
    subroutine fget_kernels_aras(x1, x2, n1, n2, nneigh1, nneigh2, sigmas, nm1, nm2, nsigmas, &
        & t_width, r_width, c_width, d_width, cut_distance, order, pd, angular_scale, kernels)
      double precision, dimension(:,:,:,:), intent(in) :: x1
      double precision, dimension(:,:,:,:), intent(in) :: x2
      integer, dimension(:), intent(in) :: n1
      integer, dimension(:), intent(in) :: n2
      integer, dimension(:,:), intent(in) :: nneigh1
      integer, dimension(:,:), intent(in) :: nneigh2
      double precision, dimension(:), intent(in) :: sigmas
      integer, intent(in) :: nm1
      integer, intent(in) :: nm2
      integer, intent(in) :: nsigmas
      double precision, intent(in) :: t_width
      double precision, intent(in) :: r_width
      double precision, intent(in) :: c_width
      double precision, intent(in) :: d_width
      double precision, intent(in) :: cut_distance
      integer, intent(in) :: order
      double precision, dimension(:,:), intent(in) :: pd
      double precision, intent(in) :: angular_scale
      double precision, dimension(nsigmas,nm1,nm2), intent(out) :: kernels
    end subroutine
    ! fml/kernels/faras_kernels.f90:486-508
    subroutine fget_symmetric_kernels_aras(x1, n1, nneigh1, sigmas, nm1, nsigmas, &
        & t_width, r_width, c_width, d_width, cut_distance, order, pd, angular_scale, kernels)
      double precision, dimension(:,:,:,:), intent(in) :: x1
      integer, dimension(:), intent(in) :: n1
      integer, dimension(:,:), intent(in) :: nneigh1
      double precision, dimension(:), intent(in) :: sigmas
      integer, intent(in) :: nm1
      integer, intent(in) :: nsigmas
      double precision, intent(in) :: t_width
      double precision, intent(in) :: r_width
      double precision, intent(in) :: c_width
      double precision, intent(in) :: d_width
      double precision, intent(in) :: cut_distance
      integer, intent(in) :: order
      double precision, dimension(:,:), intent(in) :: pd
      double precision, intent(in) :: angular_scale
      double precision, dimension(nsigmas,nm1,nm1), intent(out) :: kernels
    end subroutine
  end interface

contains

  subroutine ref_coulomb_matrix(z, r, natoms, nmax, cm) bind(C, name="ref_coulomb_matrix")
    integer(c_int), value :: natoms, nmax
    real(c_double), intent(in) :: z(natoms), r(natoms, 3)
    real(c_double), intent(out) :: cm((nmax + 1) * nmax / 2)
    integer :: n
    n = natoms
    call fgenerate_coulomb_matrix(z, r, n, nmax, cm)
  end subroutine

  subroutine ref_gaussian_kernel(a, na, b, nb, d, k, sigma) bind(C, name="ref_gaussian_kernel")
    integer(c_int), value :: na, nb, d
    real(c_double), intent(in) :: a(d, na), b(d, nb)
    real(c_double), intent(inout) :: k(na, nb)
    real(c_double), value :: sigma
    call fgaussian_kernel(a, na, b, nb, k, sigma)
  end subroutine

  subroutine ref_laplacian_kernel(a, na, b, nb, d, k, sigma) bind(C, name="ref_laplacian_kernel")
    integer(c_int), value :: na, nb, d
    real(c_double), intent(in) :: a(d, na), b(d, nb)
    real(c_double), intent(inout) :: k(na, nb)
    real(c_double), value :: sigma
    call flaplacian_kernel(a, na, b, nb, k, sigma)
  end subroutine

  subroutine ref_kernels_arad(q1, q2, z1, z2, n1, n2, sigmas, nm1, nm2, nsig, amax, &
      & width, cut, rw, cw, kernels) bind(C, name="ref_kernels_arad")
    integer(c_int), value :: nm1, nm2, nsig, amax
    real(c_double), intent(in) :: q1(nm1, amax, 5, amax), q2(nm2, amax, 5, amax)
    real(c_double), intent(in) :: z1(nm1, amax, 2), z2(nm2, amax, 2)
    integer(c_int), intent(in) :: n1(nm1), n2(nm2)
    real(c_double), intent(in) :: sigmas(nsig)
    real(c_double), value :: width, cut, rw, cw
    real(c_double), intent(out) :: kernels(nsig, nm1, nm2)
    call fget_kernels_arad(q1, q2, z1, z2, n1, n2, sigmas, nm1, nm2, nsig, &
        & width, cut, rw, cw, kernels)
  end subroutine

  subroutine ref_symmetric_kernels_arad(q1, z1, n1, sigmas, nm1, nsig, amax, &
      & width, cut, rw, cw, kernels) bind(C, name="ref_symmetric_kernels_arad")
    integer(c_int), value :: nm1, nsig, amax
    real(c_double), intent(in) :: q1(nm1, amax, 5, amax)
    real(c_double), intent(in) :: z1(nm1, amax, 2)
    integer(c_int), intent(in) :: n1(nm1)
    real(c_double), intent(in) :: sigmas(nsig)
    real(c_double), value :: width, cut, rw, cw
    real(c_double), intent(out) :: kernels(nsig, nm1, nm1)
    call fget_symmetric_kernels_arad(q1, z1, n1, sigmas, nm1, nsig, &
        & width, cut, rw, cw, kernels)
  end subroutine

  subroutine ref_kernel_arad(q, z, n, sigma, nm, amax, kernel) bind(C, name="ref_kernel_arad")
    integer(c_int), value :: nm, amax
    real(c_double), intent(in) :: q(nm, amax, 5, amax)
    real(c_double), intent(in) :: z(nm, amax, 2)
    integer(c_int), intent(in) :: n(nm)
    real(c_double), value :: sigma
    real(c_double), intent(out) :: kernel(nm, nm)
    integer :: nmi
    nmi = nm
    call fget_kernel_arad(q, z, n, sigma, nmi, kernel)
  end subroutine

  subroutine ref_atomic_kernels_arad(q1, q2, z1, z2, n1, n2, sigmas, na1, na2, nsig, amax, &
      & width, cut, rw, cw, kernels) bind(C, name="ref_atomic_kernels_arad")
    integer(c_int), value :: na1, na2, nsig, amax
    real(c_double), intent(in) :: q1(na1, 5, amax), q2(na2, 5, amax)
    real(c_double), intent(in) :: z1(na1, 2), z2(na2, 2)
    integer(c_int), intent(in) :: n1(na1), n2(na2)
    real(c_double), intent(in) :: sigmas(nsig)
    real(c_double), value :: width, cut, rw, cw
    real(c_double), intent(out) :: kernels(nsig, na1, na2)
    call fget_atomic_kernels_arad(q1, q2, z1, z2, n1, n2, sigmas, na1, na2, nsig, &
        & width, cut, rw, cw, kernels)
  end subroutine

  subroutine ref_atomic_distance_arad(q1, q2, z1, z2, n1, n2, na1, na2, amax, &
      & width, cut, rw, cw, kernel) bind(C, name="ref_atomic_distance_arad")
    integer(c_int), value :: na1, na2, amax
    real(c_double), intent(in) :: q1(na1, 5, amax), q2(na2, 5, amax)
    real(c_double), intent(in) :: z1(na1, 2), z2(na2, 2)
    integer(c_int), intent(in) :: n1(na1), n2(na2)
    real(c_double), value :: width, cut, rw, cw
    real(c_double), intent(out) :: kernel(na1, na2)
    call fget_atomic_distance_arad(q1, q2, z1, z2, n1, n2, na1, na2, &
        & width, cut, rw, cw, kernel)
  end subroutine

  subroutine ref_arad_l2_distance_all(q1, q2, z1, z2, n1, n2, nm1, nm2, amax, &
      & width, cut, rw, cw, dmatrix) bind(C, name="ref_arad_l2_distance_all")
    integer(c_int), value :: nm1, nm2, amax
    real(c_double), intent(in) :: q1(nm1, amax, 5, amax), q2(nm2, amax, 5, amax)
    real(c_double), intent(in) :: z1(nm1, amax, 2), z2(nm2, amax, 2)
    integer(c_int), intent(in) :: n1(nm1), n2(nm2)
    real(c_double), value :: width, cut, rw, cw
    real(c_double), intent(out) :: dmatrix(nm1, nm2, amax, amax)
    call atomic_arad_l2_distance_all(q1, q2, z1, z2, n1, n2, nm1, nm2, width, &
        & cut, rw, cw, amax, dmatrix)
  end subroutine

  subroutine ref_cho_solve(a, n, y, x) bind(C, name="ref_cho_solve")
    integer(c_int), value :: n
    real(c_double), intent(inout) :: a(n, n)
    real(c_double), intent(in) :: y(n)
    real(c_double), intent(inout) :: x(n)
    call fcho_solve(a, y, x)
  end subroutine

  subroutine ref_cho_invert(a, n) bind(C, name="ref_cho_invert")
    integer(c_int), value :: n
    real(c_double), intent(inout) :: a(n, n)
    call fcho_invert(a)
  end subroutine

  subroutine ref_alpha_arad(q, z, n, y, sigma, lambda, nm, amax, na, alpha) &
      & bind(C, name="ref_alpha_arad")
    integer(c_int), value :: nm, amax, na
    real(c_double), intent(in) :: q(nm, amax, 5, amax)
    real(c_double), intent(in) :: z(nm, amax, 2)
    integer(c_int), intent(in) :: n(nm)
    real(c_double), intent(in) :: y(nm)
    real(c_double), value :: sigma, lambda
    real(c_double), intent(inout) :: alpha(na)
    call fget_alpha_arad(q, z, n, y, sigma, lambda, alpha)
  end subroutine

  ! mode 0: unsorted (cm(packed)); 1: local, 2: atomic (cm(natoms, packed))
  subroutine ref_coulomb_form(mode, z, r, natoms, nmax, cm) bind(C, name="ref_coulomb_form")
    integer(c_int), value :: mode, natoms, nmax
    real(c_double), intent(in) :: z(natoms), r(natoms, 3)
    real(c_double), intent(out) :: cm(*)
    integer :: n
    n = natoms
    if (mode == 0) then
      call fgenerate_unsorted_coulomb_matrix(z, r, n, nmax, cm)
    else if (mode == 1) then
      call fgenerate_local_coulomb_matrix(z, r, n, nmax, cm)
    else
      call fgenerate_atomic_coulomb_matrix(z, r, n, nmax, cm)
    end if
  end subroutine

  ! kind 0: Gaussian, 1: Laplacian
  subroutine ref_vector_kernels(kind, q1, q2, n1, n2, sigmas, nm1, nm2, nsig, amax1, amax2, d, &
      & kernels) bind(C, name="ref_vector_kernels")
    integer(c_int), value :: kind, nm1, nm2, nsig, amax1, amax2, d
    real(c_double), intent(in) :: q1(d, amax1, nm1), q2(d, amax2, nm2)
    integer(c_int), intent(in) :: n1(nm1), n2(nm2)
    real(c_double), intent(in) :: sigmas(nsig)
    real(c_double), intent(out) :: kernels(nsig, nm1, nm2)
    if (kind == 1) then
      call fget_vector_kernels_laplacian(q1, q2, n1, n2, sigmas, nm1, nm2, nsig, kernels)
    else
      call fget_vector_kernels_gaussian(q1, q2, n1, n2, sigmas, nm1, nm2, nsig, kernels)
    end if
  end subroutine

  subroutine ref_get_alpha(q, n, y, sigma, lambda, na, nm, d, alpha) bind(C, name="ref_get_alpha")
    integer(c_int), value :: na, nm, d
    real(c_double), intent(in) :: q(d, na)
    integer(c_int), intent(in) :: n(nm)
    real(c_double), intent(in) :: y(nm)
    real(c_double), value :: sigma, lambda
    real(c_double), intent(inout) :: alpha(na)
    call fget_alpha(q, n, y, sigma, lambda, alpha)
  end subroutine

  subroutine ref_get_prediction(q, q2, n2, alpha, sigma, na, na2, nm2, d, y) &
      & bind(C, name="ref_get_prediction")
    integer(c_int), value :: na, na2, nm2, d
    real(c_double), intent(in) :: q(d, na), q2(d, na2)
    integer(c_int), intent(in) :: n2(nm2)
    real(c_double), intent(in) :: alpha(na)
    real(c_double), value :: sigma
    real(c_double), intent(inout) :: y(nm2)
    call fget_prediction(q, q2, n2, alpha, sigma, y)
  end subroutine

  subroutine ref_alpha_from_distance(dmat, n, y, dgamma, lambda, na, nm, alpha) &
      & bind(C, name="ref_alpha_from_distance")
    integer(c_int), value :: na, nm
    real(c_double), intent(in) :: dmat(na, na)
    integer(c_int), intent(in) :: n(nm)
    real(c_double), intent(in) :: y(nm)
    real(c_double), value :: dgamma, lambda
    real(c_double), intent(inout) :: alpha(na)
    call fget_alpha_from_distance(dmat, n, y, dgamma, lambda, alpha)
  end subroutine

  subroutine ref_prediction_from_distance(d2, n2, alpha, dgamma, na, na2, nm2, y) &
      & bind(C, name="ref_prediction_from_distance")
    integer(c_int), value :: na, na2, nm2
    real(c_double), intent(in) :: d2(na, na2)
    integer(c_int), intent(in) :: n2(nm2)
    real(c_double), intent(in) :: alpha(na)
    real(c_double), value :: dgamma
    real(c_double), intent(inout) :: y(nm2)
    call fget_prediction_from_distance(d2, n2, alpha, dgamma, y)
  end subroutine

  ! x: (nm, amax, 3 + amax, amax) Fortran (the numpy C array (amax, 3+amax, amax, nm) reversed)
  subroutine ref_kernels_aras(x1, x2, n1, n2, nn1, nn2, sigmas, nm1, nm2, nsig, amax, &
      & t_width, r_width, c_width, d_width, cut, order, pd, emax, angular_scale, kernels) &
      & bind(C, name="ref_kernels_aras")
    integer(c_int), value :: nm1, nm2, nsig, amax, order, emax
    real(c_double), intent(in) :: x1(nm1, amax, 3 + amax, amax), x2(nm2, amax, 3 + amax, amax)
    integer(c_int), intent(in) :: n1(nm1), n2(nm2), nn1(nm1, amax), nn2(nm2, amax)
    real(c_double), intent(in) :: sigmas(nsig), pd(emax, emax)
    real(c_double), value :: t_width, r_width, c_width, d_width, cut, angular_scale
    real(c_double), intent(out) :: kernels(nsig, nm1, nm2)
    integer :: ordi
    ordi = order
    call fget_kernels_aras(x1, x2, n1, n2, nn1, nn2, sigmas, nm1, nm2, nsig, &
        & t_width, r_width, c_width, d_width, cut, ordi, pd, angular_scale, kernels)
  end subroutine

  subroutine ref_symmetric_kernels_aras(x1, n1, nn1, sigmas, nm1, nsig, amax, &
      & t_width, r_width, c_width, d_width, cut, order, pd, emax, angular_scale, kernels) &
      & bind(C, name="ref_symmetric_kernels_aras")
    integer(c_int), value :: nm1, nsig, amax, order, emax
    real(c_double), intent(in) :: x1(nm1, amax, 3 + amax, amax)
    integer(c_int), intent(in) :: n1(nm1), nn1(nm1, amax)
    real(c_double), intent(in) :: sigmas(nsig), pd(emax, emax)
    real(c_double), value :: t_width, r_width, c_width, d_width, cut, angular_scale
    real(c_double), intent(out) :: kernels(nsig, nm1, nm1)
    integer :: ordi
    ordi = order
    call fget_symmetric_kernels_aras(x1, n1, nn1, sigmas, nm1, nsig, &
        & t_width, r_width, c_width, d_width, cut, ordi, pd, angular_scale, kernels)
  end subroutine

end module fml_ref_shim
