!! pned_lanczos_driver.f90 -- Fortran host using pned_parallelized_diag end to end.
!! Reads the data_mem tables (n, q, ibond, ipair, ibond_type, xlocal_Ham, llocal_Ham) from
!! an unformatted stream file, runs the reference's Lanczos (lnc1 + bisec) on the GPU from
!! the seeded start vector and prints E0 and the iteration count.
program pned_lanczos_driver
  use iso_c_binding
  use pned_gpu
  use pned_parallelized_diag
  implicit none
  integer :: n, q, ibond, qq, itr, u
  integer, allocatable :: ipair(:), ibond_type(:), llint(:,:,:)
  complex(kind(0d0)), allocatable :: xl(:,:,:)
  logical, allocatable :: ll(:,:,:)
  double precision :: ene(8), alpha(1024), beta(1024)
  type(c_ptr) :: v1, v0
  character(len=512) :: fname
  complex(kind(0d0)) :: prd
  call get_command_argument(1, fname)
  open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='old')
  read(u) n, q, ibond
  qq = q*q
  allocate(ipair(2*ibond), ibond_type(ibond), xl(qq, qq, 9), ll(qq, qq, 9), llint(qq, qq, 9))
  read(u) ipair, ibond_type, xl, llint
  close(u)
  ll = llint /= 0
  call pned_init(n, q, ibond, ipair, ibond_type, xl, ll)
  v1 = pned_new_vector()
  v0 = pned_new_vector()
  call parallelized_set_initial_vec_p(v1, 20261202_c_int64_t)
  call parallelized_lnc1(8, itr, ene, v1, v0, alpha, beta, 1024)
  write(*, '(A,I6)') 'ITR ', itr
  write(*, '(A,F24.16)') 'E0 ', ene(1)
  call parallelized_set_initial_vec_p(v1, 20261202_c_int64_t)
  call pned_check(pned_vec_zero(pned_context, v0), 'zero')
  call p_mltply_v3(v1, v0, prd, 0d0)
  write(*, '(A,2F24.16)') 'PRDCT ', dble(prd), dimag(prd)
end program pned_lanczos_driver
