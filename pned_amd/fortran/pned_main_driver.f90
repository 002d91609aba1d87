!! pned_main_driver.f90 -- a Fortran host that runs the reference's call sequence
!! (S_one_half/Hybrid_TPS_main.f90:89-143) through pned_parallelized_diag on the GPU:
!!   parallelized_lnc1 -> parallelized_inv1 -> readwrite_vec_data -> check1 -> calc_DSF,
!!   calc_extended_DSF (twice the DSF steps from the files calc_DSF wrote),
!!   then change_to_TPS_hamiltonian (tables given) -> parallelized_TPS.
!! Input (unformatted stream file, argument 1): n, q, ibond, ipair, ibond_type,
!! xlocal_Ham, llocal_Ham (as int), the shifted TPS tables and L, then nq, q_list(2,nq),
!! nsq, sq_q(2,nsq), site_vec(2,n), and max_itr, s_maxitr, i_range, first, last, nvec,
!! split, delt, seed.  Prints ITR, E0 and the check1 expectation; the reference's result
!! files (basisme00000, acc_log_me00000.txt, sq.txt, SalphaSbeta_*, vec_q*, dsf_res_zz.dat*)
!! land in the working directory.
program pned_main_driver
  use iso_c_binding
  use pned_gpu
  use pned_parallelized_diag
  implicit none
  integer :: n, q, ibond, qq, itr, u, nq, nsq, max_itr, s_maxitr, i_range, first, last, nvec
  integer, allocatable :: ipair(:), ibond_type(:), llint(:,:,:), llt(:,:,:)
  complex(kind(0d0)), allocatable :: xl(:,:,:), xt(:,:,:)
  double precision, allocatable :: q_list(:,:), sq_q(:,:), site_vec(:,:)
  double precision :: ene(8), L, split, delt
  integer(c_int64_t) :: seed
  type(c_ptr) :: xvec
  character(len=512) :: fname
  call get_command_argument(1, fname)
  open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='old')
  read(u) n, q, ibond
  qq = q*q
  allocate(ipair(2*ibond), ibond_type(ibond), xl(qq, qq, 9), llint(qq, qq, 9), xt(qq, qq, 9), llt(qq, qq, 9))
  read(u) ipair, ibond_type, xl, llint, xt, llt, L
  read(u) nq
  allocate(q_list(2, nq))
  read(u) q_list, nsq
  allocate(sq_q(2, nsq), site_vec(2, n))
  read(u) sq_q, site_vec
  read(u) max_itr, s_maxitr, i_range, first, last, nvec, split, delt, seed
  close(u)
  call pned_init(n, q, ibond, ipair, ibond_type, xl, llint /= 0)
  call pned_set_run(max_itr, s_maxitr, i_range, split, delt, first, last, nvec, 'zz', seed, site_vec, q_list, sq_q)
  xvec = pned_new_vector()
  ! *** Eigenvalues, ground-state eigenvector (Hybrid_TPS_main.f90:89-105)
  call parallelized_lnc1(nvec, itr, ene, xvec, .false.)
  write(*, '(A,I8)') 'ITR ', itr
  write(*, '(A,F24.16)') 'E0 ', ene(1)
  call parallelized_inv1(ene(1), xvec, .false.)
  ! *** Precision check (Hybrid_TPS_main.f90:119-120)
  call readwrite_vec_data(0, .true., ene(1), xvec)
  call check1(xvec)
  ! *** DSF (Hybrid_TPS_main.f90:139-143), then the COMP=zr/xr-style continuation of the same
  ! files to twice the steps (calc_extended_DSF); the first run's files are kept as first_*
  call calc_DSF(0, ene(1), xvec, 0)
  call execute_command_line('for f in dsf_res_zz.datcl00000lme00000 SalphaSbeta_q*me00000zz vec_q*me00000zz; '// &
                            'do cp "$f" "first_$f"; done')
  call pned_set_dsf_steps(2 * s_maxitr)
  call calc_extended_DSF(0, ene(1), xvec, 0)
  ! *** mTPQ (Hybrid_TPS_main.f90:108-114): change_to_TPS_hamiltonian, then parallelized_TPS
  call pned_update_matrices(q, xt, llt /= 0)
  localHam_max = L
  call parallelized_TPS(nvec, itr, ene, .false.)
  write(*, '(A)') 'DONE'
end program pned_main_driver
