!! pned_gpu_mod.f90 -- ISO_C_BINDING interface to libpned_gpu.so (include/pned_gpu.h).
!!
!! Module pned_gpu declares the C entry points; module pned_parallelized_diag gives them
!! the names and argument meaning of the reference's parallelized_diag, Lanczos, CG, TPS
!! and DSF routines (S_one_half/Hybrid_TPS_parallel.f90:6-8, Hybrid_TPS_Lanczos.f90:7,
!! Hybrid_TPS_cg.f90:7, Hybrid_TPS_TPS.f90:45, Hybrid_TPS_DSF.f90:18,195,263), with the
!! vectors held on the GPU as handles instead of host arrays.  A non-zero status is turned
!! into the reference's behaviour: print the message and stop.
module pned_gpu
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: PNED_C128 = 0, PNED_C64 = 1
  interface
    function pned_last_error() bind(C, name="pned_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function pned_device_count() bind(C, name="pned_device_count") result(n)
      import :: c_int
      integer(c_int) :: n
    end function
    function pned_ctx_create(n_sites, q_spin, ibond, ipair, ibond_type, xlocal_ham, llocal_ham, device, ctx) &
        bind(C, name="pned_ctx_create") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: n_sites, q_spin, ibond, device
      integer(c_int), intent(in) :: ipair(*), ibond_type(*), llocal_ham(*)
      real(c_double), intent(in) :: xlocal_ham(*)
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function
    function pned_ctx_destroy(ctx) bind(C, name="pned_ctx_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function pned_ctx_dim(ctx) bind(C, name="pned_ctx_dim") result(d)
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: d
    end function
    function pned_ctx_set_matrices(ctx, xlocal_ham, llocal_ham) bind(C, name="pned_ctx_set_matrices") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: xlocal_ham(*)
      integer(c_int), intent(in) :: llocal_ham(*)
      integer(c_int) :: rc
    end function
    function pned_vec_create(ctx, nvec, dtype, v) bind(C, name="pned_vec_create") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nvec, dtype
      type(c_ptr), intent(out) :: v
      integer(c_int) :: rc
    end function
    function pned_vec_destroy(v) bind(C, name="pned_vec_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: v
      integer(c_int) :: rc
    end function
    function pned_vec_upload(ctx, v, host, offset, count) bind(C, name="pned_vec_upload") result(rc)
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, v, host
      integer(c_int64_t), value :: offset, count
      integer(c_int) :: rc
    end function
    function pned_vec_download(ctx, v, host, offset, count) bind(C, name="pned_vec_download") result(rc)
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, v, host
      integer(c_int64_t), value :: offset, count
      integer(c_int) :: rc
    end function
    function pned_vec_zero(ctx, v) bind(C, name="pned_vec_zero") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, v
      integer(c_int) :: rc
    end function
    function pned_set_initial_vec(ctx, v, seed) bind(C, name="pned_set_initial_vec") result(rc)
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, v
      integer(c_int64_t), value :: seed
      integer(c_int) :: rc
    end function
    function pned_p_mltply_v3(ctx, v1, v0, eperbd, prdct) bind(C, name="pned_p_mltply_v3") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, v1, v0
      real(c_double), value :: eperbd
      real(c_double), intent(out) :: prdct(2)
      integer(c_int) :: rc
    end function
    function pned_calc_vec_inprdct(ctx, x, y, out) bind(C, name="pned_calc_vec_inprdct") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, x, y
      real(c_double), intent(out) :: out(2)
      integer(c_int) :: rc
    end function
    function pned_vec_copy(ctx, src, dst) bind(C, name="pned_vec_copy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, src, dst
      integer(c_int) :: rc
    end function
    function pned_calc_beta(ctx, alpha, v1, v0, beta2) bind(C, name="pned_calc_beta") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, v1, v0
      real(c_double), intent(in) :: alpha(2)
      real(c_double), intent(out) :: beta2
      integer(c_int) :: rc
    end function
    function pned_calc_alpha_beta(ctx, v1, v0, alpha, beta) bind(C, name="pned_calc_alpha_beta") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, v1, v0
      real(c_double), intent(out) :: alpha, beta
      integer(c_int) :: rc
    end function
    function pned_calc_alpha_beta2(ctx, v1, v0, alpha, beta) bind(C, name="pned_calc_alpha_beta2") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, v1, v0
      real(c_double), intent(out) :: alpha, beta
      integer(c_int) :: rc
    end function
    function pned_lnc1(ctx, v1, v0, max_itr, nvec, fixed_steps, alpha, beta, ene, itr) &
        bind(C, name="pned_lnc1") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, v1, v0
      integer(c_int), value :: max_itr, nvec, fixed_steps
      real(c_double), intent(out) :: alpha(*), beta(*), ene(*)
      integer(c_int), intent(out) :: itr
      integer(c_int) :: rc
    end function
    function pned_inv1(ctx, eig, x, b, p, r, y, max_inv, max_cg, saved_loop, iters) &
        bind(C, name="pned_inv1") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, x, b, p, r, y
      real(c_double), value :: eig
      integer(c_int), value :: max_inv, max_cg, saved_loop
      integer(c_int), intent(out) :: iters
      integer(c_int) :: rc
    end function
    function pned_check1(ctx, x, work, expect, ratios, max_ratios, nratios) bind(C, name="pned_check1") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, x, work
      real(c_double), intent(out) :: expect(2), ratios(*)
      integer(c_int), value :: max_ratios
      integer(c_int), intent(out) :: nratios
      integer(c_int) :: rc
    end function
    function pned_tps_set_observables(ctx, site_vec, nsq, sq_q) bind(C, name="pned_tps_set_observables") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: site_vec(*), sq_q(*)
      integer(c_int), value :: nsq
      integer(c_int) :: rc
    end function
    function pned_tps(ctx, v1, v0, vg, vt, lmax, nsteps, alpha, beta, coef) bind(C, name="pned_tps") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, v1, v0, vg, vt
      real(c_double), value :: lmax
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: alpha(*), beta(*), coef(*)
      integer(c_int) :: rc
    end function
    function pned_intalv_zz(ctx, x, v, q, site_vec, cnorm) bind(C, name="pned_intalv_zz") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, x, v
      real(c_double), intent(in) :: q(2), site_vec(*)
      real(c_double), intent(out) :: cnorm
      integer(c_int) :: rc
    end function
    function pned_intalv_xy(ctx, x, v, q, site_vec, ipm, cnorm) bind(C, name="pned_intalv_xy") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, x, v
      real(c_double), intent(in) :: q(2), site_vec(*)
      integer(c_int), value :: ipm
      real(c_double), intent(out) :: cnorm
      integer(c_int) :: rc
    end function
    function pned_dsflnc1z(ctx, v1, v0, cnorm, s_maxitr, salpha, sbeta) bind(C, name="pned_dsflnc1z") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, v1, v0
      real(c_double), value :: cnorm
      integer(c_int), value :: s_maxitr
      real(c_double), intent(out) :: salpha(*), sbeta(*)
      integer(c_int) :: rc
    end function
    function pned_dsflnc1z_span(ctx, v1, v0, cnorm, first, stop, s_maxitr, salpha, sbeta, icf) &
        bind(C, name="pned_dsflnc1z_span") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, v1, v0
      real(c_double), value :: cnorm
      integer(c_int), value :: first, stop, s_maxitr
      real(c_double), intent(inout) :: salpha(*), sbeta(*)
      integer(c_int), intent(out) :: icf
      integer(c_int) :: rc
    end function
    function pned_renbun(ctx, i_range, s_maxitr, split, ene, eta, salpha, sbeta, sq, w, dsf, conv) &
        bind(C, name="pned_renbun") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: i_range, s_maxitr
      real(c_double), value :: split, ene, eta, sq
      real(c_double), intent(in) :: salpha(*), sbeta(*)
      real(c_double), intent(out) :: w(*), dsf(*)
      integer(c_int), intent(out) :: conv(*)
      integer(c_int) :: rc
    end function
    function pned_ctx_sz_sector(ctx, n_down) bind(C, name="pned_ctx_sz_sector") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_down
      integer(c_int) :: rc
    end function
    function pned_cocg(ctx, e0, eta, b, x, r, p, y, max_it, tol, iters, resid) bind(C, name="pned_cocg") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, b, x, r, p, y
      real(c_double), value :: e0, eta, tol
      integer(c_int), value :: max_it
      integer(c_int), intent(out) :: iters
      real(c_double), intent(out) :: resid
      integer(c_int) :: rc
    end function
    function pned_ctpq_eval(nd1, nd2, seed_num, alpha, beta, coef, t_min, t_max, num_div, out) &
        bind(C, name="pned_ctpq_eval") result(rc)
      import :: c_int, c_double
      integer(c_int), value :: nd1, nd2, seed_num, num_div
      real(c_double), intent(in) :: alpha(*), beta(*), coef(*)
      real(c_double), value :: t_min, t_max
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function
    function pned_dump_alpha_beta_coef(path, nd1, nd2, alpha, beta, coef) &
        bind(C, name="pned_dump_alpha_beta_coef") result(rc)
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: nd1, nd2
      real(c_double), intent(in) :: alpha(*), beta(*), coef(*)
      integer(c_int) :: rc
    end function
    function pned_read_alpha_beta_coef(path, nd1, nd2, alpha, beta, coef, max1, max2) &
        bind(C, name="pned_read_alpha_beta_coef") result(rc)
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: nd1, nd2
      real(c_double), intent(out) :: alpha(*), beta(*), coef(*)
      integer(c_int), value :: max1, max2
      integer(c_int) :: rc
    end function
    function pned_dump_dsf_alpha_beta(path, sq, n_alpha, salpha, n_beta, sbeta) &
        bind(C, name="pned_dump_dsf_alpha_beta") result(rc)
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      real(c_double), intent(in) :: sq(2), salpha(*), sbeta(*)
      integer(c_int), value :: n_alpha, n_beta
      integer(c_int) :: rc
    end function
    function pned_read_dsf_alpha_beta(path, sq, n_alpha, salpha, n_beta, sbeta) &
        bind(C, name="pned_read_dsf_alpha_beta") result(rc)
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      real(c_double), intent(out) :: sq(2), salpha(*), sbeta(*)
      integer(c_int), value :: n_alpha, n_beta
      integer(c_int) :: rc
    end function
    function pned_dump_vectors(path, indx, nvec, vecs) bind(C, name="pned_dump_vectors") result(rc)
      import :: c_int, c_ptr, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: indx, nvec
      type(c_ptr), intent(in) :: vecs(*)
      integer(c_int) :: rc
    end function
    function pned_read_vectors(path, indx, nvec, vecs) bind(C, name="pned_read_vectors") result(rc)
      import :: c_int, c_ptr, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: indx
      integer(c_int), value :: nvec
      type(c_ptr), intent(in) :: vecs(*)
      integer(c_int) :: rc
    end function
    function pned_interval_dump(path, data_size, vec_size, nreal, reals, real_counts, nvec, vecs) &
        bind(C, name="pned_interval_dump") result(rc)
      import :: c_int, c_ptr, c_char, c_int64_t
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: data_size, vec_size, nreal, nvec
      type(c_ptr), intent(in) :: reals(*), vecs(*)
      integer(c_int64_t), intent(in) :: real_counts(*)
      integer(c_int) :: rc
    end function
    function pned_interval_dump_read(path, data_size, vec_size, nreal, reals, real_counts, nvec, vecs) &
        bind(C, name="pned_interval_dump_read") result(rc)
      import :: c_int, c_ptr, c_char, c_int64_t
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: data_size, vec_size
      integer(c_int), value :: nreal, nvec
      type(c_ptr), intent(in) :: reals(*), vecs(*)
      integer(c_int64_t), intent(in) :: real_counts(*)
      integer(c_int) :: rc
    end function
    function pned_dump_info_write(path, dump_name, np, npos, position, true_value, max_itr, dsf_comp, s_maxitr, &
                                  n, ibond, nvec, total_sz, itr_counter) bind(C, name="pned_dump_info_write") result(rc)
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: path(*), dump_name(*), dsf_comp(*)
      integer(c_int), value :: np, npos, true_value, max_itr, s_maxitr, n, ibond, nvec, total_sz
      integer(c_int), intent(in) :: position(*), itr_counter(12)
      integer(c_int) :: rc
    end function
    function pned_write_dsf_res(path, append, sq, q, ll, w, dsf, conv) bind(C, name="pned_write_dsf_res") result(rc)
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: append, ll
      real(c_double), value :: sq
      real(c_double), intent(in) :: q(2), w(*), dsf(*)
      integer(c_int), intent(in) :: conv(*)
      integer(c_int) :: rc
    end function
    function pned_basis_write(path, e, x) bind(C, name="pned_basis_write") result(rc)
      import :: c_int, c_ptr, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      real(c_double), value :: e
      type(c_ptr), value :: x
      integer(c_int) :: rc
    end function
    function pned_basis_read(path, e, x) bind(C, name="pned_basis_read") result(rc)
      import :: c_int, c_ptr, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      real(c_double), intent(out) :: e
      type(c_ptr), value :: x
      integer(c_int) :: rc
    end function
  end interface
contains
  !! print the library's message and stop, as the reference does on its errors
  subroutine pned_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    if (rc == 0) return
    call c_f_pointer(pned_last_error(), msg, [1024])
    write(*, '(A)', advance='no') ' #(pned)# '//where//': '
    do i = 1, 1024
      if (msg(i) == c_null_char) exit
      write(*, '(A)', advance='no') msg(i)
    end do
    write(*, *)
    stop
  end subroutine pned_check
end module pned_gpu

!! Reference-named entry points on device handles.  Every routine keeps the reference's
!! name and argument list, with type(c_ptr) vector handles where the reference passes
!! complex(8) arrays.  What the reference keeps in modules -- parameters (max_itr, s_maxitr,
!! i_range, split, delt, dsf_comp, itr_counter(9:10)), data_mem (site_vec, q_list) and
!! working_area (wwk(:,1:4), alpha, beta, coef, salpha, sbeta) -- lives here: pned_init
!! builds the context, pned_set_run the parameters, and the work vectors are created with
!! the context.
module pned_parallelized_diag
  use iso_c_binding
  use pned_gpu
  implicit none
  type(c_ptr), save :: pned_context = c_null_ptr
  ! working_area (Hybrid_TPS_params.f90:272-311): wwk(:,1:4) on the device, the rest on the host
  type(c_ptr), save :: wwk(4) = c_null_ptr
  double precision, allocatable, save :: pned_alpha(:), pned_beta(:), pned_coef(:,:)
  double precision, allocatable, save :: pned_salpha(:), pned_sbeta(:)
  ! parameters / data_mem values the Krylov routines read (defaults: Hybrid_TPS_params.f90:34-42)
  integer, save :: pned_max_itr = 1024, pned_s_maxitr = 4096, pned_i_range = 2000
  integer, save :: pned_first = 1, pned_last = 1          ! itr_counter(9:10): momenta / samples
  integer, save :: pned_nsites = 0, pned_nvec = 8
  double precision, save :: pned_split = 2.5d-3, pned_delt = 1d-2
  double precision, save :: localHam_max = 0d0            ! the TPS shift L
  integer(c_int64_t), save :: pned_seed = 20261202_c_int64_t
  character(len=2), save :: pned_dsf_comp = 'zz'
  double precision, allocatable, save :: pned_site_vec(:,:), pned_q_list(:,:), pned_sq_q(:,:)
contains
  subroutine pned_init(n, q, ibond, ipair, ibond_type, xlocal_ham, llocal_ham)
    integer, intent(in) :: n, q, ibond, ipair(:), ibond_type(:)
    complex(kind(0d0)), intent(in), target :: xlocal_ham(:,:,:)
    logical, intent(in) :: llocal_ham(:,:,:)
    integer(c_int), allocatable :: ll(:)
    real(c_double), allocatable :: xl(:)
    integer :: k
    call tables(q, xlocal_ham, llocal_ham, xl, ll)
    call pned_check(pned_ctx_create(int(n, c_int), int(q, c_int), int(ibond, c_int), int(ipair, c_int), &
                    int(ibond_type, c_int), xl, ll, 0_c_int, pned_context), 'pned_ctx_create')
    pned_nsites = n
    do k = 1, 4
      wwk(k) = pned_new_vector()
    end do
  end subroutine pned_init

  !! new bond matrices on the same cluster, e.g. after change_to_TPS_hamiltonian
  !! (Hybrid_TPS_params.f90:1600-1623) rewrote xlocal_Ham
  subroutine pned_update_matrices(q, xlocal_ham, llocal_ham)
    integer, intent(in) :: q
    complex(kind(0d0)), intent(in) :: xlocal_ham(:,:,:)
    logical, intent(in) :: llocal_ham(:,:,:)
    integer(c_int), allocatable :: ll(:)
    real(c_double), allocatable :: xl(:)
    call tables(q, xlocal_ham, llocal_ham, xl, ll)
    call pned_check(pned_ctx_set_matrices(pned_context, xl, ll), 'pned_ctx_set_matrices')
  end subroutine pned_update_matrices

  subroutine tables(q, xlocal_ham, llocal_ham, xl, ll)
    integer, intent(in) :: q
    complex(kind(0d0)), intent(in) :: xlocal_ham(:,:,:)
    logical, intent(in) :: llocal_ham(:,:,:)
    real(c_double), allocatable, intent(out) :: xl(:)
    integer(c_int), allocatable, intent(out) :: ll(:)
    integer :: i, j, k, m, qq
    qq = q*q
    allocate(xl(2*qq*qq*9), ll(qq*qq*9))
    m = 0
    do k = 1, 9
      do j = 1, qq
        do i = 1, qq
          m = m + 1
          xl(2*m-1) = dble(xlocal_ham(i, j, k))
          xl(2*m) = dimag(xlocal_ham(i, j, k))
          ll(m) = merge(1, 0, llocal_ham(i, j, k))
        end do
      end do
    end do
  end subroutine tables

  !! the run parameters the reference reads from its modules: iteration limits, the DSF
  !! frequency grid, the momentum (DSF) or sample (TPS) range, the cluster's site positions
  !! and momenta, the TPS structure-factor momenta, the start-vector seed
  subroutine pned_set_run(max_itr, s_maxitr, i_range, split, delt, first, last, nvec, dsf_comp, seed, &
                          site_vec, q_list, sq_q)
    integer, intent(in) :: max_itr, s_maxitr, i_range, first, last, nvec
    double precision, intent(in) :: split, delt
    character(len=2), intent(in) :: dsf_comp
    integer(c_int64_t), intent(in) :: seed
    double precision, intent(in) :: site_vec(:,:), q_list(:,:), sq_q(:,:)
    pned_max_itr = max_itr
    pned_s_maxitr = s_maxitr
    pned_i_range = i_range
    pned_split = split
    pned_delt = delt
    pned_first = first
    pned_last = last
    pned_nvec = nvec
    pned_dsf_comp = dsf_comp
    pned_seed = seed
    if (allocated(pned_site_vec)) deallocate(pned_site_vec, pned_q_list, pned_sq_q)
    allocate(pned_site_vec(2, size(site_vec, 2)), pned_q_list(2, size(q_list, 2)), pned_sq_q(2, size(sq_q, 2)))
    pned_site_vec = site_vec
    pned_q_list = q_list
    pned_sq_q = sq_q
    if (allocated(pned_alpha)) deallocate(pned_alpha, pned_beta, pned_coef, pned_salpha, pned_sbeta)
    ! working_area shapes (Hybrid_TPS_params.f90:285-286); coef also holds 7+sq_num columns
    allocate(pned_alpha(max_itr), pned_beta(max_itr), pned_coef(max_itr, max(nvec + 1, 7 + size(sq_q, 2))))
    allocate(pned_salpha(s_maxitr), pned_sbeta(s_maxitr))
    pned_alpha = 0d0
    pned_beta = 0d0
    pned_coef = 0d0
    pned_salpha = 0d0
    pned_sbeta = 0d0
    call pned_check(pned_tps_set_observables(pned_context, pned_site_vec, int(size(sq_q, 2), c_int), pned_sq_q), &
                    'tps_set_observables')
  end subroutine pned_set_run

  !! working_area arrays with the current parameters, when pned_set_run has not made them
  subroutine ensure_work()
    if (allocated(pned_alpha)) return
    allocate(pned_alpha(pned_max_itr), pned_beta(pned_max_itr), pned_coef(pned_max_itr, pned_nvec + 1))
    allocate(pned_salpha(pned_s_maxitr), pned_sbeta(pned_s_maxitr))
    pned_alpha = 0d0
    pned_beta = 0d0
    pned_coef = 0d0
    pned_salpha = 0d0
    pned_sbeta = 0d0
  end subroutine ensure_work

  function pned_new_vector() result(v)
    type(c_ptr) :: v
    call pned_check(pned_vec_create(pned_context, 1_c_int, PNED_C128, v), 'pned_vec_create')
  end function pned_new_vector

  !! parallelized_set_initial_vec_p (Hybrid_TPS_parallel.f90:17-102): seeded Gaussian
  !! (seed: pned_set_run's, unless given)
  subroutine parallelized_set_initial_vec_p(v, seed)
    type(c_ptr), intent(in) :: v
    integer(c_int64_t), intent(in), optional :: seed
    integer(c_int64_t) :: s
    s = pned_seed
    if (present(seed)) s = seed
    call pned_check(pned_set_initial_vec(pned_context, v, s), 'set_initial_vec')
  end subroutine

  !! p_mltply_v3 (Hybrid_TPS_parallel.f90:190-263): v0 = v0 + (H - eperbd) v1
  subroutine p_mltply_v3(v1, v0, prdct, eperbd)
    type(c_ptr), intent(in) :: v1, v0
    complex(kind(0d0)), intent(inout) :: prdct
    double precision, intent(in) :: eperbd
    real(c_double) :: p(2)
    call pned_check(pned_p_mltply_v3(pned_context, v1, v0, eperbd, p), 'p_mltply_v3')
    prdct = dcmplx(p(1), p(2))
  end subroutine

  !! calc_vec_inprdct (Hybrid_TPS_parallel.f90:104-126): norm = sum conj(x) y
  subroutine calc_vec_inprdct(ldim, x, y, norm)
    integer, intent(in) :: ldim
    type(c_ptr), intent(in) :: x, y
    complex(kind(0d0)), intent(inout) :: norm
    real(c_double) :: p(2)
    call pned_check(pned_calc_vec_inprdct(pned_context, x, y, p), 'calc_vec_inprdct')
    norm = dcmplx(p(1), p(2))
  end subroutine

  !! calc_beta (Hybrid_TPS_parallel.f90:161-188): beta = ||v0 - alpha v1||^2
  subroutine calc_beta(data_num, alpha, beta, v1, v0)
    integer, intent(in) :: data_num
    complex(kind(0d0)), intent(in) :: alpha
    double precision, intent(inout) :: beta
    type(c_ptr), intent(in) :: v1, v0
    real(c_double) :: a(2), b2
    a = [dble(alpha), dimag(alpha)]
    call pned_check(pned_calc_beta(pned_context, a, v1, v0, b2), 'calc_beta')
    beta = b2
  end subroutine

  !! calc_alpha_beta (Hybrid_TPS_parallel.f90:145-159): p_mltply_v3 then calc_beta
  subroutine calc_alpha_beta(ldim, t_alp, t_bet, v1, v0)
    integer, intent(in) :: ldim
    double precision, intent(inout) :: t_alp, t_bet
    type(c_ptr), intent(in) :: v1, v0
    call pned_check(pned_calc_alpha_beta(pned_context, v1, v0, t_alp, t_bet), 'calc_alpha_beta')
  end subroutine

  !! calc_alpha_beta2 (Hybrid_TPS_parallel.f90:128-142): p_mltply_v3 then ||v0||
  subroutine calc_alpha_beta2(ldim, t_alp, t_bet, v1, v0)
    integer, intent(in) :: ldim
    double precision, intent(inout) :: t_alp, t_bet
    type(c_ptr), intent(in) :: v1, v0
    call pned_check(pned_calc_alpha_beta2(pned_context, v1, v0, t_alp, t_bet), 'calc_alpha_beta2')
  end subroutine

  !! parallelized_lnc1 (Hybrid_TPS_Lanczos.f90:7-45): Lanczos eigenvalues from the seeded
  !! start vector, or from x when restart; wwk(1:2) are the Lanczos vectors, alpha/beta go
  !! to pned_alpha/pned_beta (working_area)
  subroutine parallelized_lnc1(nvec2, itr, ene, x, restart)
    integer, intent(in) :: nvec2
    integer, intent(inout) :: itr
    double precision :: ene(:)
    type(c_ptr), intent(in) :: x
    logical, intent(in) :: restart
    integer(c_int) :: it
    if (nvec2 < 0) then
      print *, ' #(E06)# Wrong value given to nvec in lnc1'
      print *, '         Only the eigenvalues are calculated'
      stop
    end if
    call ensure_work()
    if (.not. restart) then
      call parallelized_set_initial_vec_p(wwk(1))
    else
      call pned_check(pned_vec_copy(pned_context, x, wwk(1)), 'lnc1')
    end if
    call pned_check(pned_lnc1(pned_context, wwk(1), wwk(2), int(pned_max_itr, c_int), int(nvec2, c_int), 0_c_int, &
                    pned_alpha, pned_beta, ene, it), 'lnc1')
    itr = it
  end subroutine

  !! parallelized_inv1 (Hybrid_TPS_cg.f90:7-16): inverse iteration (inv1z + cg1, the
  !! reference's SAVEd loop flag kept) for the eigenvector of Eig into x; the start vector
  !! is seeded, or x itself when restart
  subroutine parallelized_inv1(Eig, x, restart)
    double precision, intent(inout) :: Eig
    type(c_ptr), intent(in) :: x
    logical, intent(in) :: restart
    integer(c_int) :: iters
    if (.not. restart) then
      call parallelized_set_initial_vec_p(wwk(1))
    else
      call pned_check(pned_vec_copy(pned_context, x, wwk(1)), 'inv1')
    end if
    call pned_check(pned_inv1(pned_context, Eig, x, wwk(1), wwk(2), wwk(3), wwk(4), 5000_c_int, 500_c_int, 1_c_int, &
                    iters), 'inv1')
  end subroutine

  !! check1 (Hybrid_TPS_check.f90:8-54): <x|H|x> and (Hx)_j/x_j appended to
  !! acc_log_me00000.txt with the reference's format statements
  subroutine check1(x)
    type(c_ptr), intent(in) :: x
    real(c_double) :: expect(2), ratios(2*4096)
    integer(c_int) :: nr
    integer :: i, u
    call pned_check(pned_check1(pned_context, x, wwk(1), expect, ratios, 4096_c_int, nr), 'check1')
    open(newunit=u, file='acc_log_me00000.txt', status='unknown', access='append')
    write(u, *)
    write(u, 200)
200 format(' ---------------------------- Information from check1')
    write(u, 210) dcmplx(expect(1), expect(2))
210 format(' <x*H*x> =(', 1pe16.8, ',', 1pe16.8, ')')
    write(u, 220)
220 format(' H*x(j)/x(j) (j=min(ldim/3,13),ldim,max(1,ldim/20))')
    write(u, 230) (dcmplx(ratios(2*i-1), ratios(2*i)), i = 1, nr)
230 format(4('(', 1pe16.6e3, ',', 1pe16.6e3, ')'))
    write(u, 240)
240 format(' ---------------------------------------------------')
    close(u)
  end subroutine check1

  !! parallelized_TPS (Hybrid_TPS_TPS.f90:45-81): one mTPQ run of pned_max_itr steps per
  !! sample pned_first..pned_last (the reference reuses the momentum range as the sample
  !! range), sample i starting from seed pned_seed + i; the operator must already be the
  !! shifted one (pned_update_matrices after change_to_TPS_hamiltonian) and localHam_max its
  !! shift L.  Each sample's alpha, beta, coef go to SalphaSbeta_me00000seed<i><dsf_comp>.
  subroutine parallelized_TPS(nvec2, itr, ene, restart)
    integer, intent(in) :: nvec2
    integer, intent(inout) :: itr
    double precision :: ene(:)
    logical, intent(in) :: restart
    integer :: i, k, m, ncoef
    real(c_double), allocatable :: cf(:,:)
    if (nvec2 < 0) then
      print *, ' #(E06)# Wrong value given to nvec in lnc1'
      print *, '         Only the eigenvalues are calculated'
      stop
    end if
    write(*, *) 'start TPS'
    ncoef = 7 + size(pned_sq_q, 2)
    allocate(cf(ncoef, pned_max_itr))
    do i = pned_first, pned_last
      if (.not. restart) call parallelized_set_initial_vec_p(wwk(1), pned_seed + int(i, c_int64_t))
      call pned_check(pned_tps(pned_context, wwk(1), wwk(2), wwk(3), wwk(4), localHam_max, &
                      int(pned_max_itr, c_int), pned_alpha, pned_beta, cf), 'TPS')
      pned_coef = 0d0
      do k = 1, pned_max_itr
        do m = 1, ncoef
          pned_coef(k, m) = cf(m, k)
        end do
      end do
      itr = max(pned_max_itr - 1, int(pned_ctx_dim(pned_context)))
      call dump_alpha_beta_coef(pned_dsf_comp, i, itr + 1, wwk(2), wwk(1), wwk(3))
    end do
  end subroutine parallelized_TPS

  !! calc_DSF (Hybrid_TPS_DSF.f90:446-512): for momenta pned_first..pned_last of the q_list,
  !! the excitation (S^z(q) for 'zz', S^+(q) for 'xx') of x, sq.txt, the continued-fraction
  !! Lanczos (s_maxitr steps), its files (dump_alpha_beta), the spectrum on i_range
  !! frequencies normalised to S(q), and dsf_res_<comp>.datcl<i_color>lme<Iam>
  subroutine calc_DSF(Iam, ene, x, i_color)
    integer, intent(in) :: Iam, i_color
    double precision, intent(in) :: ene
    type(c_ptr), intent(in) :: x
    integer :: i, u
    double precision :: q(2), w(pned_i_range), dsf(pned_i_range)
    real(c_double) :: cn
    real(c_double), allocatable :: sa(:), sb(:)
    integer(c_int) :: conv(pned_i_range)
    complex(kind(0d0)) :: norm
    character(len=64) :: ident, fres
    if (Iam == 0) write(*, *) 'start : calc_DSF'
    write(ident, '(A,I5.5,A,I5.5)') 'cl', i_color, 'lme', Iam
    fres = 'dsf_res_'//pned_dsf_comp//'.dat'//trim(ident)
    allocate(sa(pned_s_maxitr), sb(pned_s_maxitr + 1))
    do i = pned_first, pned_last
      q(1:2) = pned_q_list(1:2, i)
      write(*, '(A,I3,2F12.6)') ' q = ', i, q(1), q(2)
      if (pned_dsf_comp == 'zz') then
        call pned_check(pned_intalv_zz(pned_context, x, wwk(1), q, pned_site_vec, cn), 'intalv_zz')
      else if (pned_dsf_comp == 'xx') then
        call pned_check(pned_intalv_xy(pned_context, x, wwk(1), q, pned_site_vec, 1_c_int, cn), 'intalv_xy')
      else
        write(*, *) 'Error: DSF comp.'
        stop
      end if
      norm = dcmplx(cn, 0d0)
      if (Iam == 0) then
        open(newunit=u, file='sq.txt', status='unknown', access='append')
        write(u, '(A,I3,2F12.6,F25.18)') ' q = ', i, q(1), q(2), dble(norm)
        close(u)
      end if
      call pned_check(pned_dsflnc1z(pned_context, wwk(1), wwk(2), cn, int(pned_s_maxitr, c_int), sa, sb), 'dsflnc1z')
      pned_salpha = sa(1:pned_s_maxitr)
      pned_sbeta = sb(1:pned_s_maxitr)
      call dump_alpha_beta(i, pned_s_maxitr - 1, norm, pned_dsf_comp, wwk(1), wwk(2))
      call pned_check(pned_renbun(pned_context, int(pned_i_range, c_int), int(pned_s_maxitr, c_int), pned_split, &
                      ene, pned_delt, sa, sb, cn, w, dsf, conv), 'renbun')
      if (Iam == 0) call pned_check(pned_write_dsf_res(trim(fres)//c_null_char, merge(0_c_int, 1_c_int, &
                                    i == pned_first), cn, q, int(pned_i_range, c_int), w, dsf, conv), 'write_dsf_res')
    end do
  end subroutine calc_DSF

  !! calc_extended_DSF (Hybrid_TPS_DSF.f90:515-553), the COMP=zr/xr continuation: for each
  !! momentum, read_alpha_beta2 restores sq, salpha/sbeta(1..indx+1) and the two vectors of an
  !! earlier run, the Lanczos continues from step indx + 1 to pned_s_maxitr - 1
  !! (pned_set_dsf_steps raises it), then the files, the spectrum and the result block as
  !! in calc_DSF.  The result file name follows open_res_files: dsf_res_zz.dat / dsf_res_xx.dat
  !! for those components, dsf_comp_error.dat otherwise (Hybrid_TPS_output.f90:216-231).
  subroutine calc_extended_DSF(Iam, ene, x, i_color)
    integer, intent(in) :: Iam, i_color
    double precision, intent(in) :: ene
    type(c_ptr), intent(in) :: x
    integer :: i, icn
    integer(c_int) :: icf
    double precision :: q(2), w(pned_i_range), dsf(pned_i_range)
    real(c_double), allocatable :: sa(:), sb(:)
    integer(c_int) :: conv(pned_i_range)
    complex(kind(0d0)) :: norm
    character(len=64) :: ident, fres
    write(ident, '(A,I5.5,A,I5.5)') 'cl', i_color, 'lme', Iam
    if (pned_dsf_comp == 'zz' .or. pned_dsf_comp == 'xx') then
      fres = 'dsf_res_'//pned_dsf_comp//'.dat'//trim(ident)
    else
      fres = 'dsf_comp_error.dat'//trim(ident)
    end if
    allocate(sa(pned_s_maxitr), sb(pned_s_maxitr + 1))
    do i = pned_first, pned_last
      q(1:2) = pned_q_list(1:2, i)
      write(*, '(A,I3,2F12.6)') ' q = ', i, q(1), q(2)
      call read_alpha_beta2(i, icn, norm, pned_dsf_comp, wwk(1), wwk(2))
      sa = 0d0
      sb = 0d0
      sa(1:pned_s_maxitr) = pned_salpha
      sb(1:pned_s_maxitr) = pned_sbeta
      call pned_check(pned_dsflnc1z_span(pned_context, wwk(1), wwk(2), dble(norm), int(icn + 1, c_int), 0_c_int, &
                      int(pned_s_maxitr, c_int), sa, sb, icf), 'dsflnc1z')
      pned_salpha = sa(1:pned_s_maxitr)
      pned_sbeta = sb(1:pned_s_maxitr)
      call dump_alpha_beta(i, int(icf), norm, pned_dsf_comp, wwk(1), wwk(2))
      call pned_check(pned_renbun(pned_context, int(pned_i_range, c_int), int(pned_s_maxitr, c_int), pned_split, &
                      ene, pned_delt, sa, sb, dble(norm), w, dsf, conv), 'renbun')
      if (Iam == 0) call pned_check(pned_write_dsf_res(trim(fres)//c_null_char, merge(0_c_int, 1_c_int, &
                                    i == pned_first), dble(norm), q, int(pned_i_range, c_int), w, dsf, conv), &
                                    'write_dsf_res')
    end do
  end subroutine calc_extended_DSF

  !! a new s_maxitr (itr_counter(12)) for the DSF Lanczos, e.g. before calc_extended_DSF
  subroutine pned_set_dsf_steps(s_maxitr)
    integer, intent(in) :: s_maxitr
    pned_s_maxitr = s_maxitr
    if (allocated(pned_salpha)) deallocate(pned_salpha, pned_sbeta)
    allocate(pned_salpha(s_maxitr), pned_sbeta(s_maxitr))
    pned_salpha = 0d0
    pned_sbeta = 0d0
  end subroutine pned_set_dsf_steps

  !! restrict the context to the S^z sector with n_down down spins (parity_type = 0;
  !! the index tables of mk_list12, Hybrid_TPS_parallel.f90:452-578)
  subroutine pned_set_sz_sector(n_down)
    integer, intent(in) :: n_down
    call pned_check(pned_ctx_sz_sector(pned_context, int(n_down, c_int)), 'sz_sector')
  end subroutine

  !! readwrite_vec_data (Hybrid_TPS_output.f90:11-35): file "basisme<ind>", e then x;
  !! dump = .true. writes, .false. reads
  subroutine readwrite_vec_data(ind, dump, e, x)
    integer, intent(in) :: ind
    logical, intent(in) :: dump
    double precision, intent(inout) :: e
    type(c_ptr), intent(in) :: x
    character(len=32) :: fname
    write(fname, '(A,I5.5)') 'basisme', ind
    if (dump) then
      call pned_check(pned_basis_write(trim(fname)//c_null_char, e, x), 'readwrite_vec_data')
    else
      call pned_check(pned_basis_read(trim(fname)//c_null_char, e, x), 'readwrite_vec_data')
    end if
  end subroutine

  !! dump_alpha_beta_coef (Hybrid_TPS_TPS.f90:507-538): alpha, beta, coef (working_area) to
  !! SalphaSbeta_me00000seed<indx1><comp>; the vectors are not written (the reference's
  !! vector dump is commented out)
  subroutine dump_alpha_beta_coef(comp, indx1, indx2, v1, v2, v3)
    character(len=2), intent(in) :: comp
    integer, intent(in) :: indx1, indx2
    type(c_ptr), intent(in) :: v1, v2, v3
    character(len=128) :: fname
    write(fname, '(a,i5.5,a,i5.5,a)') 'SalphaSbeta_me', 0, 'seed', indx1, trim(comp)
    call pned_check(pned_dump_alpha_beta_coef(trim(fname)//c_null_char, int(ubound(pned_alpha, 1), c_int), &
                    int(ubound(pned_coef, 2), c_int), pned_alpha, pned_beta, pned_coef), 'dump_alpha_beta_coef')
  end subroutine

  !! dump_alpha_beta (Hybrid_TPS_DSF.f90:353-379): SalphaSbeta_q<iq>me00000<comp> (sq,
  !! salpha, sbeta of working_area) and vec_q<iq>me00000<comp> (indx and two vectors)
  subroutine dump_alpha_beta(iq, indx, sq, comp, v1, v2)
    integer, intent(in) :: iq, indx
    complex(kind(0d0)), intent(in) :: sq
    character(len=2), intent(in) :: comp
    type(c_ptr), intent(in) :: v1, v2
    character(len=128) :: fname
    real(c_double) :: s(2)
    type(c_ptr) :: vv(2)
    s = [dble(sq), dimag(sq)]
    write(fname, '(a,i5.5,a,i5.5,a)') 'SalphaSbeta_q', iq, 'me', 0, trim(comp)
    call pned_check(pned_dump_dsf_alpha_beta(trim(fname)//c_null_char, s, int(size(pned_salpha), c_int), &
                    pned_salpha, int(size(pned_sbeta), c_int), pned_sbeta), 'dump_alpha_beta')
    write(fname, '(a,i5.5,a,i5.5,a)') 'vec_q', iq, 'me', 0, trim(comp)
    vv = [v1, v2]
    call pned_check(pned_dump_vectors(trim(fname)//c_null_char, int(indx, c_int), 2_c_int, vv), 'dump_alpha_beta')
  end subroutine

  !! read_alpha_beta2 (Hybrid_TPS_DSF.f90:382-419): the restart counterpart of dump_alpha_beta;
  !! salpha/sbeta get the file's first indx+1 entries, the rest zero
  subroutine read_alpha_beta2(iq, indx, sq, comp, v1, v2)
    integer, intent(in) :: iq
    integer, intent(inout) :: indx
    complex(kind(0d0)), intent(inout) :: sq
    character(len=2), intent(in) :: comp
    type(c_ptr), intent(in) :: v1, v2
    character(len=128) :: fname
    real(c_double) :: s(2)
    real(c_double), allocatable :: ta(:), tb(:)
    integer(c_int) :: ix
    type(c_ptr) :: vv(2)
    write(fname, '(a,i5.5,a,i5.5,a)') 'vec_q', iq, 'me', 0, trim(comp)
    vv = [v1, v2]
    call pned_check(pned_read_vectors(trim(fname)//c_null_char, ix, 2_c_int, vv), 'read_alpha_beta2')
    indx = ix
    allocate(ta(indx + 1), tb(indx + 1))
    write(fname, '(a,i5.5,a,i5.5,a)') 'SalphaSbeta_q', iq, 'me', 0, trim(comp)
    call pned_check(pned_read_dsf_alpha_beta(trim(fname)//c_null_char, s, int(indx + 1, c_int), ta, &
                    int(indx + 1, c_int), tb), 'read_alpha_beta2')
    sq = dcmplx(s(1), s(2))
    pned_salpha = 0d0
    pned_sbeta = 0d0
    pned_salpha(1:indx + 1) = ta
    pned_sbeta(1:indx + 1) = tb
  end subroutine
end module pned_parallelized_diag
