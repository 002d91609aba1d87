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

!! Reference-named entry points on device handles.
module pned_parallelized_diag
  use iso_c_binding
  use pned_gpu
  implicit none
  type(c_ptr), save :: pned_context = c_null_ptr
contains
  subroutine pned_init(n, q, ibond, ipair, ibond_type, xlocal_ham, llocal_ham)
    integer, intent(in) :: n, q, ibond, ipair(:), ibond_type(:)
    complex(kind(0d0)), intent(in), target :: xlocal_ham(:,:,:)
    logical, intent(in) :: llocal_ham(:,:,:)
    integer(c_int), allocatable :: ll(:)
    real(c_double), allocatable :: xl(:)
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
    call pned_check(pned_ctx_create(int(n, c_int), int(q, c_int), int(ibond, c_int), int(ipair, c_int), &
                    int(ibond_type, c_int), xl, ll, 0_c_int, pned_context), 'pned_ctx_create')
  end subroutine pned_init

  function pned_new_vector() result(v)
    type(c_ptr) :: v
    call pned_check(pned_vec_create(pned_context, 1_c_int, PNED_C128, v), 'pned_vec_create')
  end function pned_new_vector

  !! parallelized_set_initial_vec_p (Hybrid_TPS_parallel.f90:17-102), seeded Gaussian
  subroutine parallelized_set_initial_vec_p(v, seed)
    type(c_ptr), intent(in) :: v
    integer(c_int64_t), intent(in) :: seed
    call pned_check(pned_set_initial_vec(pned_context, v, seed), 'set_initial_vec')
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

  !! calc_vec_inprdct (Hybrid_TPS_parallel.f90:104-126)
  subroutine calc_vec_inprdct(x, y, norm)
    type(c_ptr), intent(in) :: x, y
    complex(kind(0d0)), intent(inout) :: norm
    real(c_double) :: p(2)
    call pned_check(pned_calc_vec_inprdct(pned_context, x, y, p), 'calc_vec_inprdct')
    norm = dcmplx(p(1), p(2))
  end subroutine

  !! calc_alpha_beta (Hybrid_TPS_parallel.f90:145-159)
  subroutine calc_alpha_beta(t_alp, t_bet, v1, v0)
    double precision, intent(inout) :: t_alp, t_bet
    type(c_ptr), intent(in) :: v1, v0
    call pned_check(pned_calc_alpha_beta(pned_context, v1, v0, t_alp, t_bet), 'calc_alpha_beta')
  end subroutine

  !! parallelized_lnc1 (Hybrid_TPS_Lanczos.f90:7-45) from the start vector in v1
  subroutine parallelized_lnc1(nvec2, itr, ene, v1, v0, alpha, beta, max_itr)
    integer, intent(in) :: nvec2, max_itr
    integer, intent(inout) :: itr
    double precision, intent(inout) :: ene(:), alpha(:), beta(:)
    type(c_ptr), intent(in) :: v1, v0
    integer(c_int) :: it
    call pned_check(pned_lnc1(pned_context, v1, v0, int(max_itr, c_int), int(nvec2, c_int), 0_c_int, &
                    alpha, beta, ene, it), 'lnc1')
    itr = it
  end subroutine

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

  !! dump_alpha_beta_coef (Hybrid_TPS_TPS.f90:507-538): the mTPQ norms and observables of
  !! one sample, file SalphaSbeta_me<pme>seed<indx1><comp>
  subroutine dump_alpha_beta_coef(comp, indx1, pme, alpha, beta, coef)
    character(len=2), intent(in) :: comp
    integer, intent(in) :: indx1, pme
    double precision, intent(in) :: alpha(:), beta(:), coef(:,:)
    character(len=128) :: fname
    write(fname, '(a,i5.5,a,i5.5,a)') 'SalphaSbeta_me', pme, 'seed', indx1, trim(comp)
    call pned_check(pned_dump_alpha_beta_coef(trim(fname)//c_null_char, int(ubound(alpha, 1), c_int), &
                    int(ubound(coef, 2), c_int), alpha, beta, coef), 'dump_alpha_beta_coef')
  end subroutine

  !! dump_alpha_beta (Hybrid_TPS_DSF.f90:353-379): continued-fraction file
  !! SalphaSbeta_q<iq>me<pme><comp> and the vector file vec_q<iq>me<pme><comp>
  subroutine dump_alpha_beta(iq, pme, indx, sq, comp, salpha, sbeta, v1, v2)
    integer, intent(in) :: iq, pme, indx
    complex(kind(0d0)), intent(in) :: sq
    character(len=2), intent(in) :: comp
    double precision, intent(in) :: salpha(:), sbeta(:)
    type(c_ptr), intent(in) :: v1, v2
    character(len=128) :: fname
    real(c_double) :: s(2)
    type(c_ptr) :: vv(2)
    s = [dble(sq), dimag(sq)]
    write(fname, '(a,i5.5,a,i5.5,a)') 'SalphaSbeta_q', iq, 'me', pme, trim(comp)
    call pned_check(pned_dump_dsf_alpha_beta(trim(fname)//c_null_char, s, int(size(salpha), c_int), salpha, &
                    int(size(sbeta), c_int), sbeta), 'dump_alpha_beta')
    write(fname, '(a,i5.5,a,i5.5,a)') 'vec_q', iq, 'me', pme, trim(comp)
    vv = [v1, v2]
    call pned_check(pned_dump_vectors(trim(fname)//c_null_char, int(indx, c_int), 2_c_int, vv), 'dump_alpha_beta')
  end subroutine

  !! read_alpha_beta2 (Hybrid_TPS_DSF.f90:382-419): the restart counterpart of dump_alpha_beta
  subroutine read_alpha_beta2(iq, pme, indx, sq, comp, salpha, sbeta, v1, v2)
    integer, intent(in) :: iq, pme
    integer, intent(inout) :: indx
    complex(kind(0d0)), intent(inout) :: sq
    character(len=2), intent(in) :: comp
    double precision, intent(inout) :: salpha(:), sbeta(:)
    type(c_ptr), intent(in) :: v1, v2
    character(len=128) :: fname
    real(c_double) :: s(2)
    integer(c_int) :: ix
    type(c_ptr) :: vv(2)
    write(fname, '(a,i5.5,a,i5.5,a)') 'SalphaSbeta_q', iq, 'me', pme, trim(comp)
    call pned_check(pned_read_dsf_alpha_beta(trim(fname)//c_null_char, s, int(size(salpha), c_int), salpha, &
                    int(size(sbeta), c_int), sbeta), 'read_alpha_beta2')
    sq = dcmplx(s(1), s(2))
    write(fname, '(a,i5.5,a,i5.5,a)') 'vec_q', iq, 'me', pme, trim(comp)
    vv = [v1, v2]
    call pned_check(pned_read_vectors(trim(fname)//c_null_char, ix, 2_c_int, vv), 'read_alpha_beta2')
    indx = ix
  end subroutine
end module pned_parallelized_diag
