! ygpu_mod.f90 -- Fortran 2008 ISO_C_BINDING interface to libygpu.so.
!
! The Fortran side of the C-ABI declared in include/ygpu.h.  A Yelmo build
! links libygpu.so and uses this module in two places:
!
!   * linear_solver_matrix_solve (src/physics/solver_linear.F90:151) hands
!     its 1-based CSR arrays (lgs%a_ptr, lgs%a_index, lgs%a_value,
!     lgs%b_value, lgs%x_value) to ygpu_csr_solve unchanged;
!   * calc_velocity_ssa (src/physics/velocity_ssa.f90:60-63) is replaced by
!     one call of ygpu_velocity_ssa with the same arguments, in the same
!     order;
!   * calc_velocity_diva (src/physics/velocity_diva.f90:59-64) likewise by
!     ygpu_velocity_diva, whose last two arguments carry the optional
!     neff_hook (a bind(C) procedure, c_funloc) and its context pointer.
!
! Arrays are passed by address (assumed-size, contiguous); scalars by value
! where the C prototype takes them by value.  The handle is an opaque
! type(c_ptr).
module ygpu_mod
    use, intrinsic :: iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: YGPU_OK = 0
    integer(c_int), parameter, public :: YGPU_ERR_ARG = -1
    integer(c_int), parameter, public :: YGPU_ERR_DEVICE = -2
    integer(c_int), parameter, public :: YGPU_ERR_UNSUPPORTED = -3
    integer(c_int), parameter, public :: YGPU_ERR_NONFINITE = -6

    integer(c_int), parameter, public :: YGPU_BND_ZEROS = 0
    integer(c_int), parameter, public :: YGPU_BND_INFINITE = 1
    integer(c_int), parameter, public :: YGPU_BND_MASK = 2
    integer(c_int), parameter, public :: YGPU_BND_MISMIP3D = 3
    integer(c_int), parameter, public :: YGPU_BND_TROUGH = 4
    integer(c_int), parameter, public :: YGPU_BND_PERIODIC = 5
    integer(c_int), parameter, public :: YGPU_BND_PERIODIC_X = 6
    integer(c_int), parameter, public :: YGPU_BND_PERIODIC_Y = 7

    integer(c_int), parameter, public :: YGPU_NOALLOC_MAX_CELLS = 32768

    integer(c_int), parameter, public :: YGPU_PREC_FP64 = 0
    integer(c_int), parameter, public :: YGPU_PREC_REF = 1
    integer(c_int), parameter, public :: YGPU_PREC_MIXED = 2

    ! ygpu_ssa_params (include/ygpu.h): ssa_param_class with enums for strings
    type, bind(C), public :: ygpu_ssa_params
        integer(c_int32_t) :: ssa_solver, boundaries, visc_method, beta_method
        integer(c_int32_t) :: beta_gl_scale, beta_gl_stag, ssa_vel_lim_method
        integer(c_int32_t) :: ssa_iter_max, ssa_write_log, precision
        real(c_double) :: visc_const, beta_const, beta_q, beta_u0, beta_gl_f
        real(c_double) :: H_grnd_lim, beta_min, eps_0, ssa_vel_lim_tau, ssa_vel_max
        real(c_double) :: ssa_iter_rel, ssa_iter_conv, rho_ice, rho_sw, g
        integer(c_int32_t) :: no_slip, block
        character(kind=c_char) :: ssa_lis_opt(256)
    end type

    public :: ygpu_create, ygpu_destroy, ygpu_default_params, ygpu_csr_solve
    public :: ygpu_csr_solve_noalloc, ygpu_velocity_ssa, ygpu_velocity_diva
    public :: ygpu_set_lis_opt, ygpu_note, ygpu_error

    ! The neff_hook of calc_velocity_diva (velocity_diva.f90:99-107): called
    ! after every Picard iteration with that iteration's basal velocity in
    ! the caller's ux_b, uy_b (real_bytes each: 4 for wp = sp); it updates
    ! c_bed in place.  Implement it with this interface and pass c_funloc.
    abstract interface
        subroutine ygpu_neff_hook(ctx, nx, ny, real_bytes, c_bed, ux_b, uy_b) bind(C)
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: ctx, c_bed, ux_b, uy_b
            integer(c_int32_t), value :: nx, ny, real_bytes
        end subroutine
    end interface
    public :: ygpu_neff_hook

    interface
        integer(c_int) function ygpu_create(device, handle) bind(C, name="ygpu_create")
            import :: c_int, c_int32_t, c_ptr
            integer(c_int32_t), value :: device
            type(c_ptr) :: handle
        end function

        integer(c_int) function ygpu_destroy(handle) bind(C, name="ygpu_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function

        subroutine ygpu_default_params(par) bind(C, name="ygpu_default_params")
            import :: ygpu_ssa_params
            type(ygpu_ssa_params) :: par
        end subroutine

        ! linear_solver_matrix_solve (solver_linear.F90:151): 1-based CSR
        integer(c_int) function ygpu_csr_solve(handle, n, nnz, a_ptr, a_index, a_value, &
                b_value, x_value, opts, iters, status, rel_resid, seconds) &
                bind(C, name="ygpu_csr_solve")
            import :: c_int, c_int32_t, c_double, c_char, c_ptr
            type(c_ptr), value :: handle
            integer(c_int32_t), value :: n, nnz
            integer(c_int32_t) :: a_ptr(*), a_index(*)
            real(c_double) :: a_value(*), b_value(*), x_value(*)
            character(kind=c_char) :: opts(*)
            integer(c_int32_t) :: iters, status
            real(c_double) :: rel_resid, seconds
        end function

        ! the same without device allocation (systems of at most
        ! YGPU_NOALLOC_MAX_CELLS cells; calls that overlap on one GPU race)
        integer(c_int) function ygpu_csr_solve_noalloc(handle, n, nnz, a_ptr, a_index, &
                a_value, b_value, x_value, opts, iters, status, rel_resid, seconds) &
                bind(C, name="ygpu_csr_solve_noalloc")
            import :: c_int, c_int32_t, c_double, c_char, c_ptr
            type(c_ptr), value :: handle
            integer(c_int32_t), value :: n, nnz
            integer(c_int32_t) :: a_ptr(*), a_index(*)
            real(c_double) :: a_value(*), b_value(*), x_value(*)
            character(kind=c_char) :: opts(*)
            integer(c_int32_t) :: iters, status
            real(c_double) :: rel_resid, seconds
        end function

        type(c_ptr) function ygpu_last_error_c(handle) bind(C, name="ygpu_last_error")
            import :: c_ptr
            type(c_ptr), value :: handle
        end function

        type(c_ptr) function ygpu_last_note_c(handle) bind(C, name="ygpu_last_note")
            import :: c_ptr
            type(c_ptr), value :: handle
        end function

        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function

        ! calc_velocity_ssa (velocity_ssa.f90:60-63), wp = sp arrays
        integer(c_int) function ygpu_velocity_ssa(handle, par, nx, ny, nz, ux_b, uy_b, &
                taub_acx, taub_acy, visc_eff, visc_eff_int, ssa_mask_acx, ssa_mask_acy, &
                ssa_err_acx, ssa_err_acy, ssa_iter_now, ssa_lin_iter, ssa_lin_fail, &
                ssa_lim_n, beta, beta_acx, beta_acy, c_bed, f_slide, taud_acx, taud_acy, &
                taul_int_acx, taul_int_acy, H_ice, f_ice, H_grnd, f_grnd, f_grnd_acx, &
                f_grnd_acy, ATT, zeta_aa, z_sl, z_bed, z_srf, dx, dy, n_glen) &
                bind(C, name="ygpu_velocity_ssa")
            import :: c_int, c_int32_t, c_float, c_ptr, ygpu_ssa_params
            type(c_ptr), value :: handle
            type(ygpu_ssa_params) :: par
            integer(c_int32_t), value :: nx, ny, nz
            real(c_float) :: ux_b(*), uy_b(*), taub_acx(*), taub_acy(*), visc_eff(*)
            real(c_float) :: visc_eff_int(*)
            integer(c_int32_t) :: ssa_mask_acx(*), ssa_mask_acy(*)
            real(c_float) :: ssa_err_acx(*), ssa_err_acy(*)
            integer(c_int32_t) :: ssa_iter_now, ssa_lin_iter, ssa_lin_fail, ssa_lim_n
            real(c_float) :: beta(*), beta_acx(*), beta_acy(*)
            real(c_float) :: c_bed(*), f_slide(*), taud_acx(*), taud_acy(*)
            real(c_float) :: taul_int_acx(*), taul_int_acy(*), H_ice(*), f_ice(*)
            real(c_float) :: H_grnd(*), f_grnd(*), f_grnd_acx(*), f_grnd_acy(*)
            real(c_float) :: ATT(*), zeta_aa(*), z_sl(*), z_bed(*), z_srf(*)
            real(c_float), value :: dx, dy, n_glen
        end function

        ! calc_velocity_diva (velocity_diva.f90:59-64), wp = sp arrays; c_bed
        ! INOUT (the hook updates it); neff_hook = c_null_funptr for none
        integer(c_int) function ygpu_velocity_diva(handle, par, nx, ny, nz, ux, uy, ux_bar, &
                uy_bar, ux_b, uy_b, ux_i, uy_i, taub_acx, taub_acy, beta, beta_acx, beta_acy, &
                beta_eff, visc_eff, visc_eff_int, duxdz, duydz, ssa_mask_acx, ssa_mask_acy, &
                ssa_err_acx, ssa_err_acy, ssa_iter_now, ssa_lin_iter, ssa_lin_fail, ssa_lim_n, &
                c_bed, f_slide, taud_acx, taud_acy, taul_int_acx, taul_int_acy, H_ice, f_ice, &
                H_grnd, f_grnd, f_grnd_acx, f_grnd_acy, ATT, zeta_aa, z_sl, z_bed, z_srf, &
                dx, dy, n_glen, neff_hook, hook_ctx) bind(C, name="ygpu_velocity_diva")
            import :: c_int, c_int32_t, c_float, c_ptr, c_funptr, ygpu_ssa_params
            type(c_ptr), value :: handle
            type(ygpu_ssa_params) :: par
            integer(c_int32_t), value :: nx, ny, nz
            real(c_float) :: ux(*), uy(*), ux_bar(*), uy_bar(*), ux_b(*), uy_b(*)
            real(c_float) :: ux_i(*), uy_i(*), taub_acx(*), taub_acy(*)
            real(c_float) :: beta(*), beta_acx(*), beta_acy(*), beta_eff(*)
            real(c_float) :: visc_eff(*), visc_eff_int(*), duxdz(*), duydz(*)
            integer(c_int32_t) :: ssa_mask_acx(*), ssa_mask_acy(*)
            real(c_float) :: ssa_err_acx(*), ssa_err_acy(*)
            integer(c_int32_t) :: ssa_iter_now, ssa_lin_iter, ssa_lin_fail, ssa_lim_n
            real(c_float) :: c_bed(*), f_slide(*), taud_acx(*), taud_acy(*)
            real(c_float) :: taul_int_acx(*), taul_int_acy(*), H_ice(*), f_ice(*)
            real(c_float) :: H_grnd(*), f_grnd(*), f_grnd_acx(*), f_grnd_acy(*)
            real(c_float) :: ATT(*), zeta_aa(*), z_sl(*), z_bed(*), z_srf(*)
            real(c_float), value :: dx, dy, n_glen
            type(c_funptr), value :: neff_hook
            type(c_ptr), value :: hook_ctx
        end function
    end interface

contains

    ! Copies a Fortran string into par%ssa_lis_opt (NUL-terminated).
    subroutine ygpu_set_lis_opt(par, opt)
        type(ygpu_ssa_params), intent(inout) :: par
        character(len=*), intent(in) :: opt
        integer :: k, n
        n = min(len_trim(opt), size(par%ssa_lis_opt) - 1)
        do k = 1, n
            par%ssa_lis_opt(k) = opt(k:k)
        end do
        par%ssa_lis_opt(n + 1) = c_null_char
    end subroutine

    ! A C string returned by the library as a Fortran string.
    function c_to_string(p) result(str)
        type(c_ptr), intent(in) :: p
        character(len=:), allocatable :: str
        character(kind=c_char), pointer :: chars(:)
        integer :: k, n
        if (.not. c_associated(p)) then
            str = ""
            return
        end if
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [max(n, 1)])
        allocate(character(len=n) :: str)
        do k = 1, n
            str(k:k) = chars(k)
        end do
    end function

    ! ygpu_last_note: what the last call that read an option string did
    ! differently from what it asked (e.g. "-i minres solved with BiCGSTAB")
    function ygpu_note(handle) result(str)
        type(c_ptr), intent(in) :: handle
        character(len=:), allocatable :: str
        str = c_to_string(ygpu_last_note_c(handle))
    end function

    ! ygpu_last_error as a Fortran string
    function ygpu_error(handle) result(str)
        type(c_ptr), intent(in) :: handle
        character(len=:), allocatable :: str
        str = c_to_string(ygpu_last_error_c(handle))
    end function

end module ygpu_mod
