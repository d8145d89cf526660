! lbm_capi_mod.f90 -- iso_c_binding interface to the C ABI in include/lbm_capi.h.
!
! The reference drives its solver through argument-less subroutine calls on
! the module-global state `commondata` (natural-convection_2dim_mono.f90:1-72,
! program main at :283-380).  This module gives a Fortran program the same
! solver behind a handle: one interface per exported C function, the
! lbm_params / lbm_launch structs as bind(C) derived types, and two helpers,
! lbm_message (C string -> Fortran string) and lbm_require (stop with the
! status text on any non-zero status).
!
! Nullable array arguments (z in 2D, w in 2D, fields to skip) are passed as
! type(c_ptr), value: c_loc(array) or c_null_ptr.
module lbm_capi
    use, intrinsic :: iso_c_binding
    implicit none
    private :: c_strlen

    integer(c_int32_t), parameter :: LBM_OK = 0, LBM_ERR_ARG = 1, LBM_ERR_HIP = 2, &
        LBM_ERR_DIVERGED = 3, LBM_ERR_NOT_CONVERGED = 4, LBM_ERR_STATE = 5, LBM_ERR_NOMEM = 6
    integer(c_int32_t), parameter :: LBM_GRID_ERF = 0, LBM_GRID_TANH = 1
    integer(c_int32_t), parameter :: LBM_PREC_FP32_SHIFTED = 0, LBM_PREC_FP64 = 1
    ! lbm_get_info slots, 1-based (the C enum plus one)
    integer, parameter :: LBM_INFO_LLB = 1, LBM_INFO_DT = 2, LBM_INFO_NU = 3, &
        LBM_INFO_KAPPA = 4, LBM_INFO_GBETA = 5, LBM_INFO_TAUF = 6, LBM_INFO_SNU = 7, &
        LBM_INFO_SQ = 8, LBM_INFO_SIGK = 9, LBM_INFO_H = 10, LBM_INFO_COUNT = 10

    type, bind(C) :: lbm_params
        integer(c_int32_t) :: dim, nx, ny, nz
        real(c_double)     :: Ra, Pr, Ma, Thot, Tcold, stretch
        integer(c_int32_t) :: grid_kind, precision
        real(c_double)     :: epsU, epsT
        integer(c_int32_t) :: check_every, device
        integer(c_int64_t) :: itc_max
    end type lbm_params

    type, bind(C) :: lbm_launch
        integer(c_int32_t) :: bx, by, remap, occupancy, variant
    end type lbm_launch

    interface
        integer(c_int32_t) function lbm_create(params, h) bind(C, name='lbm_create')
            import :: c_int32_t, c_ptr, lbm_params
            type(lbm_params), intent(in) :: params
            type(c_ptr), intent(out) :: h
        end function
        integer(c_int32_t) function lbm_destroy(h) bind(C, name='lbm_destroy')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lbm_get_grid(h, x, y, z) bind(C, name='lbm_get_grid')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h, x, y, z
        end function
        integer(c_int32_t) function lbm_get_info(h, info) bind(C, name='lbm_get_info')
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: info(*)
        end function
        integer(c_int32_t) function lbm_init_rest(h) bind(C, name='lbm_init_rest')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lbm_init_fields(h, u, v, w, t) bind(C, name='lbm_init_fields')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h, u, v, w, t
        end function
        integer(c_int32_t) function lbm_step(h, nsteps) bind(C, name='lbm_step')
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: h
            integer(c_int64_t), value :: nsteps
        end function
        integer(c_int32_t) function lbm_check(h, errU, errT) bind(C, name='lbm_check')
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: errU, errT
        end function
        ! lbm_check with the reduction's partial sums in the library's static
        ! device array (no allocation; calls that overlap on one GPU race)
        integer(c_int32_t) function lbm_check_noalloc(h, errU, errT) &
                bind(C, name='lbm_check_noalloc')
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: errU, errT
        end function
        integer(c_int32_t) function lbm_nusselt(h, nu_hot, nu_cold) bind(C, name='lbm_nusselt')
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: nu_hot, nu_cold
        end function
        integer(c_int32_t) function lbm_get_fields(h, rho, u, v, w, t) bind(C, name='lbm_get_fields')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h, rho, u, v, w, t
        end function
        integer(c_int32_t) function lbm_run(h, itc, errU, errT) bind(C, name='lbm_run')
            import :: c_int32_t, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: h
            integer(c_int64_t), intent(out) :: itc
            real(c_double), intent(out) :: errU, errT
        end function
        integer(c_int32_t) function lbm_get_itc(h, itc) bind(C, name='lbm_get_itc')
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: h
            integer(c_int64_t), intent(out) :: itc
        end function
        integer(c_int32_t) function lbm_time_steps(h, nsteps, ms) bind(C, name='lbm_time_steps')
            import :: c_int32_t, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: h
            integer(c_int64_t), value :: nsteps
            real(c_double), intent(out) :: ms
        end function
        integer(c_int32_t) function lbm_synchronize(h) bind(C, name='lbm_synchronize')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lbm_state_size(h, bytes) bind(C, name='lbm_state_size')
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: h
            integer(c_int64_t), intent(out) :: bytes
        end function
        integer(c_int32_t) function lbm_save_state(h, buf, bytes) bind(C, name='lbm_save_state')
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: h, buf
            integer(c_int64_t), value :: bytes
        end function
        integer(c_int32_t) function lbm_load_state(h, buf, bytes) bind(C, name='lbm_load_state')
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: h, buf
            integer(c_int64_t), value :: bytes
        end function
        integer(c_int32_t) function lbm_copy_state(dst, src) bind(C, name='lbm_copy_state')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: dst, src
        end function
        integer(c_int32_t) function lbm_set_launch(h, launch) bind(C, name='lbm_set_launch')
            import :: c_int32_t, c_ptr, lbm_launch
            type(c_ptr), value :: h
            type(lbm_launch), intent(in) :: launch
        end function
        integer(c_int32_t) function lbm_get_launch(h, launch) bind(C, name='lbm_get_launch')
            import :: c_int32_t, c_ptr, lbm_launch
            type(c_ptr), value :: h
            type(lbm_launch), intent(out) :: launch
        end function
        integer(c_int32_t) function lbm_autotune(h, steps_per_trial, best_ms) bind(C, name='lbm_autotune')
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: steps_per_trial
            real(c_double), intent(out) :: best_ms
        end function
        integer(c_int32_t) function lbm_bytes_per_update(h, bytes) bind(C, name='lbm_bytes_per_update')
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: bytes
        end function
        ! batches of 2D cavities: arrays hold one value per cavity
        integer(c_int32_t) function lbm_batch_create(params, count, b) bind(C, name='lbm_batch_create')
            import :: c_int32_t, c_ptr, lbm_params
            type(lbm_params), intent(in) :: params(*)
            integer(c_int32_t), value :: count
            type(c_ptr), intent(out) :: b
        end function
        integer(c_int32_t) function lbm_batch_destroy(b) bind(C, name='lbm_batch_destroy')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: b
        end function
        type(c_ptr) function lbm_batch_cavity(b, index) bind(C, name='lbm_batch_cavity')
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: b
            integer(c_int32_t), value :: index
        end function
        integer(c_int32_t) function lbm_batch_init_rest(b) bind(C, name='lbm_batch_init_rest')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: b
        end function
        integer(c_int32_t) function lbm_batch_step(b, nsteps) bind(C, name='lbm_batch_step')
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: b
            integer(c_int64_t), value :: nsteps
        end function
        integer(c_int32_t) function lbm_batch_check(b, errU, errT) bind(C, name='lbm_batch_check')
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: b
            real(c_double), intent(out) :: errU(*), errT(*)
        end function
        integer(c_int32_t) function lbm_batch_nusselt(b, nu_hot, nu_cold) bind(C, name='lbm_batch_nusselt')
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: b
            real(c_double), intent(out) :: nu_hot(*), nu_cold(*)
        end function
        integer(c_int32_t) function lbm_batch_run(b, itc, errU, errT, status) bind(C, name='lbm_batch_run')
            import :: c_int32_t, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: b
            integer(c_int64_t), intent(out) :: itc(*)
            real(c_double), intent(out) :: errU(*), errT(*)
            integer(c_int32_t), intent(out) :: status(*)
        end function
        integer(c_int32_t) function lbm_batch_active(b, cavities, nodes) bind(C, name='lbm_batch_active')
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: b
            integer(c_int32_t), intent(out) :: cavities
            integer(c_int64_t), intent(out) :: nodes
        end function
        integer(c_int32_t) function lbm_batch_time_steps(b, nsteps, ms) bind(C, name='lbm_batch_time_steps')
            import :: c_int32_t, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: b
            integer(c_int64_t), value :: nsteps
            real(c_double), intent(out) :: ms
        end function
        integer(c_int32_t) function lbm_batch_set_launch(b, launch) bind(C, name='lbm_batch_set_launch')
            import :: c_int32_t, c_ptr, lbm_launch
            type(c_ptr), value :: b
            type(lbm_launch), intent(in) :: launch
        end function
        integer(c_int32_t) function lbm_batch_get_launch(b, launch) bind(C, name='lbm_batch_get_launch')
            import :: c_int32_t, c_ptr, lbm_launch
            type(c_ptr), value :: b
            type(lbm_launch), intent(out) :: launch
        end function
        integer(c_int32_t) function lbm_batch_autotune(b, steps_per_trial, best_ms) bind(C, name='lbm_batch_autotune')
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: b
            integer(c_int32_t), value :: steps_per_trial
            real(c_double), intent(out) :: best_ms
        end function
        ! the learned warm start (include/lbm_nn.h): arrays are FP32,
        ! x(G, G, 4, batch) and y(G, G, 3, batch) in Fortran order
        integer(c_int32_t) function lbm_nn_create(grid, batch_max, seed, device, net) &
                bind(C, name='lbm_nn_create')
            import :: c_int32_t, c_int64_t, c_ptr
            integer(c_int32_t), value :: grid, batch_max, device
            integer(c_int64_t), value :: seed
            type(c_ptr), intent(out) :: net
        end function
        integer(c_int32_t) function lbm_nn_create_width(grid, width, batch_max, seed, device, &
                net) bind(C, name='lbm_nn_create_width')
            import :: c_int32_t, c_int64_t, c_ptr
            integer(c_int32_t), value :: grid, width, batch_max, device
            integer(c_int64_t), value :: seed
            type(c_ptr), intent(out) :: net
        end function
        integer(c_int32_t) function lbm_nn_destroy(net) bind(C, name='lbm_nn_destroy')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: net
        end function
        integer(c_int32_t) function lbm_nn_param_count(net, count) bind(C, name='lbm_nn_param_count')
            import :: c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: net
            integer(c_int64_t), intent(out) :: count
        end function
        integer(c_int32_t) function lbm_nn_get_params(net, params) bind(C, name='lbm_nn_get_params')
            import :: c_int32_t, c_ptr, c_float
            type(c_ptr), value :: net
            real(c_float), intent(out) :: params(*)
        end function
        integer(c_int32_t) function lbm_nn_set_params(net, params) bind(C, name='lbm_nn_set_params')
            import :: c_int32_t, c_ptr, c_float
            type(c_ptr), value :: net
            real(c_float), intent(in) :: params(*)
        end function
        integer(c_int32_t) function lbm_nn_forward(net, batch, x, y) bind(C, name='lbm_nn_forward')
            import :: c_int32_t, c_ptr, c_float
            type(c_ptr), value :: net
            integer(c_int32_t), value :: batch
            real(c_float), intent(in) :: x(*)
            real(c_float), intent(out) :: y(*)
        end function
        integer(c_int32_t) function lbm_nn_loss_grad(net, batch, x, target, loss, grad) &
                bind(C, name='lbm_nn_loss_grad')
            import :: c_int32_t, c_ptr, c_float
            type(c_ptr), value :: net
            integer(c_int32_t), value :: batch
            real(c_float), intent(in) :: x(*), target(*)
            real(c_float), intent(out) :: loss, grad(*)
        end function
        integer(c_int32_t) function lbm_nn_train_step(net, batch, x, target, lr, loss) &
                bind(C, name='lbm_nn_train_step')
            import :: c_int32_t, c_ptr, c_float
            type(c_ptr), value :: net
            integer(c_int32_t), value :: batch
            real(c_float), intent(in) :: x(*), target(*)
            real(c_float), value :: lr
            real(c_float), intent(out) :: loss
        end function
        type(c_ptr) function lbm_status_string(status) bind(C, name='lbm_status_string')
            import :: c_ptr, c_int32_t
            integer(c_int32_t), value :: status
        end function
        type(c_ptr) function lbm_last_error() bind(C, name='lbm_last_error')
            import :: c_ptr
        end function
        integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

contains

    ! Copies a NUL-terminated C string into a Fortran string.
    function lbm_message(p) result(s)
        type(c_ptr), intent(in) :: p
        character(len=:), allocatable :: s
        character(kind=c_char), pointer :: chars(:)
        integer :: i, n
        if (.not. c_associated(p)) then
            s = ''
            return
        end if
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: s)
        do i = 1, n
            s(i:i) = chars(i)
        end do
    end function lbm_message

    ! Stops the program with the status text and the detail message when a
    ! C-ABI call did not return LBM_OK.
    subroutine lbm_require(status, what)
        integer(c_int32_t), intent(in) :: status
        character(len=*), intent(in) :: what
        if (status /= LBM_OK) then
            write(*,'(a,a,i0,a,a,a,a)') what, ' failed: status ', status, ' (', &
                lbm_message(lbm_status_string(status)), ') ', lbm_message(lbm_last_error())
            error stop 1
        end if
    end subroutine lbm_require

end module lbm_capi
