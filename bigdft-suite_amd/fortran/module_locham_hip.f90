!> Fortran binding of the MI355X local-Hamiltonian library (C ABI in
!! include/bigdft_hip.h).  It plays the role of the OpenCL glue of the
!! reference (bigdft/src/interface_ocl.f90): init_acceleration_OCL,
!! allocate_data_OCL, local_hamiltonian_OCL, finish_hamiltonian_OCL and
!! free_gpu_OCL have counterparts here with the same argument meaning,
!! reduced to plain arrays so the module compiles on its own.
module module_locham_hip
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: BDFT_PREC_FP64 = 0, BDFT_PREC_FP32 = 1, &
       BDFT_PREC_BF16X3 = 2, BDFT_PREC_BF16X4 = 3, BDFT_PREC_F16X3 = 4

  !> launch options (BDFT_OPT_* of bigdft_hip.h), e.g. a choice cached by
  !! the autotuner
  integer, parameter, public :: BDFT_OPT_LINE_PASSES = 1, BDFT_OPT_ZGROUP = 2, &
       BDFT_OPT_ZPOT = 3, BDFT_OPT_BWD = 4, BDFT_OPT_CHUNK_MB = 5, &
       BDFT_OPT_TILE_ORDER = 6, BDFT_OPT_PERSIST = 7

  public :: init_acceleration_hip, allocate_data_hip, local_hamiltonian_hip, &
       local_partial_density_hip, free_gpu_hip, bdft_error_message, &
       bdft_bounds_layout, bdft_locreg_bounds, bdft_locham_conf_batch_d, &
       bdft_precondition, bdft_ctx_set_option

  !> confpot_data (locregs/locreg_operations.f90:23-30) as the C struct
  !! bdft_confpot
  type, bind(C), public :: bdft_confpot
     integer(c_int) :: potorder
     integer(c_int), dimension(3) :: ioffset
     real(c_double) :: prefac
     real(c_double), dimension(3) :: hh
     real(c_double), dimension(3) :: rxyzConf
     real(c_double) :: damping
  end type bdft_confpot

  interface
     integer(c_int) function bdft_ctx_create(device, ctx) bind(C, name='bdft_ctx_create')
       import :: c_int, c_ptr
       integer(c_int), value :: device
       type(c_ptr) :: ctx
     end function bdft_ctx_create
     integer(c_int) function bdft_ctx_destroy(ctx) bind(C, name='bdft_ctx_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function bdft_ctx_destroy
     integer(c_int) function bdft_ctx_set_option(ctx, option, value) &
          bind(C, name='bdft_ctx_set_option')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: option, value
     end function bdft_ctx_set_option
     type(c_ptr) function bdft_last_error(ctx) bind(C, name='bdft_last_error')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function bdft_last_error
     integer(c_int) function bdft_lr_create(ctx, geocode, n, nfl, nfu, hgrid, &
          nseg_c, nvctr_c, keyg_c, keyv_c, nseg_f, nvctr_f, keyg_f, keyv_f, lr) &
          bind(C, name='bdft_lr_create')
       import :: c_int, c_ptr, c_char, c_double, c_int32_t
       type(c_ptr), value :: ctx
       character(kind=c_char), value :: geocode
       integer(c_int), dimension(3), intent(in) :: n, nfl, nfu
       real(c_double), dimension(3), intent(in) :: hgrid
       integer(c_int), value :: nseg_c, nvctr_c, nseg_f, nvctr_f
       integer(c_int32_t), dimension(*), intent(in) :: keyg_c, keyv_c, keyg_f, keyv_f
       type(c_ptr) :: lr
     end function bdft_lr_create
     integer(c_int) function bdft_lr_destroy(lr) bind(C, name='bdft_lr_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: lr
     end function bdft_lr_destroy
     integer(c_int) function bdft_local_hamiltonian(ctx, lr, norb, nspin, pot, spinsgn, &
          psi, hpsi, ekin, epot, precision) bind(C, name='bdft_local_hamiltonian')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, lr
       integer(c_int), value :: norb, nspin, precision
       real(c_double), dimension(*), intent(in) :: pot, spinsgn, psi
       real(c_double), dimension(*), intent(inout) :: hpsi
       real(c_double), dimension(*), intent(out) :: ekin, epot
     end function bdft_local_hamiltonian
     integer(c_int) function bdft_density(ctx, lr, norb, psi, hfac, spin, nspin, &
          rho, precision) bind(C, name='bdft_density')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx, lr
       integer(c_int), value :: norb, nspin, precision
       real(c_double), dimension(*), intent(in) :: psi, hfac
       integer(c_int32_t), dimension(*), intent(in) :: spin
       real(c_double), dimension(*), intent(inout) :: rho
     end function bdft_density
     !> batched Hamiltonian with one confining potential per orbital
     !! (device arrays as c_ptr; conf on the host)
     integer(c_int) function bdft_locham_conf_batch_d(ctx, norb, lrs, pot, pot_off, &
          potdim, ishift, psi, psi_off, hpsi, ekin, epot, conf, precision) &
          bind(C, name='bdft_locham_conf_batch_d')
       import :: c_int, c_ptr, bdft_confpot
       type(c_ptr), value :: ctx, lrs, pot, pot_off, potdim, ishift, psi, psi_off, &
            hpsi, ekin, epot
       integer(c_int), value :: norb, precision
       type(bdft_confpot), dimension(*), intent(in) :: conf
     end function bdft_locham_conf_batch_d
     !> device part of preconditionall_OCL (interface_ocl.f90:622-880) for
     !! norb orbitals of one region: x (the scaled guess left by
     !! precondition_preconditioner) <- S y after ncong CG steps; b the
     !! scaled right-hand side; cprecr from cprecr_from_eval
     integer(c_int) function bdft_precondition(ctx, lr, norb, cprecr, ncong, x, b) &
          bind(C, name='bdft_precondition')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, lr
       integer(c_int), value :: norb, ncong
       real(c_double), dimension(*), intent(in) :: cprecr, b
       real(c_double), dimension(*), intent(inout) :: x
     end function bdft_precondition
     !> offsets and index bounds of the 17 convolution-bounds arrays
     integer(c_int) function bdft_bounds_layout(n, nfl, nfu, offset, dims) &
          bind(C, name='bdft_bounds_layout')
       import :: c_int, c_int32_t, c_int64_t
       integer(c_int), dimension(3), intent(in) :: n, nfl, nfu
       integer(c_int64_t), dimension(18), intent(out) :: offset
       integer(c_int32_t), dimension(68), intent(out) :: dims
     end function bdft_bounds_layout
     !> locreg_bounds (bounds.f90:555-613) into one host buffer
     integer(c_int) function bdft_locreg_bounds(ctx, lr, bounds) &
          bind(C, name='bdft_locreg_bounds')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx, lr
       integer(c_int32_t), dimension(*), intent(out) :: bounds
     end function bdft_locreg_bounds
  end interface

contains

  !> message of the last failure on the context
  function bdft_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=512) :: msg
    character(kind=c_char), dimension(:), pointer :: s
    integer :: i
    msg = ''
    call c_f_pointer(bdft_last_error(ctx), s, [512])
    do i = 1, 512
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    end do
  end function bdft_error_message

  !> counterpart of init_acceleration_OCL (interface_ocl.f90:20-41)
  subroutine init_acceleration_hip(device, ctx, ierr)
    integer, intent(in) :: device
    type(c_ptr), intent(out) :: ctx
    integer, intent(out) :: ierr
    ctx = c_null_ptr
    ierr = bdft_ctx_create(int(device, c_int), ctx)
  end subroutine init_acceleration_hip

  !> counterpart of allocate_data_OCL (interface_ocl.f90:43-193): the region's
  !! keys (1-based keyg(2,nseg_c+nseg_f), keyv(nseg_c+nseg_f), coarse first,
  !! as in wavefunctions_descriptors) are validated and uploaded once
  subroutine allocate_data_hip(ctx, geocode, n1, n2, n3, nfl, nfu, hgrid, &
       nseg_c, nvctr_c, nseg_f, nvctr_f, keyg, keyv, lr, ierr)
    type(c_ptr), intent(in) :: ctx
    character(len=1), intent(in) :: geocode
    integer, intent(in) :: n1, n2, n3, nseg_c, nvctr_c, nseg_f, nvctr_f
    integer, dimension(3), intent(in) :: nfl, nfu
    real(kind=8), dimension(3), intent(in) :: hgrid
    integer, dimension(2, nseg_c + nseg_f), intent(in) :: keyg
    integer, dimension(nseg_c + nseg_f), intent(in) :: keyv
    type(c_ptr), intent(out) :: lr
    integer, intent(out) :: ierr
    integer(c_int32_t), dimension(2) :: dummy_g
    integer(c_int32_t), dimension(1) :: dummy_v
    lr = c_null_ptr
    dummy_g = 0
    dummy_v = 0
    if (nseg_f > 0) then
       ierr = bdft_lr_create(ctx, geocode, [n1, n2, n3], nfl, nfu, hgrid, &
            nseg_c, nvctr_c, keyg(:, 1:nseg_c), keyv(1:nseg_c), &
            nseg_f, nvctr_f, keyg(:, nseg_c + 1:), keyv(nseg_c + 1:), lr)
    else
       ierr = bdft_lr_create(ctx, geocode, [n1, n2, n3], nfl, nfu, hgrid, &
            nseg_c, nvctr_c, keyg(:, 1:nseg_c), keyv(1:nseg_c), &
            0, 0, dummy_g, dummy_v, lr)
    end if
  end subroutine allocate_data_hip

  !> counterpart of local_hamiltonian_OCL followed by finish_hamiltonian_OCL
  !! (interface_ocl.f90:391-607): hpsi(:,iorb) += H_loc psi(:,iorb) for the
  !! norbp orbitals, with pot(n1i*n2i*n3i, nspin) and the spin-down
  !! potential pot(:,2) for the orbitals whose spinsgn is not positive
  !! (rhopot_up / rhopot_down, interface_ocl.f90:487-491), and the weighted
  !! sums
  !!   ekin_sum = sum kwgt*occup(iorb)*ekin(iorb),
  !!   epot_sum = sum kwgt*occup(iorb)*epot(iorb)   (hpsiortho.f90:946-954)
  subroutine local_hamiltonian_hip(ctx, lr, ncoef, norbp, nspin, kwgt, occup, spinsgn, &
       pot, psi, hpsi, ekin_sum, epot_sum, precision, ierr)
    type(c_ptr), intent(in) :: ctx, lr
    integer, intent(in) :: ncoef, norbp, nspin, precision
    real(kind=8), intent(in) :: kwgt
    real(kind=8), dimension(norbp), intent(in) :: occup, spinsgn
    real(kind=8), dimension(*), intent(in) :: pot
    real(kind=8), dimension(ncoef, norbp), intent(in) :: psi
    real(kind=8), dimension(ncoef, norbp), intent(inout) :: hpsi
    real(kind=8), intent(out) :: ekin_sum, epot_sum
    integer, intent(out) :: ierr
    real(kind=8), dimension(norbp) :: ekin, epot
    integer :: iorb
    ierr = bdft_local_hamiltonian(ctx, lr, int(norbp, c_int), int(nspin, c_int), pot, &
         spinsgn, psi, hpsi, ekin, epot, int(precision, c_int))
    ekin_sum = 0.0d0
    epot_sum = 0.0d0
    if (ierr /= 0) return
    do iorb = 1, norbp
       ekin_sum = ekin_sum + kwgt * occup(iorb) * ekin(iorb)
       epot_sum = epot_sum + kwgt * occup(iorb) * epot(iorb)
    end do
  end subroutine local_hamiltonian_hip

  !> counterpart of local_partial_density_OCL (interface_ocl.f90:908-1040):
  !! rho_p(:,:,:,isjmp) += kwgt*occup(iorb)/volume_element * psir(iorb)**2,
  !! isjmp = 1 where spinsgn > 0, else 2 (sumrho.f90:425 and 515-519);
  !! volume_element = hxh*hyh*hzh of the fine grid; rho_p accumulated
  subroutine local_partial_density_hip(ctx, lr, ncoef, norbp, nspin, kwgt, occup, &
       spinsgn, hxh, hyh, hzh, psi, nrho, rho_p, precision, ierr)
    type(c_ptr), intent(in) :: ctx, lr
    integer, intent(in) :: ncoef, norbp, nspin, nrho, precision
    real(kind=8), intent(in) :: kwgt, hxh, hyh, hzh
    real(kind=8), dimension(norbp), intent(in) :: occup, spinsgn
    real(kind=8), dimension(ncoef, norbp), intent(in) :: psi
    real(kind=8), dimension(nrho, nspin), intent(inout) :: rho_p
    integer, intent(out) :: ierr
    real(kind=8), dimension(norbp) :: hfac
    integer(c_int32_t), dimension(norbp) :: spin
    integer :: iorb
    do iorb = 1, norbp
       hfac(iorb) = kwgt * (occup(iorb) / (hxh * hyh * hzh))
       spin(iorb) = 0
       if (spinsgn(iorb) <= 0.0d0) spin(iorb) = 1
    end do
    ierr = bdft_density(ctx, lr, int(norbp, c_int), psi, hfac, spin, &
         int(nspin, c_int), rho_p, int(precision, c_int))
  end subroutine local_partial_density_hip

  !> counterpart of free_gpu_OCL (interface_ocl.f90:196-269)
  subroutine free_gpu_hip(ctx, lr)
    type(c_ptr), intent(inout) :: ctx, lr
    integer(c_int) :: ierr
    if (c_associated(lr)) ierr = bdft_lr_destroy(lr)
    if (c_associated(ctx)) ierr = bdft_ctx_destroy(ctx)
    lr = c_null_ptr
    ctx = c_null_ptr
  end subroutine free_gpu_hip

end module module_locham_hip
