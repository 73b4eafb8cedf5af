!> Driver of module_locham_hip: a dense periodic region (keys built the way
!! segkeys does, bigdft/src/locregs/initlocregs.f90:1633-1686), deterministic
!! orbitals and potential, one call of local_hamiltonian_hip, and the
!! weighted energies printed for the test-suite to compare with the oracle.
!! Usage: locham_hip_driver n norb precision [line_passes [nspin]]
!! nspin 2: a second potential, orbitals 2,4,... spin down
program locham_hip_driver
  use iso_c_binding
  use module_locham_hip
  implicit none
  type(c_ptr) :: ctx, lr
  integer :: n, norb, prec, ierr, nseg, nvctr, iseg, i2, i3, i, iorb, ncoef, npot, nspin
  integer, dimension(:, :), allocatable :: keyg
  integer, dimension(:), allocatable :: keyv
  real(kind=8), dimension(:, :), allocatable :: psi, hpsi
  real(kind=8), dimension(:), allocatable :: pot, occup, spinsgn
  real(kind=8), dimension(:, :), allocatable :: rho
  real(kind=8) :: ekin, epot, hsum
  character(len=32) :: arg

  call get_command_argument(1, arg); read(arg, *) n
  call get_command_argument(2, arg); read(arg, *) norb
  call get_command_argument(3, arg); read(arg, *) prec

  ! every (i2,i3) row is one segment, coarse and fine alike
  nseg = (n + 1) ** 2
  nvctr = (n + 1) ** 3
  allocate(keyg(2, 2 * nseg), keyv(2 * nseg))
  do i3 = 0, n
     do i2 = 0, n
        iseg = i2 + 1 + (n + 1) * i3
        keyg(1, iseg) = i3 * (n + 1) ** 2 + i2 * (n + 1) + 1
        keyg(2, iseg) = keyg(1, iseg) + n
        keyv(iseg) = (iseg - 1) * (n + 1) + 1
     end do
  end do
  keyg(:, nseg + 1:) = keyg(:, 1:nseg)
  keyv(nseg + 1:) = keyv(1:nseg)

  nspin = 1
  if (command_argument_count() >= 5) then
     call get_command_argument(5, arg); read(arg, *) nspin
  end if
  ncoef = nvctr + 7 * nvctr
  npot = (2 * n + 2) ** 3
  allocate(psi(ncoef, norb), hpsi(ncoef, norb), pot(npot * 2), occup(norb), spinsgn(norb))
  do iorb = 1, norb
     do i = 1, ncoef
        psi(i, iorb) = sin(0.37d0 * i + 1.3d0 * iorb) * cos(0.011d0 * i)
     end do
     occup(iorb) = 2.0d0 / iorb
     spinsgn(iorb) = 1.0d0
     if (nspin == 2) spinsgn(iorb) = 1.0d0 - 2.0d0 * mod(iorb - 1, 2)
  end do
  do i = 1, npot
     pot(i) = -1.0d0 - cos(0.05d0 * i)
     pot(npot + i) = -0.5d0 - sin(0.07d0 * i)
  end do
  hpsi = 0.0d0

  call init_acceleration_hip(0, ctx, ierr)
  if (ierr /= 0) stop 'context creation failed'
  ! optional 4th argument 1: line passes for every batch (a launch option)
  if (command_argument_count() >= 4) then
     call get_command_argument(4, arg); read(arg, *) i
     ierr = bdft_ctx_set_option(ctx, BDFT_OPT_LINE_PASSES, int(i, c_int))
     if (ierr /= 0) stop 'option not accepted'
  end if
  call allocate_data_hip(ctx, 'P', n, n, n, [0, 0, 0], [n, n, n], &
       [0.45d0, 0.45d0, 0.45d0], nseg, nvctr, nseg, nvctr, keyg, keyv, lr, ierr)
  if (ierr /= 0) then
     print '(a)', trim(bdft_error_message(ctx))
     stop 'region creation failed'
  end if
  call local_hamiltonian_hip(ctx, lr, ncoef, norb, nspin, 1.0d0, occup, spinsgn, pot, &
       psi, hpsi, ekin, epot, prec, ierr)
  if (ierr /= 0) then
     print '(a)', trim(bdft_error_message(ctx))
     stop 'local hamiltonian failed'
  end if
  hsum = sum(hpsi * psi)
  print '(a,es25.17)', 'ekin_sum ', ekin
  print '(a,es25.17)', 'epot_sum ', epot
  print '(a,es25.17)', 'psi_hpsi ', hsum
  ! density, alternating spins, accumulated on a 1e-20 prefill
  allocate(rho(npot, 2))
  rho = 1.0d-20
  do iorb = 1, norb
     spinsgn(iorb) = 1.0d0 - 2.0d0 * mod(iorb - 1, 2)
  end do
  call local_partial_density_hip(ctx, lr, ncoef, norb, 2, 1.0d0, occup, spinsgn, &
       0.225d0, 0.225d0, 0.225d0, psi, npot, rho, prec, ierr)
  if (ierr /= 0) then
     print '(a)', trim(bdft_error_message(ctx))
     stop 'density failed'
  end if
  print '(a,es25.17)', 'rho_up_pot ', sum(rho(:, 1) * pot(1:npot))
  print '(a,es25.17)', 'rho_dn_sum ', sum(rho(:, 2))
  call free_gpu_hip(ctx, lr)
end program locham_hip_driver
