! Example Fortran host of the device MD loop (src/turbogap_md.f90:160-470):
! the carbon box of positions.dat with one soap_turbo block from alphas.dat /
! desc_sparse.dat, velocities and masses from md_state.dat (n lines of
! vx vy vz m), then n_steps NVT steps (Berendsen, T0 = 600 K, tau = 50 fs,
! dt = 1 fs): list, forces, remove_cm_vel, velocity Verlet, E_kin and T,
! pressure, thermostat, wrap. Prints each step's energy and temperature and the
! final positions and velocities.
program tgx_md_example
   use tgx_turbogap
   implicit none
   integer :: n, i, mid, u, step
   integer, parameter :: n_steps = 5
   real(dp) :: L, virial(3, 3), a(3), b(3), c(3), energy, ek, temp, p, pt(3, 3)
   real(dp), allocatable :: pos(:, :), vel(:, :), mass(:), rjs(:), thetas(:), phis(:), xyz(:, :)
   integer, allocatable :: spec(:), n_neigh(:), nlist(:), nspec(:)
   logical, allocatable :: do_list(:)
   integer :: n_pairs
   open (newunit=u, file="positions.dat", status="old")
   read (u, *) n, L
   allocate (pos(3, n), spec(n), vel(3, n), mass(n))
   do i = 1, n
      read (u, *) pos(:, i), spec(i)
   end do
   close (u)
   open (newunit=u, file="md_state.dat", status="old")
   do i = 1, n
      read (u, *) vel(:, i), mass(i)
   end do
   close (u)
   a = [L, 0.0_dp, 0.0_dp]; b = [0.0_dp, L, 0.0_dp]; c = [0.0_dp, 0.0_dp, L]
   call gpu_context_init(0)
   allocate (do_list(n))
   do_list = .true.
   ! the system and its first list (host copies unused)
   call build_neighbors_list(pos, a, b, c, .false., spec, 3.7_dp, n_pairs, rjs, thetas, phis, xyz, n_neigh, nlist, &
                             nspec, n, [1, 1, 1], .true., do_list, 0)
   call soap_model_from_files("alphas.dat", "desc_sparse.dat", [6], [8], 8, [3.7_dp], [3.2_dp], [0.5_dp], &
                              [0.5_dp], [1.0_dp], "poly3gauss", 0, 1.0_dp, 4.0_dp, mid)
   call md_state_init(vel, mass)
   do step = 0, n_steps - 1
      call md_build_list(3.7_dp)
      call md_forces(TGX_PATH_FP64_ADJOINT, energy, virial)
      call remove_cm_vel_device()
      call velocity_verlet_device(1.0_dp, step == 0)
      call md_kinetic(ek, temp)
      call md_pressure(virial, temp, p, pt)
      call berendsen_thermostat_device(600.0_dp, temp, 50.0_dp, 1.0_dp)
      call wrap_pbc_device()
      write (*, '(A,I0,3ES25.17)') 'step ', step, energy, temp, p
   end do
   call md_state_get(pos, vel)
   do i = 1, n
      write (*, '(A,6ES25.17)') 'atom ', pos(:, i), vel(:, i)
   end do
   call gpu_context_finalize()
end program
