! Example Fortran host: the TurboGAP call sequence (turbogap.f90:944-966,
! 1314-1344) driven through tgx_turbogap. Reads positions.dat (n, L, then
! n lines of x y z species), one soap_turbo block from alphas.dat /
! desc_sparse.dat, and prints the energy, the first force and the virial trace.
program tgx_example
   use tgx_turbogap
   implicit none
   integer :: n, i, n_pairs, mid, u
   real(dp) :: L, virial(3, 3), a(3), b(3), c(3), e_tot
   real(dp), allocatable :: pos(:, :), rjs(:), thetas(:), phis(:), xyz(:, :), energies(:), forces(:, :)
   integer, allocatable :: spec(:), n_neigh(:), nlist(:), nspec(:)
   logical, allocatable :: do_list(:)
   open (newunit=u, file="positions.dat", status="old")
   read (u, *) n, L
   allocate (pos(3, n), spec(n), energies(n), forces(3, n))
   do i = 1, n
      read (u, *) pos(:, i), spec(i)
   end do
   close (u)
   a = [L, 0.0_dp, 0.0_dp]; b = [0.0_dp, L, 0.0_dp]; c = [0.0_dp, 0.0_dp, L]
   call gpu_context_init(0)
   allocate (do_list(n))
   do_list = .true.
   call build_neighbors_list(pos, a, b, c, .false., spec, 3.7_dp, n_pairs, rjs, thetas, phis, xyz, n_neigh, nlist, &
                             nspec, n, [1, 1, 1], .true., do_list, 0)
   call soap_model_from_files("alphas.dat", "desc_sparse.dat", [6], [8], 8, [3.7_dp], [3.2_dp], [0.5_dp], &
                              [0.5_dp], [1.0_dp], "poly3gauss", 0, 1.0_dp, 4.0_dp, mid)
   energies = 0.0_dp
   forces = 0.0_dp
   call soap_backend_begin(mid)
   call get_gap_soap(TGX_PATH_FP64_PLAIN, .true., energies, forces, virial)
   call soap_backend_end()
   e_tot = sum(energies)
   write (*, '(A,I0)') 'pairs ', n_pairs
   write (*, '(A,ES24.16)') 'energy ', e_tot
   write (*, '(A,3ES24.16)') 'force1 ', forces(:, 1)
   write (*, '(A,ES24.16)') 'virial_trace ', virial(1, 1) + virial(2, 2) + virial(3, 3)
   call gpu_context_finalize()
end program
