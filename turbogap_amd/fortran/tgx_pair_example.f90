! Example Fortran host for the pair and triplet families: the gap_backend
! sequence (src/gap_backend_cpu.f90:79-267) driven through tgx_turbogap. Reads
! positions.dat (n, L, then n lines of x y z species) and one block of each
! family (2b.dat, 3b.dat, corepot.dat, formats below), builds the list at the
! largest cutoff and writes out_<family>.dat: the virial (9 values), then one
! line per atom with its energy and force.
!   2b.dat       species1 species2 rcut delta sigma n_sparse, then n_sparse lines Q alpha cutoff
!   3b.dat       species_center species1 species2 rcut delta sigma(3) kernel_type n_sparse,
!                then n_sparse lines Q1 Q2 Q3 alpha cutoff
!   corepot.dat  species1 species2 n yp1 ypn, then n lines x V
program tgx_pair_example
   use tgx_turbogap
   implicit none
   integer :: n, i, n_pairs, u, ns, s1, s2, sc, id(3), k
   real(dp) :: L, rc2, rc3, rcp, delta, sigma, sig3(3), yp1, ypn, virial(3, 3), a(3), b(3), c(3)
   real(dp), allocatable :: pos(:, :), rjs(:), thetas(:), phis(:), xyz(:, :), energies(:), forces(:, :)
   real(dp), allocatable :: q(:), q3(:, :), al(:), cut(:), x(:), v(:)
   integer, allocatable :: spec(:), n_neigh(:), nlist(:), nspec(:)
   logical, allocatable :: do_list(:)
   character(len=8) :: kt
   character(len=16), parameter :: names(3) = [character(len=16) :: "distance_2b", "angle_3b", "core_pot"]
   open (newunit=u, file="positions.dat", status="old")
   read (u, *) n, L
   allocate (pos(3, n), spec(n), energies(n), forces(3, n))
   do i = 1, n
      read (u, *) pos(:, i), spec(i)
   end do
   close (u)
   a = [L, 0.0_dp, 0.0_dp]; b = [0.0_dp, L, 0.0_dp]; c = [0.0_dp, 0.0_dp, L]
   call gpu_context_init(0)
   ! the blocks first: the list is built at the largest cutoff
   open (newunit=u, file="2b.dat", status="old")
   read (u, *) s1, s2, rc2, delta, sigma, ns
   allocate (q(ns), al(ns), cut(ns))
   do i = 1, ns
      read (u, *) q(i), al(i), cut(i)
   end do
   close (u)
   id(1) = tgx_distance_2b_model(s1, s2, rc2, delta, sigma, q, al, cut)
   deallocate (al, cut)
   open (newunit=u, file="3b.dat", status="old")
   read (u, *) sc, s1, s2, rc3, delta, sig3, kt, ns
   allocate (q3(ns, 3), al(ns), cut(ns))
   do i = 1, ns
      read (u, *) q3(i, :), al(i), cut(i)
   end do
   close (u)
   id(2) = tgx_angle_3b_model(sc, s1, s2, rc3, delta, sig3, trim(kt), q3, al, cut)
   open (newunit=u, file="corepot.dat", status="old")
   read (u, *) s1, s2, ns, yp1, ypn
   allocate (x(ns), v(ns))
   do i = 1, ns
      read (u, *) x(i), v(i)
   end do
   close (u)
   id(3) = tgx_core_pot_model(s1, s2, x, v, yp1, ypn, 1.0e10_dp, 1.0_dp)
   rcp = maxval(x)
   allocate (do_list(n))
   do_list = .true.
   call build_neighbors_list(pos, a, b, c, .false., spec, max(rc2, rc3, rcp), n_pairs, rjs, thetas, phis, xyz, &
                             n_neigh, nlist, nspec, n, [1, 1, 1], .true., do_list, 0)
   do k = 1, 3
      energies = 0.0_dp
      forces = 0.0_dp
      virial = 0.0_dp
      call add_pair_contribution(id(k), energies, forces, virial)
      open (newunit=u, file="out_"//trim(names(k))//".dat", status="replace")
      write (u, '(9ES25.16E3)') virial
      do i = 1, n
         write (u, '(4ES25.16E3)') energies(i), forces(:, i)
      end do
      close (u)
   end do
   write (*, '(A,I0)') 'pairs ', n_pairs
   call gpu_context_finalize()
end program
