! Fortran host making TurboGAP's own calls (src/turbogap.f90:432, 944-966,
! 1281-1373) with the reference's argument lists: gpu_context_init(params,
! rank, n_omp), build_neighbors_list, then per soap_turbo block and per batch
! of sites soap_backend_begin(hypers), get_gap_soap (66 arguments),
! soap_backend_end, with the per-batch energies and forces zeroed, assigned
! by the call and added to the totals as the driver does. On a one-block
! system it also calls get_soap on the whole list (every atom a centre).
!
! Input: positions.dat (n, L, then n lines "x y z species_index name"),
! blocks.dat ("n_blocks n_batches rcut rcut_soft zeta radial_enhancement",
! then per block "alphas_file desc_file central_species n_species
! compress_file" and a line of species names; compress_file "none" or a
! file_compress_soap file read as src/read_files.f90:5166-5211 reads it, with
! the descriptor size taken from it). Output: out.dat
! ("energy E", "virial" 9 values, then "force i fx fy fz" lines, and with one
! block "soap_sum S" and "soap_der_abs_sum A").
program tgx_ref_example
   use tgx_types
   use tgx_turbogap
   implicit none
   type(input_parameters) :: params
   type(soap_turbo), allocatable, target :: hyp(:)
   type(local_property_soap_turbo), allocatable :: lpm(:)
   integer :: n, i, k, u, n_omp, n_pairs, nb, nbat, bt, ib, ie, jb, je, cs, nsp, ns, dim, ios
   real(dp) :: L, a(3), b(3), c(3), virial(3, 3), this_virial(3, 3), e_tot, junk, rc, rs, zeta
   integer :: re
   real(dp), allocatable :: pos(:, :), rjs(:), thetas(:), phis(:), xyz(:, :), e_soap(:), f_soap(:, :)
   real(dp), allocatable :: this_e(:), this_f(:, :), soap(:, :), sder(:, :, :), lp0(:, :), lpd0(:, :, :)
   real(dp), allocatable :: dip0(:, :), edip0(:), s_out(:, :), d_out(:, :, :)
   integer, allocatable :: spec(:), n_neigh(:), nlist(:), nspec(:), lpi(:), der_nb(:), der_nbl(:), joff(:)
   integer, allocatable :: species(:, :), mult(:)
   logical, allocatable :: do_list(:), mask(:, :)
   character*8, allocatable :: names(:), bnames(:)
   character(len=256) :: fa, fd, fc

   open (newunit=u, file="positions.dat", status="old")
   read (u, *) n, L
   allocate (pos(3, n), spec(n), names(n))
   do i = 1, n
      read (u, *) pos(:, i), spec(i), names(i)
   end do
   close (u)
   a = [L, 0.0_dp, 0.0_dp]; b = [0.0_dp, L, 0.0_dp]; c = [0.0_dp, 0.0_dp, L]
   call gpu_context_init(params, 0, n_omp)
   allocate (do_list(n))
   do_list = .true.
   open (newunit=u, file="blocks.dat", status="old")
   read (u, *) nb, nbat, rc, rs, zeta, re
   call build_neighbors_list(pos, a, b, c, .false., spec, rc, n_pairs, rjs, thetas, phis, xyz, n_neigh, nlist, &
                             nspec, n, [1, 1, 1], .true., do_list, 0)
   allocate (joff(n + 1))
   joff(1) = 0
   do i = 1, n
      joff(i + 1) = joff(i) + n_neigh(i)
   end do

   allocate (hyp(nb))
   do k = 1, nb
      read (u, *) fa, fd, cs, nsp, fc
      allocate (bnames(nsp))
      read (u, *) bnames
      hyp(k)%n_species = nsp
      hyp(k)%species_types = bnames
      hyp(k)%alpha_max = [(8, i = 1, nsp)]
      hyp(k)%l_max = 8
      hyp(k)%rcut_hard = [(rc, i = 1, nsp)]
      hyp(k)%rcut_soft = [(rs, i = 1, nsp)]
      hyp(k)%nf = [(4.0_dp, i = 1, nsp)]
      hyp(k)%global_scaling = [(1.0_dp, i = 1, nsp)]
      hyp(k)%atom_sigma_r = [(0.5_dp, i = 1, nsp)]
      hyp(k)%atom_sigma_r_scaling = [(0.0_dp, i = 1, nsp)]
      hyp(k)%atom_sigma_t = [(0.5_dp, i = 1, nsp)]
      hyp(k)%atom_sigma_t_scaling = [(0.0_dp, i = 1, nsp)]
      hyp(k)%amplitude_scaling = [(1.0_dp, i = 1, nsp)]
      hyp(k)%central_weight = [(1.0_dp, i = 1, nsp)]
      hyp(k)%basis = "poly3gauss"
      hyp(k)%radial_enhancement = re
      hyp(k)%delta = 1.0_dp
      hyp(k)%zeta = zeta
      hyp(k)%central_species = cs
      hyp(k)%dim = (8*nsp)*(8*nsp + 1)/2*9
      if (trim(fc) == "none") then
         allocate (hyp(k)%compress_P_i(0), hyp(k)%compress_P_j(0), hyp(k)%compress_P_el(0))
      else
         call read_compress(trim(fc), nsp, hyp(k))
      end if
      ! read_alphas_and_descriptors (src/read_files.f90:660-724)
      call read_sparse(trim(fa), trim(fd), hyp(k))
      deallocate (bnames)
   end do
   close (u)

   allocate (e_soap(n), f_soap(3, n), this_e(n), this_f(3, n), lp0(n, 0), lpd0(3, n_pairs, 0), lpi(0))
   allocate (dip0(3, n), edip0(n))
   e_soap = 0.0_dp
   f_soap = 0.0_dp
   virial = 0.0_dp
   do k = 1, nb
      do bt = 1, nbat
         ib = 1 + ((bt - 1)*n)/nbat
         ie = (bt*n)/nbat
         jb = joff(ib) + 1
         je = joff(ie + 1)
         this_e = 0.0_dp
         this_f = 0.0_dp
         call soap_backend_begin(hyp(k))
         call get_gap_soap(n, ie - ib + 1, n_neigh(ib:ie), nlist(jb:je), hyp(k)%n_species, &
                           hyp(k)%species_types, rjs(jb:je), thetas(jb:je), phis(jb:je), xyz(1:3, jb:je), &
                           hyp(k)%alpha_max, hyp(k)%l_max, hyp(k)%dim, hyp(k)%rcut_hard, hyp(k)%rcut_soft, &
                           hyp(k)%nf, hyp(k)%global_scaling, hyp(k)%atom_sigma_r, hyp(k)%atom_sigma_r_scaling, &
                           hyp(k)%atom_sigma_t, hyp(k)%atom_sigma_t_scaling, hyp(k)%amplitude_scaling, &
                           hyp(k)%radial_enhancement, hyp(k)%central_weight, hyp(k)%basis, hyp(k)%scaling_mode, &
                           .false., .true., .true., .true., .false., .false., hyp(k)%compress_soap, &
                           hyp(k)%compress_P_nonzero, hyp(k)%compress_P_i, hyp(k)%compress_P_j, &
                           hyp(k)%compress_P_el, hyp(k)%delta, hyp(k)%zeta, hyp(k)%central_species, names(ib:ie), &
                           names, .true., 0, [1, 1, 1], soap, sder, der_nb, der_nbl, hyp(k)%has_local_properties, &
                           hyp(k)%n_local_properties, lpm, this_e, this_f, lp0, lpd0, lpi, ib, ie, jb, je, &
                           this_virial, 0, hyp(k)%is_dipole_model, dip0, edip0)
         call soap_backend_end()
         e_soap = e_soap + this_e
         f_soap = f_soap + this_f
         virial = virial + this_virial
      end do
   end do
   e_tot = sum(e_soap)
   open (newunit=u, file="out.dat", status="replace")
   write (u, '(A,ES25.17)') 'energy ', e_tot
   write (u, '(A,9ES25.17)') 'virial ', virial
   do i = 1, n
      write (u, '(A,I0,3ES25.17)') 'force ', i, f_soap(:, i)
   end do
   if (nb == 1) then
      ! get_soap on the marshalled list: every atom a centre, one species
      dim = hyp(1)%dim
      allocate (species(1, n), mult(n), mask(n_pairs, 1), s_out(dim, n), d_out(3, dim, n_pairs))
      species = 1
      mult = 1
      mask = .true.
      call soap_backend_begin(hyp(1))
      call get_soap(n, n_neigh, 1, species, mult, n_pairs, mask, rjs, thetas, phis, hyp(1)%alpha_max, &
                    hyp(1)%l_max, hyp(1)%rcut_hard, hyp(1)%rcut_soft, hyp(1)%nf, hyp(1)%global_scaling, &
                    hyp(1)%atom_sigma_r, hyp(1)%atom_sigma_r_scaling, hyp(1)%atom_sigma_t, &
                    hyp(1)%atom_sigma_t_scaling, hyp(1)%amplitude_scaling, hyp(1)%radial_enhancement, &
                    hyp(1)%central_weight, hyp(1)%basis, hyp(1)%scaling_mode, .false., .true., &
                    hyp(1)%compress_soap, hyp(1)%compress_P_nonzero, hyp(1)%compress_P_i, hyp(1)%compress_P_j, &
                    hyp(1)%compress_P_el, s_out, d_out)
      call soap_backend_end()
      write (u, '(A,ES25.17)') 'soap_sum ', sum(s_out)
      write (u, '(A,ES25.17)') 'soap_der_abs_sum ', sum(abs(d_out))
   end if
   close (u)
   call gpu_context_finalize(params, n_omp)

contains

   ! src/read_files.f90:5166-5211: header (alpha_max per species, l_max, dim),
   ! then a P_transformation block or one kept component per line
   subroutine read_compress(fname, nsp, h)
      character(len=*), intent(in) :: fname
      integer, intent(in) :: nsp
      type(soap_turbo), intent(inout) :: h
      integer :: uu, ij, m, nnz
      character(len=256) :: line
      open (newunit=uu, file=fname, status="old")
      read (uu, *) (ij, m=1, nsp), ij, h%dim
      read (uu, '(A)') line
      h%compress_soap = .true.
      if (trim(line) == "P_transformation") then
         nnz = 0
         do
            read (uu, '(A)') line
            if (trim(line) == "end_transformation") exit
            nnz = nnz + 1
         end do
         rewind (uu)
         read (uu, *)
         read (uu, *)
         allocate (h%compress_P_i(nnz), h%compress_P_j(nnz), h%compress_P_el(nnz))
         do m = 1, nnz
            read (uu, *) h%compress_P_i(m), h%compress_P_j(m), h%compress_P_el(m)
         end do
      else
         backspace (uu)
         nnz = h%dim
         allocate (h%compress_P_i(nnz), h%compress_P_j(nnz), h%compress_P_el(nnz))
         do m = 1, nnz
            read (uu, *) h%compress_P_j(m)
            h%compress_P_i(m) = m
            h%compress_P_el(m) = 1.0_dp
         end do
      end if
      h%compress_P_nonzero = nnz
      close (uu)
   end subroutine

   subroutine read_sparse(file_alphas, file_desc, h)
      character(len=*), intent(in) :: file_alphas, file_desc
      type(soap_turbo), intent(inout) :: h
      integer :: uu, m, j, cnt
      open (newunit=uu, file=file_alphas, status="old")
      cnt = 0
      do
         read (uu, *, iostat=ios) junk
         if (ios /= 0) exit
         cnt = cnt + 1
      end do
      rewind (uu)
      allocate (h%alphas(cnt))
      do m = 1, cnt
         read (uu, *) h%alphas(m)
      end do
      close (uu)
      h%n_sparse = cnt
      allocate (h%Qs(h%dim, cnt))
      open (newunit=uu, file=file_desc, status="old")
      do m = 1, cnt
         do j = 1, h%dim
            read (uu, *) h%Qs(j, m)
         end do
      end do
      close (uu)
   end subroutine
end program
