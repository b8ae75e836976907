! tgx_turbogap: TurboGAP's SOAP/GAP call surface over the libtgx C ABI.
!
! A Fortran host (TurboGAP's driver) keeps calling the routines by their
! reference names; each forwards to libtgx (include/tgx.h) through
! iso_c_binding. Device memory, streams and kernels stay behind the C ABI: the
! Fortran side never holds a device pointer.
!
!   gpu_context_init(params, rank, n_omp) /
!   gpu_context_finalize(params, n_omp)       src/gpu_context.f90:41-58 (also (device) / ())
!   build_neighbors_list                      src/neighbors.f90:243-605
!   number_of_unit_cells_for_given_cutoff     src/neighbors.f90:175-236
!   build_supercell                           src/read_files.f90:320-367 (the replication in read_xyz)
!   soap_model_from_files                     src/read_files.f90:660-724 (alphas / desc_sparse)
!   soap_backend_begin(hypers) / _end         src/gap.f90:49-64 (also (model_id))
!   get_gap_soap                              src/gap_interface.f90:41-480, the reference's 66
!                                             arguments (also (path, do_forces, energies0,
!                                             forces0, virial) on the resident list and block)
!   get_soap                                  soap_turbo get_soap, call shape src/gap_interface.f90:320-325
!   get_soap_energy_and_forces                src/gap.f90:66-259
!   tgx_distance_2b_model / tgx_angle_3b_model / tgx_core_pot_model and
!   add_pair_contribution                     the gap_backend seam's add_2b_ /
!                                             add_3b_ / add_core_pot_contribution
!                                             (src/gap_backend_cpu.f90:79-267)
!
!   The device MD loop of src/turbogap_md.f90:160-470 on resident state
!   (positions, velocities and forces stay on the device between steps):
!   md_state_init / md_state_get, md_build_list / md_refresh_list, md_forces,
!   wrap_pbc_device, remove_cm_vel_device, variable_time_step_device,
!   variable_time_step_adaptive_device (src/stopping/adaptive_time.f90:46-164),
!   velocity_verlet_device, gradient_descent_device, md_kinetic, md_pressure,
!   box_scaling_device, berendsen_barostat_device, berendsen_thermostat_device,
!   bussi_thermostat_device, md_displacement       src/md.f90:95-1128
!
! Errors: a non-zero status prints tgx_last_error() and stops with code 1, the
! behaviour of turbogap_abort (src/error.f90:48-61).
module tgx_turbogap
   use iso_c_binding
   use tgx_types
   implicit none
   private

   integer, parameter, public :: dp = c_double
   integer(c_int), parameter, public :: TGX_PATH_FP64_PLAIN = 0, TGX_PATH_FP64_ADJOINT = 1, &
                                        TGX_PATH_MIXED = 2, TGX_PATH_SURROGATE = 3

   type, bind(C), public :: tgx_soap_hypers
      integer(c_int) :: n_species, l_max, basis, radial_enhancement, n_quad_soft, n_quad_buffer
      type(c_ptr) :: alpha_max, rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r, &
                     atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, amplitude_scaling, central_weight
   end type

   type(c_ptr), save :: ctx = c_null_ptr

   ! The reference surface's state: the list build_neighbors_list last made
   ! resident (checked against the lists get_gap_soap is handed), species names
   ! to the integer codes of that build, the hypers of soap_backend_begin, and
   ! the blocks created for them (keyed by the address of their sparse set).
   integer, allocatable, save :: res_species(:), res_n_neigh(:), res_nlist(:)
   integer, save :: res_n_sites = 0
   character(len=8), save :: sp_name(64)
   integer, save :: sp_code(64), n_sp = 0
   type(soap_turbo), pointer, save :: cur_hypers => null()
   type(c_ptr), save :: reg_key(128), reg_desc_key(128)
   integer, save :: reg_id(128), reg_desc_id(128), n_reg = 0, n_reg_desc = 0
   integer(c_int), save :: gap_path = 0

   interface gpu_context_init
      module procedure gpu_context_init_device, gpu_context_init_params
   end interface
   interface gpu_context_finalize
      module procedure gpu_context_finalize_default, gpu_context_finalize_params
   end interface
   interface soap_backend_begin
      module procedure soap_backend_begin_id, soap_backend_begin_hypers
   end interface
   interface get_gap_soap
      module procedure get_gap_soap_resident, get_gap_soap_ref
   end interface

   public :: gpu_context_init, gpu_context_finalize, build_neighbors_list, soap_model_from_files, &
             get_soap, tgx_set_gap_path, tgx_set_memory_budget, &
             number_of_unit_cells_for_given_cutoff, build_supercell, &
             soap_backend_begin, soap_backend_end, get_gap_soap, get_soap_energy_and_forces, tgx_check, &
             tgx_distance_2b_model, tgx_angle_3b_model, tgx_core_pot_model, add_pair_contribution, &
             md_state_init, md_state_get, md_build_list, md_refresh_list, md_forces, wrap_pbc_device, &
             remove_cm_vel_device, variable_time_step_device, variable_time_step_adaptive_device, &
             velocity_verlet_device, gradient_descent_device, &
             md_kinetic, md_pressure, box_scaling_device, berendsen_barostat_device, berendsen_thermostat_device, &
             bussi_thermostat_device, md_displacement

   interface
      integer(c_int) function c_tgx_init(device, ctx_out) bind(C, name="tgx_init")
         import :: c_int, c_ptr
         integer(c_int), value :: device
         type(c_ptr) :: ctx_out
      end function
      integer(c_int) function c_tgx_finalize(c) bind(C, name="tgx_finalize")
         import :: c_int, c_ptr
         type(c_ptr), value :: c
      end function
      type(c_ptr) function c_tgx_last_error() bind(C, name="tgx_last_error")
         import :: c_ptr
      end function
      integer(c_int) function c_tgx_set_system(c, n, nsc, pos, spec, box) bind(C, name="tgx_set_system")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         integer(c_int), value :: n, nsc
         real(c_double) :: pos(*), box(*)
         integer(c_int) :: spec(*)
      end function
      integer(c_int) function c_tgx_build(c, rcut, indices, do_list, npairs) bind(C, name="tgx_build_neighbors_list")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: c, do_list  ! do_list: int[n_sites], or null for every site
         real(c_double), value :: rcut
         integer(c_int) :: indices(*)
         integer(c_int64_t) :: npairs
      end function
      integer(c_int) function c_tgx_set_positions(c, pos) bind(C, name="tgx_set_positions")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double) :: pos(*)
      end function
      integer(c_int) function c_tgx_refresh(c) bind(C, name="tgx_refresh_neighbors_list")
         import :: c_int, c_ptr
         type(c_ptr), value :: c
      end function
      integer(c_int) function c_tgx_get_nl(c, nn, nl, rjs, th, ph, xyz, ns) bind(C, name="tgx_get_neighbors_list")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         integer(c_int) :: nn(*), nl(*), ns(*)
         real(c_double) :: rjs(*), th(*), ph(*), xyz(*)
      end function
      integer(c_int) function c_tgx_model(c, h, st, cs, delta, zeta, e0, ns, dim, qs, al, mid) &
         bind(C, name="tgx_soap_model_create")
         import :: c_int, c_ptr, c_double, tgx_soap_hypers
         type(c_ptr), value :: c
         type(tgx_soap_hypers) :: h
         integer(c_int) :: st(*)
         integer(c_int), value :: cs, ns, dim
         real(c_double), value :: delta, zeta, e0
         real(c_double) :: qs(*), al(*)
         integer(c_int) :: mid
      end function
      integer(c_int) function c_tgx_model_c(c, h, st, cs, delta, zeta, e0, ns, dim, qs, al, nnz, pi, pj, pel, mid) &
         bind(C, name="tgx_soap_model_create_compressed")
         import :: c_int, c_ptr, c_double, tgx_soap_hypers
         type(c_ptr), value :: c
         type(tgx_soap_hypers) :: h
         integer(c_int) :: st(*), pi(*), pj(*)
         integer(c_int), value :: cs, ns, dim, nnz
         real(c_double), value :: delta, zeta, e0
         real(c_double) :: qs(*), al(*), pel(*)
         integer(c_int) :: mid
      end function
      integer(c_int) function c_tgx_begin(c, mid) bind(C, name="tgx_soap_backend_begin")
         import :: c_int, c_ptr
         type(c_ptr), value :: c
         integer(c_int), value :: mid
      end function
      integer(c_int) function c_tgx_end(c) bind(C, name="tgx_soap_backend_end")
         import :: c_int, c_ptr
         type(c_ptr), value :: c
      end function
      integer(c_int) function c_tgx_gap_soap(c, path, do_forces, en, f, vir, nsb, npb) bind(C, name="tgx_get_gap_soap")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: c
         integer(c_int), value :: path, do_forces
         real(c_double) :: en(*), f(*), vir(*)
         integer(c_int) :: nsb
         integer(c_int64_t) :: npb
      end function
      integer(c_int) function c_tgx_eaf(c, n, nsoap, np, soap, sder, delta, zeta0, e0, nn, nl, xyz, n0, dof, en, &
                                        f, vir) bind(C, name="tgx_get_soap_energy_and_forces")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: c
         integer(c_int), value :: n, nsoap, n0, dof
         integer(c_int64_t), value :: np
         real(c_double), value :: delta, zeta0, e0
         real(c_double) :: soap(*), sder(*), xyz(*), en(*), f(*), vir(*)
         integer(c_int) :: nn(*), nl(*)
      end function
      integer(c_int) function c_tgx_2b(c, s1, s2, rcut, buffer, delta, sigma, e0, ns, qs, al, cut, mid) &
         bind(C, name="tgx_distance_2b_create")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         integer(c_int), value :: s1, s2, ns
         real(c_double), value :: rcut, buffer, delta, sigma, e0
         real(c_double) :: qs(*), al(*), cut(*)
         integer(c_int) :: mid
      end function
      integer(c_int) function c_tgx_3b(c, sc, s1, s2, rcut, buffer, delta, sigma, e0, kt, ns, qs, al, cut, mid) &
         bind(C, name="tgx_angle_3b_create")
         import :: c_int, c_ptr, c_double, c_char
         type(c_ptr), value :: c
         integer(c_int), value :: sc, s1, s2, ns
         real(c_double), value :: rcut, buffer, delta, e0
         real(c_double) :: sigma(*), qs(*), al(*), cut(*)
         character(kind=c_char) :: kt(*)
         integer(c_int) :: mid
      end function
      integer(c_int) function c_tgx_core_pot(c, s1, s2, n, x, v, yp1, ypn, cut, buf, mid) bind(C, name="tgx_core_pot_create")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         integer(c_int), value :: s1, s2, n
         real(c_double), value :: yp1, ypn, cut, buf
         real(c_double) :: x(*), v(*)
         integer(c_int) :: mid
      end function
      integer(c_int) function c_tgx_pair_eaf(c, mid, en, f, vir) bind(C, name="tgx_pair_energy_and_forces")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         integer(c_int), value :: mid
         real(c_double) :: en(*), f(*), vir(*)
      end function
      integer(c_int) function c_md_set_state(c, v, m) bind(C, name="tgx_md_set_state")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double), intent(in) :: v(*), m(*)
      end function
      integer(c_int) function c_md_set_fixed(c, f) bind(C, name="tgx_md_set_fixed")
         import :: c_int, c_ptr
         type(c_ptr), value :: c
         integer(c_int), intent(in) :: f(*)
      end function
      integer(c_int) function c_md_get_state(c, x, v) bind(C, name="tgx_md_get_state")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double), intent(out) :: x(*), v(*)
      end function
      integer(c_int) function c_compute(c, path, e, vir) bind(C, name="tgx_compute")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         integer(c_int), value :: path
         real(c_double), intent(out) :: e, vir(*)
      end function
      integer(c_int) function c_wrap(c) bind(C, name="tgx_wrap_pbc")
         import :: c_int, c_ptr
         type(c_ptr), value :: c
      end function
      integer(c_int) function c_remove_cm(c) bind(C, name="tgx_remove_cm_vel")
         import :: c_int, c_ptr
         type(c_ptr), value :: c
      end function
      integer(c_int) function c_vdt(c, init, target, tau, dt0, dt) bind(C, name="tgx_variable_time_step")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         integer(c_int), value :: init
         real(c_double), value :: target, tau, dt0
         real(c_double), intent(inout) :: dt
      end function
      integer(c_int) function c_adt(c, init, tmin, tmax, xmax, emax, dt0, dt) bind(C, name="tgx_adaptive_time_step")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         integer(c_int), value :: init
         real(c_double), value :: tmin, tmax, xmax, emax, dt0
         real(c_double), intent(inout) :: dt
      end function
      integer(c_int) function c_vv(c, dt, first) bind(C, name="tgx_velocity_verlet")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double), value :: dt
         integer(c_int), value :: first
      end function
      integer(c_int) function c_gd(c, e, step, first, gamma, restored) bind(C, name="tgx_gradient_descent")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double), value :: e, step
         integer(c_int), value :: first
         real(c_double), intent(out) :: gamma
         integer(c_int), intent(out) :: restored
      end function
      integer(c_int) function c_kinetic(c, ek, t) bind(C, name="tgx_kinetic")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double), intent(out) :: ek, t
      end function
      integer(c_int) function c_pressure(c, vir, t, p, pt) bind(C, name="tgx_md_pressure")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double), intent(in) :: vir(*)
         real(c_double), value :: t
         real(c_double), intent(out) :: p, pt(*)
      end function
      integer(c_int) function c_box_scaling(c, g, istep, nsteps, box) bind(C, name="tgx_box_scaling")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double), intent(in) :: g(*)
         integer(c_int), value :: istep, nsteps
         real(c_double), intent(out) :: box(*)
      end function
      integer(c_int) function c_barostat(c, p0, p, sym, tau, gamma, dt, box) bind(C, name="tgx_berendsen_barostat")
         import :: c_int, c_ptr, c_double, c_char
         type(c_ptr), value :: c
         real(c_double), value :: p0, tau, gamma, dt
         real(c_double), intent(in) :: p(*)
         character(kind=c_char), intent(in) :: sym(*)
         real(c_double), intent(out) :: box(*)
      end function
      integer(c_int) function c_berendsen(c, t0, t, tau, dt) bind(C, name="tgx_berendsen")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double), value :: t0, t, tau, dt
      end function
      integer(c_int) function c_bussi(c, t0, tau, dt, g, chi2, ek) bind(C, name="tgx_bussi_thermostat")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double), value :: t0, tau, dt, g, chi2
         real(c_double), intent(out) :: ek
      end function
      integer(c_int) function c_displacement(c, buffer, reset, rebuild) bind(C, name="tgx_md_displacement")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double), value :: buffer
         integer(c_int), value :: reset
         integer(c_int), intent(out) :: rebuild
      end function
      integer(c_int) function c_unit_cells(box, rcut, ind) bind(C, name="tgx_number_of_unit_cells")
         import :: c_int, c_double
         real(c_double) :: box(*)
         real(c_double), value :: rcut
         integer(c_int) :: ind(*)
      end function
      integer(c_int) function c_supercell(n, pos, spec, box, ind, psc, ssc, bsc) bind(C, name="tgx_build_supercell")
         import :: c_int, c_double
         integer(c_int), value :: n
         real(c_double) :: pos(*), box(*), psc(*), bsc(*)
         integer(c_int) :: spec(*), ind(*), ssc(*)
      end function
      integer(c_int) function c_tgx_gap_soap_sites(c, path, do_forces, ib, ie, en, f, vir, nsb, npb) &
         bind(C, name="tgx_get_gap_soap_sites")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: c
         integer(c_int), value :: path, do_forces, ib, ie
         real(c_double) :: en(*), f(*), vir(*)
         integer(c_int) :: nsb
         integer(c_int64_t) :: npb
      end function
      integer(c_int) function c_tgx_soap_list(c, n, nn, np, ps, rjs, xyz, der, soap, sder) &
         bind(C, name="tgx_get_soap_list")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: c
         integer(c_int), value :: n, der
         integer(c_int64_t), value :: np
         integer(c_int) :: nn(*), ps(*)
         real(c_double) :: rjs(*), xyz(*), soap(*), sder(*)
      end function
      integer(c_int) function c_tgx_device_count(n) bind(C, name="tgx_device_count")
         import :: c_int
         integer(c_int) :: n
      end function
      integer(c_int) function c_tgx_set_budget(c, gib) bind(C, name="tgx_set_memory_budget")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c
         real(c_double), value :: gib
      end function
      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function
   end interface

contains

   subroutine tgx_check(status, what)
      integer(c_int), intent(in) :: status
      character(len=*), intent(in) :: what
      character(kind=c_char), pointer :: msg(:)
      integer :: n
      if (status == 0) return
      n = int(c_strlen(c_tgx_last_error()))
      call c_f_pointer(c_tgx_last_error(), msg, [n])
      write (*, '(A,A,A,I0,A,*(A))') 'ERROR in ', what, ' (status ', status, '): ', msg(1:n)
      error stop 1
   end subroutine

   subroutine gpu_context_init_device(device)
      integer, intent(in) :: device
      call tgx_check(c_tgx_init(int(device, c_int), ctx), "gpu_context_init")
   end subroutine

   ! gpu_context_init with the reference's arguments (src/gpu_context.f90:41-48):
   ! rank r binds device mod(r, number of devices); one host thread per rank.
   subroutine gpu_context_init_params(params, rank, n_omp)
      type(input_parameters), intent(in) :: params
      integer, intent(in) :: rank
      integer, intent(out) :: n_omp
      integer(c_int) :: nd
      call tgx_check(c_tgx_device_count(nd), "gpu_context_init")
      if (nd < 1) then
         write (*, '(A)') 'ERROR in gpu_context_init: no HIP device'
         error stop 1
      end if
      call tgx_check(c_tgx_init(int(mod(rank, nd), c_int), ctx), "gpu_context_init")
      n_omp = 1
   end subroutine

   subroutine gpu_context_finalize_default()
      if (c_associated(ctx)) call tgx_check(c_tgx_finalize(ctx), "gpu_context_finalize")
      ctx = c_null_ptr
      n_reg = 0
      n_reg_desc = 0
      nullify (cur_hypers)
   end subroutine

   ! gpu_context_finalize with the reference's arguments (src/gpu_context.f90:52-58).
   subroutine gpu_context_finalize_params(params, n_omp)
      type(input_parameters), intent(in) :: params
      integer, intent(in) :: n_omp
      call gpu_context_finalize_default()
   end subroutine

   ! The device path's choice for get_gap_soap (TGX_PATH_*; FP64 plain by default).
   subroutine tgx_set_gap_path(path)
      integer, intent(in) :: path
      gap_path = int(path, c_int)
   end subroutine

   ! Device memory (GiB) a block of the FP64 plain path may hold before its centres
   ! run in batches (0: 80% of the device).
   subroutine tgx_set_memory_budget(gib)
      real(dp), intent(in) :: gib
      call tgx_check(c_tgx_set_budget(ctx, gib), "tgx_set_memory_budget")
   end subroutine

   ! number_of_unit_cells_for_given_cutoff with the reference's arguments
   ! (src/neighbors.f90:175-236); PBC is accepted and, as there, not used.
   subroutine number_of_unit_cells_for_given_cutoff(a, b, c, rcut, PBC, indices)
      real(dp), intent(in) :: a(1:3), b(1:3), c(1:3), rcut
      logical, intent(in) :: PBC(1:3)
      integer, intent(out) :: indices(1:3)
      real(dp) :: box(9)
      integer(c_int) :: ind(3)
      box(1:3) = a; box(4:6) = b; box(7:9) = c
      call tgx_check(c_unit_cells(box, rcut, ind), "number_of_unit_cells_for_given_cutoff")
      indices = int(ind)
   end subroutine

   ! The supercell replication of read_xyz (src/read_files.f90:320-367):
   ! positions(3, n_sites) and species(n_sites) become positions(3, n_sites_supercell)
   ! and species_supercell, copies in the reference's order (a slowest, the cell
   ! first), and a_box, b_box, c_box are scaled by indices. Nothing changes when
   ! indices is (1, 1, 1) beyond species_supercell = species.
   subroutine build_supercell(positions, species, a_box, b_box, c_box, indices, species_supercell, n_sites_supercell)
      real(dp), allocatable, intent(inout) :: positions(:, :)
      integer, intent(in) :: species(:), indices(1:3)
      real(dp), intent(inout) :: a_box(1:3), b_box(1:3), c_box(1:3)
      integer, allocatable, intent(inout) :: species_supercell(:)
      integer, intent(out) :: n_sites_supercell
      real(dp), allocatable :: psc(:, :)
      integer(c_int), allocatable :: ssc(:), spec(:)
      integer(c_int) :: ind(3)
      real(dp) :: box(9), bsc(9)
      integer :: n
      n = size(positions, 2)
      n_sites_supercell = n*indices(1)*indices(2)*indices(3)
      allocate (psc(3, n_sites_supercell), ssc(n_sites_supercell), spec(n))
      spec = int(species(1:n), c_int); ind = int(indices, c_int)
      box(1:3) = a_box; box(4:6) = b_box; box(7:9) = c_box
      call tgx_check(c_supercell(int(n, c_int), positions, spec, box, ind, psc, ssc, bsc), "build_supercell")
      call move_alloc(psc, positions)
      if (allocated(species_supercell)) deallocate (species_supercell)
      allocate (species_supercell(n_sites_supercell))
      species_supercell = int(ssc)
      a_box = bsc(1:3); b_box = bsc(4:6); c_box = bsc(7:9)
   end subroutine

   ! build_neighbors_list with the reference's 20 arguments in its order
   ! (src/neighbors.f90:243-246) and the same outputs: callee-allocated arrays,
   ! 1-based supercell indices in neighbors_list, xyz(3, n_atom_pairs) = r_j - r_i.
   ! With rebuild_neighbors_list = .false. and a list already returned, the list
   ! keeps its pairs and order and rjs, xyz, thetas, phis are recomputed from
   ! positions into the same arrays (src/neighbors.f90:554-585); the box and
   ! species are those of the last build. do_timing prints the build time; rank
   ! only selects who prints.
   subroutine build_neighbors_list(positions, a_box, b_box, c_box, do_timing, species_supercell, rcut_max, &
                                   n_atom_pairs, rjs, thetas, phis, xyz, n_neigh, neighbors_list, neighbor_species, &
                                   n_sites, indices, rebuild_neighbors_list, do_list, rank)
      real(dp), intent(in) :: positions(:, :), rcut_max
      real(dp), intent(inout) :: a_box(1:3), b_box(1:3), c_box(1:3)
      logical, intent(in) :: do_timing, rebuild_neighbors_list, do_list(:)
      integer, intent(in) :: species_supercell(:), n_sites, indices(1:3), rank
      integer, intent(out) :: n_atom_pairs
      real(dp), allocatable, intent(inout) :: rjs(:), thetas(:), phis(:), xyz(:, :)
      integer, allocatable, intent(inout) :: n_neigh(:), neighbors_list(:), neighbor_species(:)
      real(dp) :: box(9)
      integer(c_int) :: ind(3)
      integer(c_int), allocatable :: spec(:)
      integer(c_int), allocatable, target :: dl(:)
      integer(c_int64_t) :: np
      integer(8) :: t0, t1, rate
      if (do_timing) call system_clock(t0, rate)
      if (.not. rebuild_neighbors_list .and. allocated(n_neigh) .and. allocated(rjs)) then
         call tgx_check(c_tgx_set_positions(ctx, positions), "build_neighbors_list")
         call tgx_check(c_tgx_refresh(ctx), "build_neighbors_list")
         n_atom_pairs = size(rjs)
         call tgx_check(c_tgx_get_nl(ctx, n_neigh, neighbors_list, rjs, thetas, phis, xyz, neighbor_species), &
                        "build_neighbors_list")
         return
      end if
      box(1:3) = a_box; box(4:6) = b_box; box(7:9) = c_box
      ind = int(indices, c_int)
      allocate (spec(size(species_supercell)), dl(n_sites))
      spec = int(species_supercell, c_int)
      dl = merge(1_c_int, 0_c_int, do_list(1:n_sites))
      call tgx_check(c_tgx_set_system(ctx, int(n_sites, c_int), int(size(positions, 2), c_int), positions, spec, box), &
                     "build_neighbors_list")
      call tgx_check(c_tgx_build(ctx, rcut_max, ind, c_loc(dl), np), "build_neighbors_list")
      n_atom_pairs = int(np)
      if (allocated(rjs)) deallocate (rjs)
      if (allocated(thetas)) deallocate (thetas)
      if (allocated(phis)) deallocate (phis)
      if (allocated(xyz)) deallocate (xyz)
      if (allocated(n_neigh)) deallocate (n_neigh)
      if (allocated(neighbors_list)) deallocate (neighbors_list)
      if (allocated(neighbor_species)) deallocate (neighbor_species)
      allocate (rjs(n_atom_pairs), thetas(n_atom_pairs), phis(n_atom_pairs), xyz(3, n_atom_pairs))
      allocate (n_neigh(n_sites), neighbors_list(n_atom_pairs), neighbor_species(n_atom_pairs))
      call tgx_check(c_tgx_get_nl(ctx, n_neigh, neighbors_list, rjs, thetas, phis, xyz, neighbor_species), &
                     "build_neighbors_list")
      res_n_sites = n_sites
      res_species = species_supercell
      res_n_neigh = n_neigh
      res_nlist = neighbors_list
      if (do_timing .and. rank == 0) then
         call system_clock(t1)
         write (*, '(A, F9.3, A)') '  *) Neighbors build: ', real(t1 - t0, dp)/real(rate, dp), ' seconds |'
      end if
   end subroutine

   ! One soap_turbo block from the reference's sparse-set files
   ! (src/read_files.f90:660-724: one value per line, alphas then desc_sparse
   ! with the descriptor component fastest). Returns the model id.
   subroutine soap_model_from_files(file_alphas, file_desc, species_types, alpha_max, l_max, rcut_hard, rcut_soft, &
                                    atom_sigma_r, atom_sigma_t, central_weight, basis, central_species, delta, zeta, &
                                    model_id)
      character(len=*), intent(in) :: file_alphas, file_desc, basis
      integer, intent(in) :: species_types(:), alpha_max(:), l_max, central_species
      real(dp), intent(in) :: rcut_hard(:), rcut_soft(:), atom_sigma_r(:), atom_sigma_t(:), central_weight(:)
      real(dp), intent(in) :: delta, zeta
      integer, intent(out) :: model_id
      real(dp), allocatable :: alphas(:), qs(:, :)
      real(dp), allocatable, target :: r_h(:), r_s(:), nf(:), gs(:), sr(:), srs(:), stt(:), sts(:), amp(:), cw(:)
      integer(c_int), allocatable, target :: am(:)
      integer(c_int), allocatable :: st(:)
      type(tgx_soap_hypers) :: h
      integer :: n_sparse, dim, i, j, ios, u, ns
      integer(c_int) :: mid
      real(dp) :: junk
      open (newunit=u, file=file_alphas, status="old")
      n_sparse = 0
      do
         read (u, *, iostat=ios) junk
         if (ios /= 0) exit
         n_sparse = n_sparse + 1
      end do
      close (u)
      open (newunit=u, file=file_desc, status="old")
      i = 0
      do
         read (u, *, iostat=ios) junk
         if (ios /= 0) exit
         i = i + 1
      end do
      close (u)
      dim = i/n_sparse
      allocate (alphas(n_sparse), qs(dim, n_sparse))
      open (newunit=u, file=file_alphas, status="old")
      do i = 1, n_sparse
         read (u, *) alphas(i)
      end do
      close (u)
      open (newunit=u, file=file_desc, status="old")
      do i = 1, n_sparse
         do j = 1, dim
            read (u, *) qs(j, i)
         end do
      end do
      close (u)
      ns = size(species_types)
      allocate (am(ns), st(ns), r_h(ns), r_s(ns), nf(ns), gs(ns), sr(ns), srs(ns), stt(ns), sts(ns), amp(ns), cw(ns))
      am = int(alpha_max, c_int); st = int(species_types, c_int)
      r_h = rcut_hard; r_s = rcut_soft; nf = 4.0_dp; gs = 1.0_dp; sr = atom_sigma_r; srs = 0.0_dp
      stt = atom_sigma_t; sts = 0.0_dp; amp = 1.0_dp; cw = central_weight
      h%n_species = ns; h%l_max = l_max; h%radial_enhancement = 0; h%n_quad_soft = 32; h%n_quad_buffer = 16
      select case (trim(basis))
      case ("poly3")
         h%basis = 0
      case ("poly3gauss")
         h%basis = 1
      case default
         write (*, '(A,A,A)') 'ERROR in soap_model_from_files: basis "', trim(basis), &
            '" is not implemented (poly3, poly3gauss)'
         error stop 1
      end select
      h%alpha_max = c_loc(am); h%rcut_hard = c_loc(r_h); h%rcut_soft = c_loc(r_s); h%nf = c_loc(nf)
      h%global_scaling = c_loc(gs); h%atom_sigma_r = c_loc(sr); h%atom_sigma_r_scaling = c_loc(srs)
      h%atom_sigma_t = c_loc(stt); h%atom_sigma_t_scaling = c_loc(sts); h%amplitude_scaling = c_loc(amp)
      h%central_weight = c_loc(cw)
      call tgx_check(c_tgx_model(ctx, h, st, int(central_species, c_int), delta, zeta, 0.0_dp, int(n_sparse, c_int), &
                                 int(dim, c_int), qs, alphas, mid), "soap_model_from_files")
      model_id = int(mid)
   end subroutine

   subroutine soap_backend_begin_id(model_id)
      integer, intent(in) :: model_id
      call tgx_check(c_tgx_begin(ctx, int(model_id, c_int)), "soap_backend_begin")
   end subroutine

   ! soap_backend_begin(hypers) (src/gap.f90:49-58): the block get_gap_soap and
   ! get_soap use next; its sparse set goes to the device on first use and stays.
   subroutine soap_backend_begin_hypers(hypers)
      type(soap_turbo), intent(in), target :: hypers
      cur_hypers => hypers
   end subroutine

   subroutine soap_backend_end()
      nullify (cur_hypers)
      call tgx_check(c_tgx_end(ctx), "soap_backend_end")
   end subroutine

   ! get_gap_soap for the current block on the list last built: energies0 and
   ! forces0(3, n_sites) are ASSIGNED for the atoms the block touches
   ! (src/gap_interface.f90:422-431); virial(3,3) is written.
   subroutine get_gap_soap_resident(path, do_forces, energies0, forces0, virial)
      integer, intent(in) :: path
      logical, intent(in) :: do_forces
      real(dp), intent(inout) :: energies0(:), forces0(:, :)
      real(dp), intent(out) :: virial(3, 3)
      integer(c_int) :: nsb
      integer(c_int64_t) :: npb
      call tgx_check(c_tgx_gap_soap(ctx, int(path, c_int), merge(1_c_int, 0_c_int, do_forces), energies0, forces0, &
                                    virial, nsb, npb), "get_gap_soap")
   end subroutine

   ! ---- the reference call surface (src/gap_interface.f90:41-480) --------------

   ! Species names of the system to the integer codes of the last build, from
   ! the names the caller passes per supercell atom.
   subroutine learn_species(names)
      character*8, intent(in) :: names(:)
      integer :: k, t
      logical :: known
      do k = 1, min(size(names), size(res_species))
         known = .false.
         do t = 1, n_sp
            if (sp_name(t) == names(k)) then
               known = .true.
               exit
            end if
         end do
         if (.not. known .and. n_sp < size(sp_name)) then
            n_sp = n_sp + 1
            sp_name(n_sp) = names(k)
            sp_code(n_sp) = res_species(k)
         end if
      end do
   end subroutine

   integer function species_code(name, slot)
      character*8, intent(in) :: name
      integer, intent(in) :: slot
      integer :: t
      species_code = 1000000 + slot  ! a species absent from the system
      do t = 1, n_sp
         if (sp_name(t) == name) then
            species_code = sp_code(t)
            return
         end if
      end do
   end function

   subroutine require(ok, what)
      logical, intent(in) :: ok
      character(len=*), intent(in) :: what
      if (ok) return
      write (*, '(A,A)') 'ERROR in get_gap_soap: ', what
      error stop 1
   end subroutine

   ! The block of the current hypers' sparse set with the given SOAP hypers,
   ! created on first use (the sparse set stays on the device). `for_energy`
   ! blocks carry the system's species codes; descriptor-only blocks
   ! (get_soap) any distinct codes.
   integer function block_for(for_energy, species_types, alpha_max, l_max, n_soap, rcut_hard, rcut_soft, nf, &
                              global_scaling, atom_sigma_r, atom_sigma_r_scaling, atom_sigma_t, &
                              atom_sigma_t_scaling, amplitude_scaling, radial_enhancement, central_weight, basis, &
                              scaling_mode, delta, zeta, central_species, compress_soap, compress_P_nonzero, &
                              compress_P_i, compress_P_j, compress_P_el) result(mid_out)
      logical, intent(in) :: for_energy, compress_soap
      character*8, intent(in) :: species_types(:)
      integer, intent(in) :: alpha_max(:), l_max, n_soap, radial_enhancement, central_species, compress_P_nonzero, &
                             compress_P_i(:), compress_P_j(:)
      real(dp), intent(in) :: compress_P_el(:)
      real(dp), intent(in) :: rcut_hard(:), rcut_soft(:), nf(:), global_scaling(:), atom_sigma_r(:), &
                              atom_sigma_r_scaling(:), atom_sigma_t(:), atom_sigma_t_scaling(:), &
                              amplitude_scaling(:), central_weight(:), delta, zeta
      character(len=*), intent(in) :: basis, scaling_mode
      type(c_ptr) :: key
      integer :: k, ns, nsp
      integer(c_int), allocatable, target :: am(:)
      integer(c_int), allocatable :: st(:)
      real(dp), allocatable, target :: r_h(:), r_s(:), nf_(:), gs(:), sr(:), srs(:), stt(:), sts(:), amp(:), cw(:)
      real(dp), allocatable :: qz(:, :), az(:)
      integer(c_int), allocatable :: cpi(:), cpj(:)
      type(tgx_soap_hypers) :: h
      integer(c_int) :: mid, cs
      call require(associated(cur_hypers), "no soap_backend_begin(hypers) before this call")
      if (allocated(cur_hypers%Qs)) then
         key = c_loc(cur_hypers%Qs)
      else
         key = c_loc(cur_hypers)
      end if
      if (for_energy) then
         do k = 1, n_reg
            if (c_associated(reg_key(k), key)) then
               mid_out = reg_id(k)
               return
            end if
         end do
      else
         do k = 1, n_reg_desc
            if (c_associated(reg_desc_key(k), key)) then
               mid_out = reg_desc_id(k)
               return
            end if
         end do
      end if
      call require(trim(scaling_mode) == "polynomial", 'scaling_mode "'//trim(scaling_mode)//'" is not implemented')
      nsp = size(species_types)
      allocate (am(nsp), st(nsp), r_h(nsp), r_s(nsp), nf_(nsp), gs(nsp), sr(nsp), srs(nsp), stt(nsp), sts(nsp), &
                amp(nsp), cw(nsp))
      do k = 1, nsp
         st(k) = int(merge(species_code(species_types(k), k), 1000000 + k, for_energy), c_int)
      end do
      am = int(alpha_max, c_int)
      r_h = rcut_hard; r_s = rcut_soft; nf_ = nf; gs = global_scaling; sr = atom_sigma_r
      srs = atom_sigma_r_scaling; stt = atom_sigma_t; sts = atom_sigma_t_scaling; amp = amplitude_scaling
      cw = central_weight
      h%n_species = nsp; h%l_max = l_max; h%radial_enhancement = radial_enhancement
      h%n_quad_soft = 32; h%n_quad_buffer = 16
      select case (trim(basis))
      case ("poly3")
         h%basis = 0
      case ("poly3gauss")
         h%basis = 1
      case default
         call require(.false., 'basis "'//trim(basis)//'" is not implemented (poly3, poly3gauss)')
      end select
      h%alpha_max = c_loc(am); h%rcut_hard = c_loc(r_h); h%rcut_soft = c_loc(r_s); h%nf = c_loc(nf_)
      h%global_scaling = c_loc(gs); h%atom_sigma_r = c_loc(sr); h%atom_sigma_r_scaling = c_loc(srs)
      h%atom_sigma_t = c_loc(stt); h%atom_sigma_t_scaling = c_loc(sts); h%amplitude_scaling = c_loc(amp)
      h%central_weight = c_loc(cw)
      if (compress_soap) then
         call require(compress_P_nonzero >= 1 .and. size(compress_P_i) >= compress_P_nonzero .and. &
                      size(compress_P_j) >= compress_P_nonzero .and. size(compress_P_el) >= compress_P_nonzero, &
                      "compress_soap without compress_P_i / compress_P_j / compress_P_el")
         allocate (cpi(compress_P_nonzero), cpj(compress_P_nonzero))
         cpi = int(compress_P_i(1:compress_P_nonzero), c_int)
         cpj = int(compress_P_j(1:compress_P_nonzero), c_int)
      end if
      cs = int(merge(central_species, 0, for_energy), c_int)
      if (allocated(cur_hypers%Qs) .and. allocated(cur_hypers%alphas)) then
         ns = size(cur_hypers%alphas)
         call require(size(cur_hypers%Qs, 1) == n_soap, "the sparse set's descriptor size differs from n_soap")
         allocate (qz(n_soap, ns), az(ns))
         qz = cur_hypers%Qs(1:n_soap, 1:ns)
         az = cur_hypers%alphas(1:ns)
      else  ! a descriptor-only use needs no sparse set
         ns = 1
         cs = 0_c_int
         allocate (qz(n_soap, 1), az(1))
         qz = 0.0_dp; az = 0.0_dp
      end if
      if (compress_soap) then
         call tgx_check(c_tgx_model_c(ctx, h, st, cs, delta, zeta, 0.0_dp, int(ns, c_int), int(n_soap, c_int), qz, &
                                      az, int(compress_P_nonzero, c_int), cpi, cpj, compress_P_el, mid), &
                        "soap_backend_begin")
      else
         call tgx_check(c_tgx_model(ctx, h, st, cs, delta, zeta, 0.0_dp, int(ns, c_int), int(n_soap, c_int), qz, az, &
                                    mid), "soap_backend_begin")
      end if
      mid_out = int(mid)
      if (for_energy) then
         call require(n_reg < size(reg_key), "too many soap_turbo blocks")
         n_reg = n_reg + 1
         reg_key(n_reg) = key
         reg_id(n_reg) = mid_out
      else
         call require(n_reg_desc < size(reg_desc_key), "too many soap_turbo blocks")
         n_reg_desc = n_reg_desc + 1
         reg_desc_key(n_reg_desc) = key
         reg_desc_id(n_reg_desc) = mid_out
      end if
   end function

   ! get_gap_soap with the reference's arguments in its order
   ! (src/gap_interface.f90:41-57) for the block of soap_backend_begin(hypers):
   ! the centres among sites this_i_beg..this_i_end of the list that
   ! build_neighbors_list made resident (n_neigh0 and neighbors_list0 must be
   ! that list's slices), energies0 and forces0 ASSIGNED for the atoms the block
   ! touches (src/gap_interface.f90:422-431), virial written. The descriptors
   ! stay on the device: soap and soap_cart_der are returned unallocated, and
   ! write_soap / write_derivatives, local properties, dipole models, species
   ! multiplicity and the single-atom mode (all_atoms false) stop with an
   ! error. Compressed blocks (compress_soap, compress_P_*) are evaluated.
   subroutine get_gap_soap_ref(n_total_sites, n_sites0, n_neigh0, neighbors_list0, n_species, species_types, &
                               rjs0, thetas0, phis0, xyz0, alpha_max, l_max, n_soap, &
                               rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r, &
                               atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, &
                               amplitude_scaling, radial_enhancement, central_weight, basis, &
                               scaling_mode, do_timing, do_derivatives, do_forces, do_prediction, &
                               write_soap, write_derivatives, compress_soap, &
                               compress_P_nonzero, compress_P_i, compress_P_j, compress_P_el, &
                               delta, zeta, central_species, &
                               xyz_species, xyz_species_supercell, all_atoms, &
                               which_atom, indices, soap, soap_cart_der, der_neighbors, der_neighbors_list, &
                               has_local_properties, n_local_properties, local_property_models, &
                               energies0, forces0, local_properties0, local_properties_cart_der0, &
                               local_property_indexes, this_i_beg, this_i_end, this_j_beg, &
                               this_j_end, virial, lp_index, is_dipole_model, local_dipoles0, energies_dipole0)
      type(local_property_soap_turbo), allocatable, intent(in) :: local_property_models(:)
      real(dp), intent(in) :: rjs0(:), thetas0(:), phis0(:), xyz0(:, :), rcut_hard(:), rcut_soft(:), nf(:), &
                              global_scaling(:), atom_sigma_r(:), atom_sigma_r_scaling(:), atom_sigma_t(:), &
                              atom_sigma_t_scaling(:), amplitude_scaling(:), central_weight(:), delta, zeta, &
                              compress_P_el(:)
      integer, intent(in) :: n_sites0, n_neigh0(:), neighbors_list0(:), n_species, central_species, &
                             radial_enhancement, which_atom, indices(1:3), alpha_max(:), l_max, n_total_sites, &
                             compress_P_nonzero, compress_P_i(:), compress_P_j(:), n_local_properties, &
                             local_property_indexes(:), this_i_beg, this_i_end, this_j_beg, this_j_end, lp_index, &
                             n_soap
      logical, intent(in) :: is_dipole_model, do_timing, do_derivatives, compress_soap, do_forces, do_prediction, &
                             all_atoms, write_soap, write_derivatives, has_local_properties
      character*64, intent(in) :: basis
      character*32, intent(in) :: scaling_mode
      character*8, intent(in) :: xyz_species(:), xyz_species_supercell(:), species_types(:)
      real(dp), allocatable, intent(out) :: soap(:, :), soap_cart_der(:, :, :)
      real(dp), intent(out) :: virial(1:3, 1:3)
      real(dp), intent(inout) :: energies0(:), forces0(:, :), local_properties0(:, :), &
                                 local_properties_cart_der0(:, :, :), local_dipoles0(:, :), energies_dipole0(:)
      integer, allocatable :: der_neighbors(:), der_neighbors_list(:)
      integer :: mid, i, j
      integer(c_int) :: nsb
      integer(c_int64_t) :: npb
      integer(8) :: t0, t1, rate
      if (do_timing) call system_clock(t0, rate)
      call require(.not. is_dipole_model, "dipole models are not implemented")
      call require(.not. (write_soap .or. write_derivatives), "write_soap / write_derivatives are not implemented")
      call require(.not. has_local_properties, "local properties through get_gap_soap are not implemented")
      call require(all_atoms, "all_atoms = .false. (which_atom) is not implemented")
      do i = 1, n_species
         do j = i + 1, n_species
            call require(species_types(i) /= species_types(j), "species multiplicity above 1 is not implemented")
         end do
      end do
      ! the caller's list must be the resident one
      call require(allocated(res_n_neigh) .and. n_total_sites == res_n_sites, &
                   "no resident neighbour list of this system (build_neighbors_list first)")
      call require(this_i_beg >= 1 .and. this_i_end <= res_n_sites .and. &
                   n_sites0 == this_i_end - this_i_beg + 1 .and. size(n_neigh0) == n_sites0, "bad site range")
      call require(this_j_beg >= 1 .and. this_j_end <= size(res_nlist) .and. &
                   size(neighbors_list0) == this_j_end - this_j_beg + 1, "bad pair range")
      call require(all(n_neigh0 == res_n_neigh(this_i_beg:this_i_end)) .and. &
                   all(neighbors_list0 == res_nlist(this_j_beg:this_j_end)), &
                   "the neighbour list differs from the one build_neighbors_list made resident")
      call require(size(energies0) >= res_n_sites .and. size(forces0, 2) >= res_n_sites, &
                   "energies0 / forces0 smaller than the system")
      call learn_species(xyz_species_supercell)
      mid = block_for(.true., species_types, alpha_max, l_max, n_soap, rcut_hard, rcut_soft, nf, global_scaling, &
                      atom_sigma_r, atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, amplitude_scaling, &
                      radial_enhancement, central_weight, basis, scaling_mode, delta, zeta, central_species, &
                      compress_soap, compress_P_nonzero, compress_P_i, compress_P_j, compress_P_el)
      virial = 0.0_dp
      if (do_prediction) then
         call tgx_check(c_tgx_begin(ctx, int(mid, c_int)), "get_gap_soap")
         call tgx_check(c_tgx_gap_soap_sites(ctx, gap_path, merge(1_c_int, 0_c_int, do_forces), &
                                             int(this_i_beg, c_int), int(this_i_end, c_int), energies0, forces0, &
                                             virial, nsb, npb), "get_gap_soap")
      end if
      if (do_timing) then
         call system_clock(t1)
         write (*, '(A, F9.3, A)') '  *) soap_turbo (device): ', real(t1 - t0, dp)/real(rate, dp), ' seconds |'
      end if
   end subroutine

   ! soap_turbo's get_soap (call shape src/gap_interface.f90:320-325) on a list
   ! the caller marshalled: n_sites centres with n_neigh entries each (self
   ! first), the neighbours' species in mask(pair, species), rjs, thetas and
   ! phis; soap(n_soap, n_sites) and, with do_derivatives,
   ! soap_cart_der(3, n_soap, n_atom_pairs) are written. The hypers must be
   ! those of soap_backend_begin(hypers)'s block.
   subroutine get_soap(n_sites, n_neigh, n_species, species, species_multiplicity, n_atom_pairs, mask, rjs, &
                       thetas, phis, alpha_max, l_max, rcut_hard, rcut_soft, nf, global_scaling, &
                       atom_sigma_r, atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, &
                       amplitude_scaling, radial_enhancement, central_weight, basis, scaling_mode, do_timing, &
                       do_derivatives, compress_soap, compress_P_nonzero, compress_P_i, compress_P_j, &
                       compress_P_el, soap, soap_cart_der)
      integer, intent(in) :: n_sites, n_neigh(:), n_species, species(:, :), species_multiplicity(:), n_atom_pairs, &
                             alpha_max(:), l_max, radial_enhancement, compress_P_nonzero, compress_P_i(:), &
                             compress_P_j(:)
      logical, intent(in) :: mask(:, :), do_timing, do_derivatives, compress_soap
      real(dp), intent(in) :: rjs(:), thetas(:), phis(:), rcut_hard(:), rcut_soft(:), nf(:), global_scaling(:), &
                              atom_sigma_r(:), atom_sigma_r_scaling(:), atom_sigma_t(:), atom_sigma_t_scaling(:), &
                              amplitude_scaling(:), central_weight(:), compress_P_el(:)
      character(len=*), intent(in) :: basis, scaling_mode
      real(dp), intent(inout) :: soap(:, :), soap_cart_der(:, :, :)
      integer :: mid, k, s
      integer(c_int), allocatable :: nn(:), ps(:)
      real(dp), allocatable :: xyz(:, :), der(:, :, :)
      character*8, allocatable :: names(:)
      call require(all(species_multiplicity(1:n_sites) == 1), "species multiplicity above 1 is not implemented")
      call require(size(soap, 2) >= n_sites, "soap smaller than n_sites")
      allocate (names(n_species))
      names = ""
      mid = block_for(.false., names, alpha_max, l_max, size(soap, 1), rcut_hard, rcut_soft, nf, global_scaling, &
                      atom_sigma_r, atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, amplitude_scaling, &
                      radial_enhancement, central_weight, basis, scaling_mode, 1.0_dp, 2.0_dp, 0, compress_soap, &
                      compress_P_nonzero, compress_P_i, compress_P_j, compress_P_el)
      allocate (nn(n_sites), ps(max(n_atom_pairs, 1)), xyz(3, max(n_atom_pairs, 1)))
      nn = int(n_neigh(1:n_sites), c_int)
      do k = 1, n_atom_pairs
         ps(k) = -1
         do s = 1, n_species
            if (mask(k, s)) then
               ps(k) = int(s - 1, c_int)
               exit
            end if
         end do
         xyz(1, k) = rjs(k)*sin(thetas(k))*cos(phis(k))
         xyz(2, k) = rjs(k)*sin(thetas(k))*sin(phis(k))
         xyz(3, k) = rjs(k)*cos(thetas(k))
      end do
      call tgx_check(c_tgx_begin(ctx, int(mid, c_int)), "get_soap")
      if (do_derivatives .and. n_atom_pairs > 0) then
         call require(size(soap_cart_der, 3) >= n_atom_pairs, "soap_cart_der smaller than n_atom_pairs")
         call tgx_check(c_tgx_soap_list(ctx, int(n_sites, c_int), nn, int(n_atom_pairs, c_int64_t), ps, rjs, xyz, &
                                        1_c_int, soap, soap_cart_der), "get_soap")
      else
         allocate (der(3, 1, 1))
         call tgx_check(c_tgx_soap_list(ctx, int(n_sites, c_int), nn, int(n_atom_pairs, c_int64_t), ps, rjs, xyz, &
                                        0_c_int, soap, der), "get_soap")
      end if
      call tgx_check(c_tgx_end(ctx), "get_soap")
   end subroutine

   ! get_soap_energy_and_forces with the reference's arguments in its order
   ! (src/gap.f90:66-68) on host arrays, for the current block. delta and zeta0
   ! belong to the block (given when it was created): other values stop with an
   ! error.
   subroutine get_soap_energy_and_forces(soap, soap_der, delta, zeta0, e0, n_neigh, neighbors_list, xyz, do_forces, &
                                         do_timing, energies, forces, virial)
      real(dp), intent(in) :: soap(:, :), soap_der(:, :, :), delta, zeta0, e0, xyz(:, :)
      integer, intent(in) :: n_neigh(:), neighbors_list(:)
      logical, intent(in) :: do_forces, do_timing
      real(dp), intent(out) :: energies(:), forces(:, :), virial(3, 3)
      integer(c_int), allocatable :: nn(:), nl(:)
      integer(8) :: t0, t1, rate
      if (do_timing) call system_clock(t0, rate)
      allocate (nn(size(n_neigh)), nl(size(neighbors_list)))
      nn = int(n_neigh, c_int); nl = int(neighbors_list, c_int)
      call tgx_check(c_tgx_eaf(ctx, int(size(soap, 2), c_int), int(size(soap, 1), c_int), &
                               int(size(soap_der, 3), c_int64_t), soap, soap_der, delta, zeta0, e0, nn, nl, xyz, &
                               int(size(forces, 2), c_int), merge(1_c_int, 0_c_int, do_forces), energies, forces, &
                               virial), "get_soap_energy_and_forces")
      if (do_timing) then
         call system_clock(t1)
         write (*, '(A, F9.3, A)') '  *) GAP prediction: ', real(t1 - t0, dp)/real(rate, dp), ' seconds |'
      end if
   end subroutine

   ! A distance_2b block (species as the codes given to build_neighbors_list),
   ! with the buffer 0.5 and e0 = 0 the gap_backend passes
   ! (src/gap_backend_cpu.f90:120-126). Qs(n_sparse) distances.
   function tgx_distance_2b_model(species1, species2, rcut, delta, sigma, Qs, alphas, cutoff) result(id)
      integer, intent(in) :: species1, species2
      real(dp), intent(in) :: rcut, delta, sigma, Qs(:), alphas(:), cutoff(:)
      integer :: id
      integer(c_int) :: mid
      real(c_double), allocatable :: q(:), a(:), c(:)
      q = Qs
      a = alphas
      c = cutoff
      call tgx_check(c_tgx_2b(ctx, int(species1, c_int), int(species2, c_int), rcut, 0.5_dp, delta, sigma, 0.0_dp, &
                              int(size(q), c_int), q, a, c, mid), "distance_2b")
      id = int(mid)
   end function

   ! An angle_3b block: Qs(n_sparse, 3) as the reference stores it
   ! (src/read_files.f90:743-770), sigma(3), kernel_type "exp" or "pol"; buffer
   ! 0.5 and e0 = 0 as the gap_backend passes (src/gap_backend_cpu.f90:250-258).
   function tgx_angle_3b_model(species_center, species1, species2, rcut, delta, sigma, kernel_type, Qs, alphas, &
                               cutoff) result(id)
      integer, intent(in) :: species_center, species1, species2
      real(dp), intent(in) :: rcut, delta, sigma(3), Qs(:, :), alphas(:), cutoff(:)
      character(len=*), intent(in) :: kernel_type
      integer :: id
      integer(c_int) :: mid
      real(c_double), allocatable :: qt(:, :), a(:), c(:)
      real(c_double) :: sg(3)
      qt = transpose(Qs)  ! (3, n_sparse): each sparse point's components contiguous
      a = alphas
      c = cutoff
      sg = sigma
      call tgx_check(c_tgx_3b(ctx, int(species_center, c_int), int(species1, c_int), int(species2, c_int), rcut, &
                              0.5_dp, delta, sg, 0.0_dp, trim(kernel_type)//c_null_char, int(size(a), c_int), qt, a, &
                              c, mid), "angle_3b")
      id = int(mid)
   end function

   ! A core_pot block: the table V(x), end slopes (>= 0.99e30: natural end) and
   ! core_pot_cutoff / core_pot_buffer (src/read_files.f90:5392-5510).
   function tgx_core_pot_model(species1, species2, x, V, yp1, ypn, cutoff, buffer) result(id)
      integer, intent(in) :: species1, species2
      real(dp), intent(in) :: x(:), V(:), yp1, ypn, cutoff, buffer
      integer :: id
      integer(c_int) :: mid
      real(c_double), allocatable :: xx(:), vv(:)
      xx = x
      vv = V
      call tgx_check(c_tgx_core_pot(ctx, int(species1, c_int), int(species2, c_int), int(size(xx), c_int), xx, vv, &
                                    yp1, ypn, cutoff, buffer, mid), "core_pot")
      id = int(mid)
   end function

   ! One model's energies, forces and virial on the resident list, accumulated
   ! into the driver's arrays as add_2b_contribution / add_3b_contribution /
   ! add_core_pot_contribution accumulate into energies_2b etc.
   subroutine add_pair_contribution(id, energies, forces, virial)
      integer, intent(in) :: id
      real(dp), intent(inout) :: energies(:), forces(:, :), virial(3, 3)
      real(c_double), allocatable :: e(:), f(:, :)
      real(c_double) :: w(3, 3)
      allocate (e(size(energies)), f(3, size(forces, 2)))
      call tgx_check(c_tgx_pair_eaf(ctx, int(id, c_int), e, f, w), "pair_energy_and_forces")
      energies = energies + e
      forces = forces + f
      virial = virial + w
   end subroutine

   ! ---- device MD (src/turbogap_md.f90:160-470) ----------------------------
   ! The system is the one of the last build_neighbors_list; velocities(3, n) in
   ! A/fs, masses(n) in eV fs^2/A^2, fix_atom(3, n) as velocity_verlet takes it.
   subroutine md_state_init(velocities, masses, fix_atom)
      real(dp), intent(in) :: velocities(:, :), masses(:)
      logical, intent(in), optional :: fix_atom(:, :)
      integer(c_int), allocatable :: f(:, :)
      call tgx_check(c_md_set_state(ctx, velocities, masses), "md_state_init")
      if (present(fix_atom)) then
         allocate (f(size(fix_atom, 1), size(fix_atom, 2)))
         f = merge(1_c_int, 0_c_int, fix_atom)
      else
         allocate (f(3, size(masses)))
         f = 0_c_int
      end if
      ! without fix_atom every component is free (a mask of an earlier run is dropped)
      call tgx_check(c_md_set_fixed(ctx, f), "md_state_init")
   end subroutine

   subroutine md_state_get(positions, velocities)
      real(dp), intent(out) :: positions(:, :), velocities(:, :)
      call tgx_check(c_md_get_state(ctx, positions, velocities), "md_state_get")
   end subroutine

   ! A resident list of every site within rcut_max, nothing copied to the host.
   subroutine md_build_list(rcut_max)
      real(dp), intent(in) :: rcut_max
      integer(c_int) :: ind(3)
      integer(c_int64_t) :: np
      ind = 1_c_int
      call tgx_check(c_tgx_build(ctx, rcut_max, ind, c_null_ptr, np), "md_build_list")
   end subroutine

   subroutine md_refresh_list()
      call tgx_check(c_tgx_refresh(ctx), "md_refresh_list")
   end subroutine

   ! Energies, forces and virial of every block on the resident list; the
   ! forces stay on the device for the integrator.
   subroutine md_forces(path, energy, virial)
      integer, intent(in) :: path
      real(dp), intent(out) :: energy, virial(3, 3)
      call tgx_check(c_compute(ctx, int(path, c_int), energy, virial), "md_forces")
   end subroutine

   subroutine wrap_pbc_device()
      call tgx_check(c_wrap(ctx), "wrap_pbc")
   end subroutine

   subroutine remove_cm_vel_device()
      call tgx_check(c_remove_cm(ctx), "remove_cm_vel")
   end subroutine

   subroutine variable_time_step_device(init, target_pos_step, tau_dt, dt0, dt)
      logical, intent(in) :: init
      real(dp), intent(in) :: target_pos_step, tau_dt, dt0
      real(dp), intent(inout) :: dt
      call tgx_check(c_vdt(ctx, merge(1_c_int, 0_c_int, init), target_pos_step, tau_dt, dt0, dt), &
                     "variable_time_step")
   end subroutine

   subroutine variable_time_step_adaptive_device(init, tmin, tmax, xmax, emax, dt0, dt)
      logical, intent(in) :: init
      real(dp), intent(in) :: tmin, tmax, xmax, emax, dt0
      real(dp), intent(inout) :: dt
      call tgx_check(c_adt(ctx, merge(1_c_int, 0_c_int, init), tmin, tmax, xmax, emax, dt0, dt), &
                     "variable_time_step_adaptive")
   end subroutine

   subroutine velocity_verlet_device(dt, first_step)
      real(dp), intent(in) :: dt
      logical, intent(in) :: first_step
      call tgx_check(c_vv(ctx, dt, merge(1_c_int, 0_c_int, first_step)), "velocity_verlet")
   end subroutine

   subroutine gradient_descent_device(energy, max_opt_step, first_step, gamma, restored)
      real(dp), intent(in) :: energy, max_opt_step
      logical, intent(in) :: first_step
      real(dp), intent(out) :: gamma
      logical, intent(out) :: restored
      integer(c_int) :: r
      call tgx_check(c_gd(ctx, energy, max_opt_step, merge(1_c_int, 0_c_int, first_step), gamma, r), &
                     "gradient_descent")
      restored = r /= 0
   end subroutine

   subroutine md_kinetic(e_kinetic, instant_temp)
      real(dp), intent(out) :: e_kinetic, instant_temp
      call tgx_check(c_kinetic(ctx, e_kinetic, instant_temp), "md_kinetic")
   end subroutine

   subroutine md_pressure(virial, instant_temp, instant_pressure, instant_pressure_tensor)
      real(dp), intent(in) :: virial(3, 3), instant_temp
      real(dp), intent(out) :: instant_pressure, instant_pressure_tensor(3, 3)
      call tgx_check(c_pressure(ctx, virial, instant_temp, instant_pressure, instant_pressure_tensor), "md_pressure")
   end subroutine

   ! box_scaling_factor(3, 3) as the reference holds it; the box vectors are updated.
   subroutine box_scaling_device(box_scaling_factor, md_istep, md_nsteps, a_box, b_box, c_box)
      real(dp), intent(in) :: box_scaling_factor(3, 3)
      integer, intent(in) :: md_istep, md_nsteps
      real(dp), intent(inout) :: a_box(3), b_box(3), c_box(3)
      real(dp) :: g(3, 3), box(3, 3)
      g = transpose(box_scaling_factor)  ! the C ABI reads it row by row
      call tgx_check(c_box_scaling(ctx, g, int(md_istep, c_int), int(md_nsteps, c_int), box), "box_scaling")
      a_box = box(:, 1); b_box = box(:, 2); c_box = box(:, 3)
   end subroutine

   subroutine berendsen_barostat_device(p0, instant_pressure_tensor, barostat_sym, tau_p, gamma_p, dt, a_box, &
                                        b_box, c_box)
      real(dp), intent(in) :: p0, instant_pressure_tensor(3, 3), tau_p, gamma_p, dt
      character(len=*), intent(in) :: barostat_sym
      real(dp), intent(inout) :: a_box(3), b_box(3), c_box(3)
      real(dp) :: box(3, 3)
      call tgx_check(c_barostat(ctx, p0, instant_pressure_tensor, trim(barostat_sym)//c_null_char, tau_p, gamma_p, &
                                dt, box), "berendsen_barostat")
      a_box = box(:, 1); b_box = box(:, 2); c_box = box(:, 3)
   end subroutine

   subroutine berendsen_thermostat_device(target_temp, instant_temp, tau_t, dt)
      real(dp), intent(in) :: target_temp, instant_temp, tau_t, dt
      call tgx_check(c_berendsen(ctx, target_temp, instant_temp, tau_t, dt), "berendsen_thermostat")
   end subroutine

   ! gauss and chi2 are resamplekin's gasdev() and sumnoises(3 n_sites - 4).
   subroutine bussi_thermostat_device(target_temp, tau_t, dt, gauss, chi2)
      real(dp), intent(in) :: target_temp, tau_t, dt, gauss, chi2
      real(dp) :: ek
      call tgx_check(c_bussi(ctx, target_temp, tau_t, dt, gauss, chi2, ek), "bussi_thermostat")
   end subroutine

   ! The neighbors_buffer test (src/turbogap_md.f90:451-471): reset after a build.
   subroutine md_displacement(neighbors_buffer, reset, rebuild_neighbors_list)
      real(dp), intent(in) :: neighbors_buffer
      logical, intent(in) :: reset
      logical, intent(out) :: rebuild_neighbors_list
      integer(c_int) :: r
      call tgx_check(c_displacement(ctx, neighbors_buffer, merge(1_c_int, 0_c_int, reset), r), "md_displacement")
      rebuild_neighbors_list = r /= 0
   end subroutine

end module tgx_turbogap
