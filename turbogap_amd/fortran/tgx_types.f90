! The derived types TurboGAP passes at its SOAP call surface (src/types.f90:
! input_parameters, soap_turbo, local_property_soap_turbo), reduced to the
! fields tgx_turbogap reads, with the reference's names, kinds and defaults.
! Built into TurboGAP, tgx_turbogap takes these types from the reference's own
! module `types` instead (`use types` in place of `use tgx_types`): the field
! names are the same.
module tgx_types
   use iso_c_binding, only: c_double
   implicit none
   private
   integer, parameter :: tdp = c_double

   type, public :: local_property_soap_turbo
      real(tdp), allocatable :: Qs(:, :)
      real(tdp), allocatable :: alphas(:)
      real(tdp) :: zeta
      real(tdp) :: delta
      real(tdp) :: V0
      integer :: n_sparse
      integer :: dim
      logical :: do_derivatives = .false.
   end type

   type, public :: soap_turbo
      real(tdp), allocatable :: nf(:), rcut_hard(:), rcut_soft(:), atom_sigma_r(:), atom_sigma_t(:)
      real(tdp), allocatable :: atom_sigma_r_scaling(:), atom_sigma_t_scaling(:), amplitude_scaling(:)
      real(tdp), allocatable :: central_weight(:), global_scaling(:)
      real(tdp), allocatable :: alphas(:), Qs(:, :), compress_P_el(:)
      real(tdp) :: zeta = 2.d0
      real(tdp) :: delta = 1.d0
      integer, allocatable :: alpha_max(:), compress_P_i(:), compress_P_j(:)
      integer :: n_species
      integer :: central_species = 0
      integer :: dim
      integer :: l_max
      integer :: radial_enhancement = 0
      integer :: n_sparse
      integer :: compress_P_nonzero = 0
      integer :: n_local_properties = 0
      character*64 :: basis = "poly3"
      character*32 :: scaling_mode = "polynomial"
      character*8, allocatable :: species_types(:)
      logical :: compress_soap = .false.
      logical :: has_local_properties = .false.
      logical :: is_dipole_model = .false.
      type(local_property_soap_turbo), allocatable :: local_property_models(:)
   end type

   type, public :: input_parameters
      real(tdp) :: max_Gbytes_per_process = 1.d0
      logical :: do_timing = .false.
      logical :: do_derivatives = .true.
      logical :: do_forces = .true.
      logical :: do_prediction = .true.
      logical :: write_soap = .false.
      logical :: write_derivatives = .false.
      logical :: all_atoms = .true.
      integer :: which_atom = 0
   end type
end module tgx_types
