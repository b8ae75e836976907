/* tgx — MI355X-native SOAP descriptor and GAP energy/force path.
 *
 * C ABI of libtgx.so. Plain pointers and sizes only; every function returns an
 * int status (TGX_OK = 0) and tgx_last_error() gives the message. Arrays at this
 * boundary keep TurboGAP's conventions: Fortran column-major layouts, FP64, and
 * 1-based atom indices in neighbour lists. Host arrays are caller-owned; device
 * buffers are owned by the context.
 *
 * Reference interfaces replaced (paths relative to the TurboGAP tree):
 *   tgx_init / tgx_finalize             src/gpu_context.f90:41-58   gpu_context_init / _finalize
 *   tgx_memory_budget                   src/gpu_context.f90:100-162 gpu_memory_budget_init
 *   tgx_set_memory_budget               src/neighbors.f90:637-778   the batch split of get_number_of_atom_pairs
 *   tgx_set_system                      src/turbogap.f90:871-886    positions/species/box handed to the build
 *   tgx_number_of_unit_cells            src/neighbors.f90:175-236   number_of_unit_cells_for_given_cutoff
 *   tgx_build_supercell                 src/read_files.f90:320-367  the supercell replication of read_xyz
 *   tgx_build_neighbors_list            src/neighbors.f90:243-605   build_neighbors_list (count phase)
 *   tgx_get_neighbors_list              src/neighbors.f90:544-591   the allocatable outputs of build_neighbors_list
 *   tgx_refresh_neighbors_list          src/neighbors.f90:554-585   build_neighbors_list, rebuild_neighbors_list = .false.
 *   tgx_md_displacement                 src/turbogap_md.f90:451-471 the neighbors_buffer rebuild test
 *   tgx_soap_model_create               src/read_files.f90:660-724, 4641-5236  the soap_turbo block + sparse set
 *   tgx_soap_model_create_compressed    src/read_files.f90:5162-5231            the same, compress_soap
 *   tgx_soap_backend_begin / _end       src/gap.f90:49-64           soap_backend_begin / soap_backend_end
 *   tgx_get_gap_soap                    src/gap_interface.f90:41-480 get_gap_soap (marshal, get_soap, prediction, fold)
 *   tgx_get_gap_soap_sites              src/gap_interface.f90:41-480 one this_i_beg:this_i_end batch of it
 *   tgx_get_soap                        soap_turbo get_soap, call shape src/gap_interface.f90:320-325
 *   tgx_get_soap_list                   soap_turbo get_soap on a caller-marshalled list (the same call shape)
 *   tgx_device_count                    src/gpu_context.f90:41-48   the device a rank binds
 *   tgx_get_soap_energy_and_forces      src/gap.f90:66-259          get_soap_energy_and_forces
 *   tgx_compute                         src/turbogap.f90:1253-1380,1941-1954  the per-step SOAP loop and totals
 *   tgx_local_property_predict          src/local_properties.f90:174-297 local_property_predict
 *   tgx_velocity_verlet etc.            src/md.f90:226-345          velocity_verlet / berendsen_thermostat / remove_cm_vel
 *   tgx_md_set_fixed                    src/md.f90:226-283          velocity_verlet's fix_atom argument
 *   tgx_md_pressure                     src/turbogap_md.f90:287-295 instant_pressure, instant_pressure_tensor
 *   tgx_berendsen_barostat              src/md.f90:384-417, src/turbogap_md.f90:413-426 berendsen_barostat on box and positions
 *   tgx_box_scaling                     src/md.f90:421-458          box_scaling (scale_box, box_scaling_factor)
 *   tgx_gradient_descent                src/md.f90:540-664          gradient_descent (optimize = "gd")
 *   tgx_variable_time_step              src/md.f90:464-535, src/turbogap_md.f90:178-181 variable_time_step
 *   tgx_adaptive_time_step              src/stopping/adaptive_time.f90:46-164, src/turbogap_md.f90:204-210
 *                                       variable_time_step_adaptive (adaptive_time)
 *   tgx_bussi_thermostat                src/turbogap_md.f90:438-449, src/third_party/bussi_thermostat/resamplekin.f90:50-84
 *                                       thermostat = "bussi" (resamplekin on the device E_kin)
 *   tgx_distance_2b_create              src/read_files.f90:722-742,5238-5305  a distance_2b block and its sparse set
 *   tgx_core_pot_create                 src/read_files.f90:5392-5510 a core_pot block, its table and spline
 *   tgx_angle_3b_create                 src/read_files.f90:743-770,5305-5384  an angle_3b block and its sparse set
 *   tgx_pair_energy_and_forces          src/gap.f90:400-563 / 565-721 / 724-1112 get_2b_ / get_core_pot_ /
 *                                       get_3b_energy_and_forces
 */
#ifndef TGX_H
#define TGX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TGX_OK = 0,
  TGX_ERR_ARG = 1,         /* bad argument, size or state */
  TGX_ERR_HIP = 2,         /* a HIP runtime call failed */
  TGX_ERR_UNSUPPORTED = 3, /* a configuration this build does not implement */
  TGX_ERR_NOMEM = 4        /* device allocation failed */
};

/* Evaluation paths for tgx_compute / tgx_get_gap_soap. */
enum {
  TGX_PATH_FP64_PLAIN = 0,   /* FP64; dq/dr materialised, then contracted (the reference's data flow) */
  TGX_PATH_FP64_ADJOINT = 1, /* FP64; forces from dE/dq contracted with recomputed dq/dr */
  TGX_PATH_MIXED = 2,        /* FP32 descriptors, BF16 MFMA kernel with FP32 accumulation, FP64 sums */
  TGX_PATH_SURROGATE = 3     /* MLP surrogate per species (FP32 descriptors, BF16 MLP) */
};

typedef struct tgx_ctx tgx_ctx;

/* Hypers of one soap_turbo block (src/types.f90:75-134). Per-species arrays have
 * n_species entries. basis: 0 = "poly3", 1 = "poly3gauss". rcut_soft is the soft
 * cutoff (rcut - buffer, src/read_files.f90:5101-5109). n_quad_soft/n_quad_buffer:
 * Gauss-Legendre nodes of the radial projection on [0, rcut_soft] and
 * [rcut_soft, rcut_hard] (48 nodes on [0, rcut_hard] when there is no buffer);
 * their sum must not exceed 64. */
typedef struct tgx_soap_hypers {
  int n_species;
  int l_max;
  int basis;
  int radial_enhancement;
  int n_quad_soft;
  int n_quad_buffer;
  const int* alpha_max;
  const double* rcut_hard;
  const double* rcut_soft;
  const double* nf;
  const double* global_scaling;
  const double* atom_sigma_r;
  const double* atom_sigma_r_scaling;
  const double* atom_sigma_t;
  const double* atom_sigma_t_scaling;
  const double* amplitude_scaling;
  const double* central_weight;
} tgx_soap_hypers;

/* ---- context ------------------------------------------------------------ */
/* gpu_context_init: bind a HIP device, create the context's stream. */
int tgx_init(int device, tgx_ctx** ctx);
/* gpu_context_finalize: free every device buffer and the stream. */
int tgx_finalize(tgx_ctx* ctx);
/* Message of the last failing call on this thread (never NULL). */
const char* tgx_last_error(void);
/* gpu_memory_budget_init: free and total device memory in GiB. */
int tgx_memory_budget(tgx_ctx* ctx, double* free_gib, double* total_gib);
/* The device memory the FP64 plain path may hold for one block (the role of
 * max_Gbytes_per_process, src/neighbors.f90:637-778): a block whose
 * materialised dq/dr and per-centre rows exceed it runs in batches of
 * consecutive centres that fit. 0 (the default) means 80% of the device. */
int tgx_set_memory_budget(tgx_ctx* ctx, double gib);
/* Number of HIP devices visible (the rank -> device map of gpu_context_init). */
int tgx_device_count(int* n);
/* Name of the device the context is bound to. */
int tgx_device_name(tgx_ctx* ctx, char* buf, int buflen);

/* ---- system and neighbour list ---------------------------------------- */
/* Upload positions (3, n_sites_sc) column-major, the integer species code of
 * every supercell atom, and the supercell box (a, b, c as three 3-vectors). The
 * first n_sites atoms are the unit cell (src/read_files.f90:320-367 order). */
int tgx_set_system(tgx_ctx* ctx, int n_sites, int n_sites_sc, const double* positions, const int* species_sc,
                   const double* box);
/* number_of_unit_cells_for_given_cutoff (host only, no context): the repetitions
 * along a, b, c (box = a, b, c as three 3-vectors) after which the supercell's
 * opposite faces are at least 2 rcut apart; (1, 1, 1) when the cell already holds
 * a cutoff sphere. */
int tgx_number_of_unit_cells(const double* box, double rcut, int* indices);
/* The supercell of read_xyz (host only): positions (3, n_sites) and species
 * repeated indices[0] x indices[1] x indices[2] times, copy (i, j, k), i slowest,
 * shifted by i a + j b + k c, into positions_sc (3, n_sites * prod(indices)) and
 * species_sc; box_sc = (indices[0] a, indices[1] b, indices[2] c). The first
 * n_sites atoms are the cell itself, as tgx_set_system expects. */
int tgx_build_supercell(int n_sites, const double* positions, const int* species, const double* box,
                        const int* indices, double* positions_sc, int* species_sc, double* box_sc);
/* Replace positions (host pointer, same shape as in tgx_set_system). */
int tgx_set_positions(tgx_ctx* ctx, const double* positions);
/* Replace positions from a device pointer holding (3, n_sites_sc) doubles. */
int tgx_set_positions_device(tgx_ctx* ctx, const double* d_positions);
/* build_neighbors_list on the device, in the reference's order. indices is the
 * supercell repetition count (1,1,1 when the box holds a cutoff sphere);
 * do_list may be NULL (all sites). Writes the number of pairs. */
int tgx_build_neighbors_list(tgx_ctx* ctx, double rcut_max, const int* indices, const int* do_list,
                             int64_t* n_atom_pairs);
/* build_neighbors_list with rebuild_neighbors_list = .false.: the resident list
 * keeps its pairs and order, and rjs/xyz are recomputed from the current
 * positions with get_distance (thetas/phis on the next tgx_get_neighbors_list).
 * A list built for a supercell (indices > 1) must be rebuilt instead. */
int tgx_refresh_neighbors_list(tgx_ctx* ctx);
/* Copy the list to host arrays (any may be NULL): n_neigh(n_sites),
 * neighbors_list(P) 1-based supercell indices, rjs(P), thetas(P), phis(P),
 * xyz(3,P), neighbor_species(P). */
int tgx_get_neighbors_list(tgx_ctx* ctx, int* n_neigh, int* neighbors_list, double* rjs, double* thetas,
                           double* phis, double* xyz, int* neighbor_species);

/* ---- models -------------------------------------------------------------- */
/* One soap_turbo block and its sparse set: species_types holds the species
 * codes of the block's species list; central_species is a 1-based entry of it
 * (0 = every atom is a centre). Qs is (dim, n_sparse) column-major, alphas
 * (n_sparse). The sparse set is uploaded once and stays resident. */
int tgx_soap_model_create(tgx_ctx* ctx, const tgx_soap_hypers* hypers, const int* species_types,
                          int central_species, double delta, double zeta, double e0, int n_sparse, int dim,
                          const double* Qs, const double* alphas, int* model_id);
/* The same for a compressed block (compress_soap; soap_turbo_hypers%compress_P_*
 * as src/read_files.f90:5162-5211 fills them): dim is the compressed size, and
 * each of the compress_P_nonzero entries adds compress_P_el times component
 * compress_P_j of the raw power spectrum to component compress_P_i (1-based)
 * before normalisation. Descriptors, derivatives, energies and forces of the
 * block are those of the compressed descriptor; it runs on the FP64 paths. */
int tgx_soap_model_create_compressed(tgx_ctx* ctx, const tgx_soap_hypers* hypers, const int* species_types,
                                     int central_species, double delta, double zeta, double e0, int n_sparse,
                                     int dim, const double* Qs, const double* alphas, int compress_P_nonzero,
                                     const int* compress_P_i, const int* compress_P_j,
                                     const double* compress_P_el, int* model_id);
/* soap_backend_begin / soap_backend_end: select the block the next
 * tgx_get_gap_soap / tgx_get_soap / tgx_get_soap_energy_and_forces call uses. */
int tgx_soap_backend_begin(tgx_ctx* ctx, int model_id);
int tgx_soap_backend_end(tgx_ctx* ctx);
/* Descriptor size of a model (N(N+1)/2 (l_max+1), N = sum alpha_max). */
int tgx_soap_model_dim(tgx_ctx* ctx, int model_id, int* dim);

/* get_gap_soap for the current block over the resident neighbour list:
 * energies(n_sites) and forces(3, n_sites) are ASSIGNED for the atoms this
 * block touches (src/gap_interface.f90:422-431) and left alone elsewhere;
 * virial(3,3) is written. path: TGX_PATH_FP64_PLAIN or TGX_PATH_FP64_ADJOINT or
 * TGX_PATH_MIXED. n_sites_block / n_pairs_block may be NULL. */
int tgx_get_gap_soap(tgx_ctx* ctx, int path, int do_forces, double* energies, double* forces, double* virial,
                     int* n_sites_block, int64_t* n_pairs_block);
/* The same for the centres among sites i_beg..i_end (1-based, inclusive): one
 * batch of the reference's loop over this_i_beg:this_i_end
 * (src/turbogap.f90:1281-1344). */
int tgx_get_gap_soap_sites(tgx_ctx* ctx, int path, int do_forces, int i_beg, int i_end, double* energies,
                           double* forces, double* virial, int* n_sites_block, int64_t* n_pairs_block);
/* get_soap for the current block: soap(dim, n_sites_block) in the block's
 * central-site order and, when soap_cart_der is not NULL, soap_cart_der(3, dim,
 * n_pairs_block) in the block's pair order (self pair first per site). Sizes
 * come from a preceding tgx_get_gap_soap or tgx_soap_block_size call. */
int tgx_soap_block_size(tgx_ctx* ctx, int* n_sites_block, int64_t* n_pairs_block);
int tgx_get_soap(tgx_ctx* ctx, double* soap, double* soap_cart_der);
/* get_soap on a list the caller marshalled, soap_turbo's call shape
 * (src/gap_interface.f90:320-325), with the current block's hypers: n_sites
 * centres, n_neigh(n_sites) entries each (self pair first), per pair the
 * 0-based index of its species in the block's species list (-1: a species the
 * block does not see), rjs(P) and xyz(3, P). Writes soap(dim, n_sites) and, with
 * do_derivatives, soap_cart_der(3, dim, P). */
int tgx_get_soap_list(tgx_ctx* ctx, int n_sites, const int* n_neigh, int64_t n_pairs, const int* pair_species,
                      const double* rjs, const double* xyz, int do_derivatives, double* soap, double* soap_cart_der);
/* local_property_predict (src/local_properties.f90:174-297) with the current
 * block as the property model (its e0 is V0): V (n_sites_block) =
 * max(0, delta^2 sum_s alpha_s K_is^zeta + V0) and, with do_derivatives,
 * V_der (3, n_pairs_block) = dV_i/dq . soap_cart_der(:, :, p), zero where
 * V_i = 0, on the FP64 plain path. Sizes as tgx_soap_block_size gives them. */
int tgx_local_property_predict(tgx_ctx* ctx, int do_derivatives, double* V, double* V_der, int* n_sites_block,
                               int64_t* n_pairs_block);
/* get_soap_energy_and_forces for the current block, host inputs: soap(n_soap,
 * n_sites), soap_der(3, n_soap, n_pairs), neighbors_list 1-based local indices,
 * xyz(3, n_pairs); outputs energies(n_sites), forces(3, n_sites0), virial(3,3).
 * delta and zeta0 must equal the block's (TGX_ERR_ARG otherwise). */
int tgx_get_soap_energy_and_forces(tgx_ctx* ctx, int n_sites, int n_soap, int64_t n_pairs, const double* soap,
                                   const double* soap_der, double delta, double zeta0, double e0,
                                   const int* n_neigh, const int* neighbors_list, const double* xyz, int n_sites0,
                                   int do_forces, double* energies, double* forces, double* virial);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tgx_get_soap_energy_and_forces on DEVICE arrays (same
 * shapes and meaning) that calls no hipMalloc, hipFree or allocating routine.
 * Its work buffers live in a 64 MiB device array that every call on a GPU shares
 * (held until the process ends); inputs whose work space exceeds it, or more
 * than 16,384 sites (n_sites or n_sites0), are rejected with TGX_ERR_ARG: split
 * the sites over several calls. The work is enqueued on the context's stream
 * and the call returns without waiting (tgx_synchronize waits). Calls that
 * overlap on one GPU (several contexts or host threads) race on the shared
 * array. Energies and forces equal tgx_get_soap_energy_and_forces's bit for bit;
 * the virial is a fixed-order device sum. */
int tgx_get_soap_energy_and_forces_noalloc(tgx_ctx* ctx, int n_sites, int n_soap, int64_t n_pairs,
                                           const double* d_soap, const double* d_soap_der, double delta,
                                           double zeta0, double e0, const int* d_n_neigh,
                                           const int* d_neighbors_list, const double* d_xyz, int n_sites0,
                                           int do_forces, double* d_energies, double* d_forces, double* d_virial);

/* ---- whole-step evaluation ------------------------------------------------ */
/* E, F and virial of every model on the resident positions (neighbour list must
 * have been built). Per-atom energies include each model's e0 for its central
 * atoms. Results stay on the device until tgx_get_results. */
int tgx_compute(tgx_ctx* ctx, int path, double* energy, double* virial);
/* Copy energies(n_sites) and forces(3, n_sites) to the host (either may be NULL). */
int tgx_get_results(tgx_ctx* ctx, double* energies, double* forces);
/* Device pointer of the resident forces (3, n_sites) and energies (n_sites). */
int tgx_results_device(tgx_ctx* ctx, double** d_forces, double** d_energies);

/* ---- surrogate (no counterpart in the reference) ------------------------- */
/* A per-block MLP, dim -> hidden -> hidden -> 1 with SiLU, trained by the
 * library's own kernels on (q, E_i, dE_i/dq) samples that an FP64 tgx_compute
 * produced, and run with BF16 weights and FP32 accumulation by
 * tgx_compute(TGX_PATH_SURROGATE). Forces go through the same adjoint kernel as
 * the GAP's with Qss = -dE/dq. */
typedef struct tgx_mlp tgx_mlp;
/* hidden and max_batch must be multiples of 128. */
int tgx_mlp_create(tgx_ctx* ctx, int dim, int hidden, int max_batch, uint64_t seed, tgx_mlp** mlp);
int tgx_mlp_destroy(tgx_mlp* mlp);
/* Append every stride-th centre (from offset) of the model's last FP64
 * tgx_compute to the device-resident dataset; writes the dataset size. */
int tgx_mlp_add_samples(tgx_mlp* mlp, int model_id, int stride, int offset, int* n_total);
/* Mean and standard deviation of the energies of the first n_use samples. */
int tgx_mlp_dataset_stats(tgx_mlp* mlp, int n_use, double* e_mean, double* e_std);
/* Targets are trained as (E - e_shift)/e_scale and g/e_scale. */
int tgx_mlp_set_normalization(tgx_mlp* mlp, double e_shift, double e_scale);
int tgx_mlp_get_normalization(tgx_mlp* mlp, double* e_shift, double* e_scale);
/* Adam (beta 0.9/0.999) over shuffled batches of samples [0, n_train):
 * loss = w_energy mean (E - E_ref)^2 + w_grad mean |P (g - g_ref)|^2 with
 * P = 1 - q q^T, in normalised units. Writes the last epoch's mean losses. */
int tgx_mlp_train(tgx_mlp* mlp, int n_train, int epochs, int batch, float lr, float w_energy, float w_grad,
                  uint64_t seed, double* loss_e, double* loss_g);
/* Errors on samples [begin, end): energy MAE and RMSE (eV), RMS of |P (g - g_ref)| (eV). */
int tgx_mlp_evaluate(tgx_mlp* mlp, int begin, int end, double* e_mae, double* e_rmse, double* g_rmse);
/* Host-array forms for tests: q (batch, dim) FP32; outputs E (batch) and dE/dq (batch, dim). */
int tgx_mlp_forward_host(tgx_mlp* mlp, int batch, const float* q, float* e, float* g);
/* One batch's losses and weight gradients (batch a multiple of 128); targets
 * in normalised units. Gradient layout as tgx_mlp_get_weights. */
int tgx_mlp_gradients_host(tgx_mlp* mlp, int batch, const float* q, const float* e_ref, const float* g_ref,
                           float w_energy, float w_grad, float* dW1, float* db1, float* dW2, float* db2,
                           float* dw3, float* db3, double* loss_e, double* loss_g);
/* Energy plus force loss on whole structures (the configuration's "energy
 * plus force loss"): the centres of blocks model_ids[0..n_blocks) of the
 * context's current system, each block one batch of its network nets[b],
 * after a whole-block FP64 plain evaluation with forces (tgx_compute with
 * TGX_PATH_FP64_PLAIN). Loss = w_energy sum_blocks mean_i ((E_i -
 * e_shift)/e_scale - E_hat_i)^2 + w_force / n_atoms sum_j |F_hat_j - F_j|^2,
 * with the force error
 * of every block together (eV/A) formed from the networks' dE/dq and the
 * block's dq/dr. Leaves each network's weight gradients for
 * tgx_mlp_get_gradients / tgx_mlp_adam_step; overwrites the context's pair
 * forces (the computed results stay). */
int tgx_mlp_force_gradients(tgx_ctx* ctx, int n_blocks, tgx_mlp* const* nets, const int* model_ids, double w_energy,
                            double w_force, double* loss_energy, double* loss_force);
/* The weight gradients of the network's last backward pass (layout of
 * tgx_mlp_get_weights), and one Adam step (beta 0.9/0.999) with them. */
int tgx_mlp_get_gradients(tgx_mlp* mlp, float* dW1, float* db1, float* dW2, float* db2, float* dw3, float* db3);
int tgx_mlp_adam_step(tgx_mlp* mlp, float lr);
/* FP32 master weights, row-major: W1 (hidden, dim), b1 (hidden), W2 (hidden,
 * hidden), b2 (hidden), w3 (hidden), b3 (1). */
int tgx_mlp_get_weights(tgx_mlp* mlp, float* W1, float* b1, float* W2, float* b2, float* w3, float* b3);
int tgx_mlp_set_weights(tgx_mlp* mlp, const float* W1, const float* b1, const float* W2, const float* b2,
                        const float* w3, const float* b3);
/* Copy a network in as ensemble member `member` of model_id's surrogate. */
int tgx_surrogate_install(tgx_ctx* ctx, int model_id, int member, tgx_mlp* mlp);
/* Use n_members per model; centres whose member energies spread by more than
 * threshold eV are re-evaluated by the FP64 GAP (energy and gradient). */
int tgx_surrogate_gate(tgx_ctx* ctx, int n_members, double threshold);
/* Of the last surrogate compute: fraction of centres gated, largest spread (eV). */
int tgx_surrogate_stats(tgx_ctx* ctx, double* gated_fraction, double* max_spread);

/* ---- pair and triplet GAP families on the resident list ------------------ */
/* distance_2b: centres of species1 or species2 with a neighbour of the other
 * (species codes as in tgx_set_system), r < rcut, cosine cutoff over the last
 * `buffer` Angstrom; E_i = e0 + delta^2 sum_s alpha_s cutoff_s fcut
 * exp(-(r - Qs_s)^2 / (2 sigma^2)) over the centre's pairs. */
int tgx_distance_2b_create(tgx_ctx* ctx, int species1, int species2, double rcut, double buffer, double delta,
                           double sigma, double e0, int n_sparse, const double* Qs, const double* alphas,
                           const double* cutoff, int* model_id);
/* core_pot: the tabulated pair potential V(x) (n nodes, ascending; end slopes
 * yp1, ypn, >= 0.99e30 for a natural end) as a cubic spline; nodes beyond
 * `cutoff` are dropped and those in [cutoff - buffer, cutoff] damped
 * (core_pot_cutoff / core_pot_buffer; 1e10 and 1 by default). E_i = 1/2 sum V(r)
 * over the centre's pairs with r < max(x). */
int tgx_core_pot_create(tgx_ctx* ctx, int species1, int species2, int n, const double* x, const double* V, double yp1,
                        double ypn, double cutoff, double buffer, int* model_id);
/* angle_3b: centres of species_center; every unordered pair of neighbours with
 * species (species1, species2) in either order and both distances < rcut (cosine
 * cutoff over the last `buffer` Angstrom on each); the triplet descriptor
 * q = (r_ij + r_ik, (r_ij - r_ik)^2, r_jk) against the sparse points Qs (row-major
 * n_sparse x 3) scaled by sigma[3]; kernel_type "exp" (exp(-r^2/2)) or "pol"
 * (cov_pp, d = 3, q = 1); E_i = e0 + fc_ij fc_ik sum_s 2 delta^2 alpha_s cutoff_s
 * k(r_s). Forces act on the centre and on both neighbours (their home atoms).
 * A centre may have at most 1024 such neighbours. */
int tgx_angle_3b_create(tgx_ctx* ctx, int species_center, int species1, int species2, double rcut, double buffer,
                        double delta, const double* sigma, double e0, const char* kernel_type, int n_sparse,
                        const double* Qs, const double* alphas, const double* cutoff, int* model_id);
/* Energies (n_sites), forces (3, n_sites) and virial (3, 3) of one pair or
 * triplet model on the resident list, as the reference's subroutines return
 * them (any output may be NULL). distance_2b and core_pot forces act on the
 * centres only, as in the reference; angle_3b needs every atom's own list. */
int tgx_pair_energy_and_forces(tgx_ctx* ctx, int model_id, double* energies, double* forces, double* virial);

/* ---- MD (src/md.f90:226-345, src/turbogap_md.f90:160-470) --------------- */
/* Velocities (3, n_sites) and masses (n_sites) in eV fs^2/A^2 live on the device. */
int tgx_md_set_state(tgx_ctx* ctx, const double* velocities, const double* masses);
int tgx_md_get_state(tgx_ctx* ctx, double* positions, double* velocities);
/* velocity_verlet (md.f90:226-286) with the resident forces and forces_prev. */
int tgx_velocity_verlet(tgx_ctx* ctx, double dt, int first_step);
/* fix_atom (3, n_sites), nonzero = that Cartesian component is fixed: velocity_verlet
 * then zeroes its velocity (after the first step) and leaves its position. NULL
 * (or all zero) frees every component. */
int tgx_md_set_fixed(tgx_ctx* ctx, const int* fix_atom);
/* Instantaneous pressure and pressure tensor in bar from the virial (3, 3) of the
 * last force evaluation and the temperature: (kB (N - 1) T + tr(W)/3) / V and
 * W / V + kB (N - 1) T / V on the diagonal, V the box volume. */
int tgx_md_pressure(tgx_ctx* ctx, const double* virial, double temperature, double* pressure,
                    double* pressure_tensor);
/* berendsen_barostat: box vectors and positions scaled by
 * (1 + dt/tau 4.5e-5/gamma (P - P0))^(1/3), P the mean of the pressure tensor's
 * diagonal (sym "iso") or per Cartesian component from its diagonal ("diag");
 * pressures in bar. The new box (a, b, c) goes to box_out (9, may be NULL). The
 * resident list then needs a rebuild or refresh. */
int tgx_berendsen_barostat(tgx_ctx* ctx, double P0, const double* P, const char* sym, double tau, double gamma,
                           double dt, double* box_out);
/* E_kin and instantaneous T = 2/(3 (N-1) kB) E_kin (turbogap_md.f90:281-295). */
int tgx_kinetic(tgx_ctx* ctx, double* e_kin, double* temperature);
/* berendsen_thermostat (md.f90:289-313). */
int tgx_berendsen(tgx_ctx* ctx, double T0, double T, double tau, double dt);
/* remove_cm_vel (md.f90:316-345), on the device: the centre-of-mass velocity from
 * fixed-order partial sums, subtracted without a host round trip. */
int tgx_remove_cm_vel(tgx_ctx* ctx);
/* wrap positions back into the orthorhombic box. */
int tgx_wrap_pbc(tgx_ctx* ctx);
/* The neighbors_buffer test of turbogap_md.f90:451-471, after a velocity-Verlet
 * step: adds this step's displacement (minimum image, so wrapping is not motion)
 * to each atom's total since the last list build and sets *rebuild when an atom
 * has moved more than neighbors_buffer / 2. (The reference compares the squared
 * displacement with buffer / 2, which lets two atoms approach each other by more
 * than the buffer before a rebuild.) reset = 1 after a rebuild zeroes the totals. */
int tgx_md_displacement(tgx_ctx* ctx, double neighbors_buffer, int reset, int* rebuild);
/* box_scaling: positions += (gamma - I) positions / n_steps and the box vectors set to
 * f . (a0, b0, c0), f = I + (gamma - I)(i_step + 1)/n_steps, where (a0, b0, c0) is the
 * box when i_step was 0. gamma is the reference's box_scaling_factor(3, 3) row-major
 * (gamma[3 r + c] = box_scaling_factor(r + 1, c + 1)); the new box goes to box_out
 * (9, may be NULL). */
int tgx_box_scaling(tgx_ctx* ctx, const double* gamma, int i_step, int n_steps, double* box_out);
/* variable_time_step, from the device velocities and forces: the step new_dt at
 * which the largest |v dt + F/(2m) dt^2| over atoms first crosses target_pos_step,
 * searched from *dt in steps of 1% (down while too large, up while too small);
 * then *dt = min(new_dt, dt0) when init, else min(dt tau_dt/(tau_dt + dt - new_dt), dt0).
 * An error when 16384 steps of 1% do not reach the target (no motion at all). */
int tgx_variable_time_step(tgx_ctx* ctx, int init, double target_pos_step, double tau_dt, double dt0, double* dt);
/* variable_time_step_adaptive from the device velocities and forces: per atom the
 * smallest of xmax/|v|, sqrt(2 xmax m/|F|) and (emax > 0) emax/(|v||F|), an atom
 * without velocity or force taking those of the last atom before it that has one
 * (else the starting step: dt0 when init, else *dt), scaled down by xmax/d when the
 * displacement d = |v dt + F/(2m) dt^2| exceeds xmax; then *dt = min(smallest, tmax)
 * (the reference's tmin bound is overwritten, so tmin only has to be valid), and
 * min(that, dt0) when init. Errors as the reference's checks: xmax, tmin, tmax > 0,
 * tmax > tmin. */
int tgx_adaptive_time_step(tgx_ctx* ctx, int init, double tmin, double tmax, double xmax, double emax, double dt0,
                           double* dt);
/* optimize = "gd": one relaxation step in place of velocity Verlet, from the device
 * forces and the total energy at the current positions. Velocities are zeroed. On
 * first_step the positions and forces are kept and the step is
 * max_opt_step / max_i |F_i| (or the step the last line search accepted, for the
 * context's life); while that line search runs, a step that fails the Armijo test
 * E <= E0 - gamma/2 sum F^2 goes back to the kept positions and forces (restored = 1;
 * the device forces are replaced too) with half the step. After it, the
 * Barzilai-Borwein step |s.y|/|y|^2 with s and y the changes of positions (minimum
 * image) and forces. Fixed components stay put. Orthorhombic boxes; gamma_out and
 * restored may be NULL. */
int tgx_gradient_descent(tgx_ctx* ctx, double energy, double max_opt_step, int first_step, double* gamma_out,
                         int* restored);
/* thermostat = "bussi": velocities scaled by sqrt(K'/K), K the device E_kin and K'
 * the Bussi-Donadio-Parrinello resample of it (resamplekin) towards T0 with
 * ndeg = 3 N - 3, relaxation time tau and step dt. The caller supplies the random
 * draws: gauss ~ N(0, 1) and chi2 ~ chi-squared with ndeg - 1 degrees of freedom
 * (resamplekin's gasdev() and sumnoises(ndeg - 1)). K = 0 zeroes the velocities.
 * K' goes to e_kin_out (may be NULL). */
int tgx_bussi_thermostat(tgx_ctx* ctx, double T0, double tau, double dt, double gauss, double chi2,
                         double* e_kin_out);
/* No counterpart in the reference (a driver shortcut): after the step's forces,
 * tgx_remove_cm_vel (remove_cm != 0), tgx_velocity_verlet, tgx_kinetic,
 * tgx_berendsen to T0 (T0 <= 0: none), tgx_wrap_pbc and, with
 * neighbors_buffer > 0, the tgx_md_displacement test in one call with one
 * synchronisation; E_kin and T are those after the Verlet update, as
 * tgx_kinetic returns them. */
int tgx_md_advance(tgx_ctx* ctx, double dt, int first_step, int remove_cm, double T0, double tau,
                   double neighbors_buffer, double* e_kin, double* temperature, int* rebuild);

/* ---- profiling ----------------------------------------------------------- */
/* Record HIP events around every kernel launch on the context's stream. While
 * on, the surrogate path's blocks run one after the other. */
int tgx_set_profiling(tgx_ctx* ctx, int on);
/* Execution option with no counterpart in the reference: tgx_compute runs blocks
 * of distinct central species concurrently, each on a stream of its own.
 * 1 (the default): on the surrogate and FP64 paths, the mixed path's blocks in
 * turn (there the overlap measured slower); 2: on every path; 0: every path in
 * turn on the context's stream. Results are the same bits either way. */
int tgx_set_block_streams(tgx_ctx* ctx, int on);
/* Launch count and total milliseconds of kernels whose name contains `name`. */
int tgx_kernel_time(tgx_ctx* ctx, const char* name, int* count, double* total_ms);
int tgx_reset_profile(tgx_ctx* ctx);
int tgx_synchronize(tgx_ctx* ctx);
/* HBM reference measurement (no counterpart in the reference): copies `bytes`
 * between two device buffers with a streaming kernel (k_hbm_copy), `reps` times,
 * and reports the read + write bandwidth of the fastest copy in GB/s. The
 * kernel also calibrates the FETCH_SIZE / WRITE_SIZE counters. */
int tgx_hbm_probe(tgx_ctx* ctx, size_t bytes, int reps, double* gbps);

/* ---- synthetic inputs (host only; no GPU needed) -------------------------- */
/* Random sequential addition of n points in an orthorhombic box with minimum
 * distance dmin (periodic). Returns TGX_ERR_ARG if the packing fails. */
int tgx_generate_rsa(int n, const double* box_lengths, double dmin, uint64_t seed, double* positions);
/* Seeded Fisher-Yates permutation of 0..n-1. */
int tgx_generate_permutation(int n, uint64_t seed, int* perm);
/* Seeded normal deviates with the given standard deviation. */
int tgx_generate_normal(int n, double stddev, uint64_t seed, double* out);

#ifdef __cplusplus
}
#endif

#endif /* TGX_H */
