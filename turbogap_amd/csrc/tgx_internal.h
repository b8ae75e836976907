// Internal declarations shared by the libtgx translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <cstring>

#include "../../include/tgx.h"

// Limits of this build. Checked on the host before any launch.
#define TGX_MAX_SPECIES 4
#define TGX_MAX_NB 16    // radial functions per species
#define TGX_MAX_NRAD 32  // stacked radial functions
#define TGX_MAX_L 10
#define TGX_MAX_Q 64     // radial quadrature nodes per species
#define TGX_PAD 32       // descriptor / sparse padding granule

void tgx_set_error(const char* fmt, ...);

#define TGX_CHECK_HIP(call)                                                                     \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      tgx_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return TGX_ERR_HIP;                                                                       \
    }                                                                                           \
  } while (0)

#define TGX_REQUIRE(cond, ...)      \
  do {                              \
    if (!(cond)) {                  \
      tgx_set_error(__VA_ARGS__);   \
      return TGX_ERR_ARG;           \
    }                               \
  } while (0)

#define TGX_TRY(call)          \
  do {                         \
    int s_ = (call);           \
    if (s_ != TGX_OK) return s_; \
  } while (0)

static inline int64_t tgx_round_up(int64_t x, int64_t g) { return (x + g - 1) / g * g; }

// Growable device buffer.
struct dbuf {
  void* p = nullptr;
  size_t bytes = 0;
  // A buffer that has outgrown its allocation once (list and block sizes
  // drift during MD) is regrown with 25% headroom, so that a slowly rising
  // size does not free and reallocate it, synchronising the device, on every
  // new maximum; exact size if the headroom does not fit.
  int ensure(size_t n) {
    if (n <= bytes && p) return TGX_OK;
    const bool regrow = p != nullptr;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t want = n < 256 ? 256 : n;
    hipError_t e = hipErrorOutOfMemory;
    if (regrow) {
      e = hipMalloc(&p, want + want / 4);
      if (e == hipSuccess) want += want / 4;
      else (void)hipGetLastError();
    }
    if (e != hipSuccess) e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      tgx_set_error("hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
      p = nullptr;
      return TGX_ERR_NOMEM;
    }
    bytes = want;
    return TGX_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Per-species SOAP parameters as the kernels see them (lengths in Å unless
// marked as scaled by rcut_hard).
struct SpeciesParams {
  double rc;         // rcut_hard
  double inv_rc;
  double sig_r;      // atom_sigma_r (Å)
  double sig_r_sc;   // atom_sigma_r_scaling (1/Å)
  double sig_t;      // atom_sigma_t (Å)
  double sig_t_sc;   // atom_sigma_t_scaling (1/Å)
  double amp_a;      // amplitude_scaling
  double gs;         // global_scaling
  int nb;            // alpha_max
  int off;           // first stacked radial index
  int nq;            // quadrature nodes
  int pad_;
};

struct SoapParams {
  int n_species;
  int L;     // l_max
  int NLM;   // (L+1)^2
  int N;     // stacked radial functions
  int D;     // descriptor size
  int Dpad;  // padded row stride of q
  SpeciesParams sp[TGX_MAX_SPECIES];
  const double* B;    // [species][TGX_MAX_NB][TGX_MAX_Q]
  const double* xq;   // [species][TGX_MAX_Q]
  const double* cen;  // [species][TGX_MAX_NB] central-atom coefficients (weights and Y00 folded in)
  const int* dtab;    // [D] packed n | n' << 8 | l << 16
  const double* ynorm;  // [(L+1)^2] real spherical harmonic constants
};

struct tgx_model_t {
  // hypers (host copies)
  int n_species = 0, l_max = 0, basis = 0, radial_enhancement = 0;
  int nq_soft = 32, nq_buffer = 16;
  std::vector<int> alpha_max, species_types;
  std::vector<double> rcut_hard, rcut_soft, nf, global_scaling, sig_r, sig_r_sc, sig_t, sig_t_sc, amp_a, cw;
  int central_species = 0;
  double delta = 1.0, zeta = 2.0, e0 = 0.0;
  int N = 0, D = 0, Dpad = 0, S = 0, Spad = 0;
  double rcut_max = 0.0;
  SoapParams P{};
  // device copies of the tables
  dbuf B, xq, cen, dtab, ynorm, rtab;  // rtab: FP32 Hermite table of R_n, dR_n/dr
  dbuf low_idx, low_coef;              // lowering table of the solid harmonics [NLM][3][2]
  dbuf raise_idx, raise_coef;          // its transpose onto degree l_max - 1 and below [L^2][3][4]
  bool have_raise = false;
  // sparse set: Qs[Spad][Dpad], QsT[Dpad][Spad], alphas[Spad] (FP64) and BF16 copies
  dbuf Qs, QsT, alpha, Qs_bf, QsT_bf, alpha_f;
  // per-step block state (device)
  int n_sites_b = 0;
  int64_t n_pairs_b = 0;
  dbuf site_atom;   // int[n_sites_b] global 0-based atom of every centre
  dbuf site_beg;    // int64[n_sites_b+1] pair offsets
  dbuf pr_gidx;     // int64[n_pairs_b] global pair index of every kept pair
  dbuf pr_nbr;      // int[n_pairs_b] folded 0-based neighbour atom
  dbuf pr_spec;     // int[n_pairs_b] descriptor species (0-based) or -1
  dbuf pr_r;        // double[n_pairs_b]
  dbuf pr_xyz;      // double[3 n_pairs_b]
  dbuf q64, q32, qbf;   // descriptors [n_sites_pad][Dpad]
  dbuf pnorm;       // double[n_sites] |p|
  dbuf e_site;      // double[n_sites] energies
  dbuf qss64;       // [n_sites_pad][Dpad] FP64 -dE/dq
  dbuf qss32;       // FP32 -dE/dq
  dbuf kd64, kdbf;  // [n_sites_pad][Spad] K^(zeta-1) alpha (scaled)
  dbuf epart;       // energy partials [n_sites_pad][n_tiles]
  dbuf soap_der;    // [n_pairs_b][D][3] FP64 (plain path)
  dbuf gate_pn, gate_e;  // surrogate fallback: |p| of the gated sites, saved energies
  dbuf gate_stat, gate_list;  // surrogate gate: {count, max spread bits}, gated sites
  dbuf pbasis;      // [n_pairs_b][ldb] FP32 pair basis: A_l Y_lm | R_n | dR_n (FP32 paths)
  dbuf fpair;       // [n_pairs_b][3] pair forces
  dbuf sel_scratch; // block selection: flags, pair counts and their scans
  dbuf scan_tmp;    // block totals of the selection scans
  dbuf pair_site;   // int[n_pairs_b]: centre of every kept pair (plain force pass)
  int64_t* sel_host = nullptr;  // pinned: the selection's site and pair totals
  int sel_lo = 0, sel_hi = -1;  // centres are taken from atoms [sel_lo, sel_hi) only (sel_hi < 0: all)
  bool der_whole = false;       // soap_der holds dq/dr of every centre of the block (last FP64 plain evaluation)
  dbuf mlp_ws;      // surrogate work space (activations, member energies)
  // SOAP compression (compress_soap): q_c = P p / |P p| from the full power
  // spectrum p. D and Dpad above are then the compressed sizes, P.D and P.Dpad
  // the full ones. P as CSR over compressed rows (cp_*) and over full
  // components (cpt_*), 0-based.
  bool compress = false;
  dbuf cp_row, cp_col, cp_el, cpt_col, cpt_row, cpt_el;
  dbuf qf64;        // [n_sites_pad][P.Dpad] full normalised descriptors of a compressed block
  dbuf qssf64;      // [n_sites_pad][P.Dpad] -dE/dq of the full descriptors (FP64 adjoint)
  dbuf derf;        // [n_pairs_b][P.D][3] full dq/dr (FP64 plain)
  dbuf cnorm;       // double[n_sites] |P q| of the normalised full q
  hipStream_t stream = nullptr;  // the block's own stream when blocks run concurrently
  hipEvent_t done = nullptr;     // recorded on that stream after the block's last launch
};

struct prof_rec { std::string name; hipEvent_t a, b; };

// Entry-point prologue: a non-null context, its device current.
#define TGX_CTX_CHECK(ctx)                                      \
  do {                                                          \
    if (!(ctx)) {                                               \
      tgx_set_error("%s: null context", __func__);              \
      return TGX_ERR_ARG;                                       \
    }                                                           \
    TGX_CHECK_HIP(hipSetDevice((ctx)->device));                 \
  } while (0)

struct tgx_pair_model;  // distance_2b / angle_3b / core_pot model (tgx_pair.hip)
void tgx_pair_models_release(tgx_ctx* ctx);

struct tgx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // system
  int n_sites = 0, n_sites_sc = 0;
  double box[9] = {0};
  std::vector<int> species_sc_h;
  dbuf pos, species_sc;
  // neighbour list
  double rcut_max = 0.0;
  int64_t n_pairs = 0;
  dbuf n_neigh, nl_off, nlist, rjs, xyz, thetas, phis, nspec, rev;
  dbuf cell_of, cell_cnt, cell_start, cell_atoms, scan_tmp, do_list;
  bool have_list = false;
  bool have_angles = false;  // thetas / phis of the resident list computed
  dbuf cell_pos;             // positions in cell order (cell-list build)
  dbuf nbr_hits;             // cell slots of each atom's first hits (count pass, read by the fill)
  // models
  std::vector<tgx_model_t*> models;
  int current = -1;
  std::vector<tgx_pair_model*> pair_models;
  // results
  dbuf energies, forces, fglob, virial_part, scratch, red;
  dbuf fres;  // per-atom force residual of a force-loss training step
  int results_n = 0;  // atoms the energies / forces buffers hold results for (0: none)
  double mem_budget = 0.0;  // device bytes the FP64 plain path may hold for one block (0: 80% of the device)
  // md
  int md_n = 0;       // atoms of the MD state (velocities, masses, previous step); 0: none
  dbuf vel, mass, forces_prev, pos_prev;
  dbuf md_disp, md_flag;      // displacement since the last list build, rebuild flag
  dbuf md_fix;                 // unsigned char[3 n]: fixed Cartesian components
  bool md_fixed = false;
  double md_box0[9] = {0};     // box at step 0 of a box-scaling run
  bool md_box0_set = false;
  // gradient descent: positions and forces of its first step, then the line
  // search and Barzilai-Borwein state (gd_initialized outlives a restart)
  dbuf md_gd0;
  dbuf md_cm;                  // centre-of-mass partial sums
  bool gd_started = false, gd_backtracking = false, gd_initialized = false;
  double gd_energy0 = 0.0, gd_gamma_prev = 0.0, gd_m_prev = 0.0, gd_gamma_back0 = 0.0;
  int* md_flag_host = nullptr;  // pinned
  double dt_prev = 0.0;
  // surrogate gate
  int gate_members = 1;
  double gate_threshold = 1e30;
  double gated_fraction = 0.0, max_spread = 0.0;  // over every block of the last tgx_compute
  int64_t gate_sites = 0, gate_gated = 0;
  int* gate_host = nullptr;  // pinned: {count, max spread bits} of every block
  double* red_host = nullptr;  // pinned: virial and energy partials of the force gather
  int gate_host_n = 0;
  dbuf gate_mask;
  hipEvent_t ev_blocks = nullptr;  // recorded on the context stream before concurrent blocks start
  int block_streams = 1;           // blocks of distinct central species on streams of their own: 0 off, 1 default, 2 every path
  // profiling
  bool profiling = false;
  std::vector<prof_rec> prof;
  std::vector<hipEvent_t> ev_pool;
};

// Profiled launch: records events around the launch when profiling is on.
int tgx_prof_begin(tgx_ctx* ctx, const char* name, int* slot);
int tgx_prof_end(tgx_ctx* ctx, int slot);

#define TGX_LAUNCH(ctx, name, kernel, grid, block, shmem, ...)                       \
  do {                                                                              \
    int slot_ = -1;                                                                 \
    if ((ctx)->profiling) TGX_TRY(tgx_prof_begin((ctx), (name), &slot_));           \
    hipLaunchKernelGGL(kernel, grid, block, shmem, (ctx)->stream, __VA_ARGS__);     \
    hipError_t le_ = hipGetLastError();                                             \
    if (le_ != hipSuccess) {                                                        \
      tgx_set_error("launch of %s failed: %s", (name), hipGetErrorString(le_));     \
      return TGX_ERR_HIP;                                                           \
    }                                                                               \
    if (slot_ >= 0) TGX_TRY(tgx_prof_end((ctx), slot_));                            \
  } while (0)

// Host-side helpers shared between units (tgx_host.cpp).
int tgx_build_radial_tables(tgx_model_t* m, std::vector<double>& B, std::vector<double>& xq,
                            std::vector<double>& cen);
void tgx_gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);

// Kernels' host entry points (one per translation unit).
int tgx_neighbors_run(tgx_ctx* ctx, double rcut_max, const int* indices, const int* do_list);
int tgx_neighbor_angles(tgx_ctx* ctx);
int tgx_neighbors_refresh(tgx_ctx* ctx);
int tgx_local_property_finish(tgx_ctx* ctx, tgx_model_t* m, bool with_der, double** d_V, double** d_vder);
int tgx_block_select(tgx_ctx* ctx, tgx_model_t* m);
// The two halves of tgx_block_select: counts and scans with the totals copied to
// pinned memory (no wait), then, after the caller has synchronised the stream,
// allocation and fill.
int tgx_block_select_count(tgx_ctx* ctx, tgx_model_t* m);
// Host copies of the count half's per-atom centre flags and pair offsets
// (n and n + 1 values), after the stream has been synchronised.
int tgx_block_select_counts_host(tgx_ctx* ctx, tgx_model_t* m, std::vector<int>& flag, std::vector<int64_t>& pair_off);
int tgx_block_select_fill(tgx_ctx* ctx, tgx_model_t* m);
int tgx_soap_descriptors(tgx_ctx* ctx, tgx_model_t* m, int path, bool with_der);
int tgx_gap_contract(tgx_ctx* ctx, tgx_model_t* m, int path, bool do_forces);
int tgx_plain_pair_forces(tgx_ctx* ctx, tgx_model_t* m);
// The FP64 contraction of ns sites on explicit buffers (no allocation): q, kd,
// epart as the model's padded layouts, energies to e_site, and Qss when qss is
// not NULL.
int tgx_gap_contract_f64(tgx_ctx* ctx, const tgx_model_t* m, int ns, const double* q, double* kd, double* epart,
                         double* e_site, double* qss);
int tgx_plain_pair_forces_raw(tgx_ctx* ctx, int ns, const int64_t* site_beg, int64_t np, int D, int Dpad,
                              const double* qss, const double* der, int* pair_site, double* fpair);
// Device address of this device's allocation-free work array (set up by tgx_init)
// and its size in bytes.
int tgx_scratch_setup(int device);
void* tgx_scratch_ptr(int device);
size_t tgx_scratch_bytes();
int tgx_adjoint_pair_forces(tgx_ctx* ctx, tgx_model_t* m, int path);
// plain pair forces of the block with the given -dE/dq rows ([n_sites_b][Dpad]) into ctx->fglob
int tgx_plain_pair_forces_q(tgx_ctx* ctx, tgx_model_t* m, const double* qss);
// compressed blocks: q64 / soap_der from qf64 / derf, and qssf64 from qss64
int tgx_compress_descriptors(tgx_ctx* ctx, tgx_model_t* m, bool with_der);
int tgx_compress_gradient(tgx_ctx* ctx, tgx_model_t* m);
int tgx_gather_forces(tgx_ctx* ctx, double* energy_out, double* virial_out);
// F (device, n x 3) and the pair virial (host, 9, optional) from per-entry pair
// forces fglob (device, n_pairs x 3) on the resident list.
int tgx_gather_pair_forces(tgx_ctx* ctx, const double* fglob, double* F, double* virial_out,
                           const double* d_energies = nullptr, double* energy_out = nullptr);
int tgx_zero(tgx_ctx* ctx, void* p, size_t bytes);
int tgx_surrogate_energy_grad(tgx_ctx* ctx, tgx_model_t* m);
int tgx_device_sum(tgx_ctx* ctx, const double* d_x, int64_t n, double* out_host);
int tgx_gemm_bf16_nt(tgx_ctx* ctx, const char* name, int M, int N, int K, const __bf16* A, int lda, const __bf16* B,
                     int ldb, float* C, int ldc, double scale);
int tgx_f32_to_bf16(tgx_ctx* ctx, const float* x, __bf16* y, int64_t n);

// SiLU and its first two derivatives (FP32), shared by the MLP kernels and the
// fused GEMM epilogues. The sigmoid takes the hardware reciprocal (v_rcp_f32,
// within one ulp; 1 / inf = 0 and 1 / 1 = 1 exactly) instead of the IEEE
// division, which costs nine instructions per element in the epilogues.
__device__ __forceinline__ float tgx_sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float tgx_silu(float z) { return z * tgx_sigm(z); }
__device__ __forceinline__ float tgx_silu1(float z) {
  const float s = tgx_sigm(z);
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ float tgx_silu2(float z) {
  const float s = tgx_sigm(z);
  return s * (1.0f - s) * (2.0f + z * (1.0f - 2.0f * s));
}

// Batched BF16 NT GEMM with a fused epilogue: for z < batch,
//   acc = A_z[M][K] . B_z[N][K]^T (FP32 accumulation), X_z = X + z * sX, and
//   TGX_EPI_STORE_F32 : C_z (FP32)  = scale * acc
//   TGX_EPI_SILU      : v = acc + bias_z[col]; C_z (BF16) = silu(v); aux_z (FP32) = silu'(v)
//   TGX_EPI_SCALE_AUX : C_z (BF16)  = acc * aux_z * (zscale ? zscale[z] : scale)
//   TGX_EPI_ACT2      : v = acc + bias_z[col]; C_z (BF16) = w3_z[col] silu'(v);
//                       epart_z[row][col / 128] (FP64) = sum over the 128-column
//                       tile of w3_z[col] silu(v)
// M, N multiples of 128, K of 64, leading dimensions and strides in elements,
// multiples of 8 for A and B.
enum { TGX_EPI_STORE_F32 = 0, TGX_EPI_SILU = 2, TGX_EPI_SCALE_AUX = 3, TGX_EPI_ACT2 = 4 };
struct tgx_gemm_desc {
  int epi = TGX_EPI_STORE_F32;
  int M = 0, N = 0, K = 0, batch = 1;
  const __bf16* A = nullptr;
  int lda = 0;
  int64_t sA = 0;
  const __bf16* B = nullptr;
  int ldb = 0;
  int64_t sB = 0;
  void* C = nullptr;
  int ldc = 0;
  int64_t sC = 0;
  double scale = 1.0;
  const float* bias = nullptr;
  int64_t sBias = 0;
  float* aux = nullptr;
  int ldaux = 0;
  int64_t sAux = 0;
  const float* zscale = nullptr;
  const float* w3 = nullptr;  // TGX_EPI_ACT2
  int64_t sW3 = 0;
  double* epart = nullptr;    // TGX_EPI_ACT2: [batch][M][N / 128]
  int64_t sEpart = 0;
};
int tgx_gemm_bf16(tgx_ctx* ctx, const char* name, const tgx_gemm_desc& d);
