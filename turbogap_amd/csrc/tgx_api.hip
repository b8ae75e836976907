// C ABI of libtgx (declared in include/tgx.h): context, system, neighbour list,
// models, and the SOAP/GAP evaluation entry points.
#include <cmath>
#include <algorithm>
#include <cstdio>
#include <cstring>

#include "tgx_internal.h"
#include "tgx_util.h"

int tgx_energy_scatter(tgx_ctx* ctx, tgx_model_t* m);
int tgx_block_seen(tgx_ctx* ctx, tgx_model_t* m, int* d_seen, bool clear);
int tgx_radial_table_build(tgx_ctx* ctx, tgx_model_t* m);


// ---- profiling ----------------------------------------------------------------
int tgx_prof_begin(tgx_ctx* ctx, const char* name, int* slot) {
  prof_rec r;
  r.name = name;
  for (hipEvent_t* e : {&r.a, &r.b}) {
    if (!ctx->ev_pool.empty()) {
      *e = ctx->ev_pool.back();
      ctx->ev_pool.pop_back();
    } else {
      TGX_CHECK_HIP(hipEventCreate(e));
    }
  }
  TGX_CHECK_HIP(hipEventRecord(r.a, ctx->stream));
  ctx->prof.push_back(r);
  *slot = (int)ctx->prof.size() - 1;
  return TGX_OK;
}

int tgx_prof_end(tgx_ctx* ctx, int slot) {
  TGX_CHECK_HIP(hipEventRecord(ctx->prof[slot].b, ctx->stream));
  return TGX_OK;
}

extern "C" int tgx_set_profiling(tgx_ctx* ctx, int on) {
  TGX_CTX_CHECK(ctx);
  ctx->profiling = on != 0;
  return TGX_OK;
}

extern "C" int tgx_set_block_streams(tgx_ctx* ctx, int on) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(on >= 0 && on <= 2, "tgx_set_block_streams: mode %d (0 off, 1 default, 2 every path)", on);
  ctx->block_streams = on;
  return TGX_OK;
}

extern "C" int tgx_reset_profile(tgx_ctx* ctx) {
  TGX_CTX_CHECK(ctx);
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->prof) { ctx->ev_pool.push_back(r.a); ctx->ev_pool.push_back(r.b); }
  ctx->prof.clear();
  return TGX_OK;
}

extern "C" int tgx_kernel_time(tgx_ctx* ctx, const char* name, int* count, double* total_ms) {
  TGX_CTX_CHECK(ctx);
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  int c = 0;
  double t = 0.0;
  for (auto& r : ctx->prof) {
    if (name && name[0] && r.name.find(name) == std::string::npos) continue;
    float ms = 0.f;
    TGX_CHECK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    ++c;
    t += ms;
  }
  if (count) *count = c;
  if (total_ms) *total_ms = t;
  return TGX_OK;
}

extern "C" int tgx_synchronize(tgx_ctx* ctx) {
  TGX_CTX_CHECK(ctx);
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

extern "C" int tgx_hbm_probe(tgx_ctx* ctx, size_t bytes, int reps, double* gbps) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(gbps && reps >= 1 && bytes >= 16 && bytes % 16 == 0 && bytes <= ((size_t)64 << 30),
              "tgx_hbm_probe: need 16 <= bytes <= 64 GiB, a multiple of 16, and reps >= 1");
  dbuf a, b;
  TGX_TRY(a.ensure(bytes));
  TGX_TRY(b.ensure(bytes));
  TGX_TRY(tgx_zero(ctx, a.p, bytes));
  hipEvent_t e0, e1;
  TGX_CHECK_HIP(hipEventCreate(&e0));
  TGX_CHECK_HIP(hipEventCreate(&e1));
  double best = 1e30;
  int st = TGX_OK;
  for (int r = 0; r < reps && st == TGX_OK; ++r) {
    (void)hipEventRecord(e0, ctx->stream);
    st = tgx_hbm_copy(ctx, a.p, b.p, bytes);
    (void)hipEventRecord(e1, ctx->stream);
    if (hipEventSynchronize(e1) != hipSuccess) st = TGX_ERR_HIP;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    best = std::min(best, (double)ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(ctx->stream);
  a.release();
  b.release();
  TGX_TRY(st);
  *gbps = 2.0 * (double)bytes / (best * 1e-3) / 1e9;
  return TGX_OK;
}

// ---- context ------------------------------------------------------------------
extern "C" int tgx_init(int device, tgx_ctx** out) {
  TGX_REQUIRE(out, "tgx_init: null output");
  int n = 0;
  TGX_CHECK_HIP(hipGetDeviceCount(&n));
  TGX_REQUIRE(device >= 0 && device < n, "tgx_init: device %d not present (%d visible)", device, n);
  TGX_CHECK_HIP(hipSetDevice(device));
  // the device address of the allocation-free entry point's work array, once per device
  TGX_TRY(tgx_scratch_setup(device));
  tgx_ctx* ctx = new tgx_ctx();
  ctx->device = device;
  hipError_t e = hipStreamCreate(&ctx->stream);
  if (e != hipSuccess) {
    delete ctx;
    tgx_set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
    return TGX_ERR_HIP;
  }
  *out = ctx;
  return TGX_OK;
}

static void free_model(tgx_model_t* m) {
  dbuf* bufs[] = {&m->B, &m->xq, &m->cen, &m->dtab, &m->ynorm, &m->rtab, &m->Qs, &m->QsT, &m->alpha, &m->Qs_bf, &m->QsT_bf, &m->alpha_f,
                  &m->site_atom, &m->site_beg, &m->pr_gidx, &m->pr_nbr, &m->pr_spec, &m->pr_r, &m->pr_xyz, &m->q64,
                  &m->q32, &m->qbf, &m->pnorm, &m->e_site, &m->qss64, &m->qss32, &m->kd64, &m->kdbf, &m->epart,
                  &m->soap_der, &m->fpair, &m->pbasis, &m->low_idx, &m->low_coef, &m->raise_idx, &m->raise_coef,
                  &m->gate_pn, &m->gate_e,
                  &m->gate_stat, &m->gate_list, &m->sel_scratch, &m->mlp_ws, &m->scan_tmp, &m->pair_site,
                  &m->cp_row, &m->cp_col, &m->cp_el, &m->cpt_col, &m->cpt_row, &m->cpt_el, &m->qf64, &m->qssf64,
                  &m->derf, &m->cnorm};
  for (dbuf* b : bufs) b->release();
  if (m->sel_host) (void)hipHostFree(m->sel_host);
  if (m->done) (void)hipEventDestroy(m->done);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

void tgx_surrogate_release(tgx_ctx* ctx);

extern "C" int tgx_finalize(tgx_ctx* ctx) {
  if (!ctx) return TGX_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  tgx_surrogate_release(ctx);
  tgx_pair_models_release(ctx);
  for (auto* m : ctx->models) free_model(m);
  dbuf* bufs[] = {&ctx->pos, &ctx->species_sc, &ctx->n_neigh, &ctx->nl_off, &ctx->nlist, &ctx->rjs, &ctx->xyz,
                  &ctx->thetas, &ctx->phis, &ctx->nspec, &ctx->rev, &ctx->cell_of, &ctx->cell_cnt, &ctx->cell_start,
                  &ctx->cell_atoms, &ctx->scan_tmp, &ctx->do_list, &ctx->energies, &ctx->forces, &ctx->fglob, &ctx->fres,
                  &ctx->virial_part, &ctx->scratch, &ctx->red, &ctx->vel, &ctx->mass, &ctx->forces_prev,
                  &ctx->pos_prev, &ctx->gate_mask, &ctx->cell_pos, &ctx->nbr_hits, &ctx->md_disp, &ctx->md_flag,
                  &ctx->md_fix, &ctx->md_gd0, &ctx->md_cm};
  for (dbuf* b : bufs) b->release();
  if (ctx->gate_host) (void)hipHostFree(ctx->gate_host);
  if (ctx->md_flag_host) (void)hipHostFree(ctx->md_flag_host);
  if (ctx->red_host) (void)hipHostFree(ctx->red_host);
  if (ctx->ev_blocks) (void)hipEventDestroy(ctx->ev_blocks);
  for (auto& r : ctx->prof) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return TGX_OK;
}

extern "C" int tgx_memory_budget(tgx_ctx* ctx, double* free_gib, double* total_gib) {
  TGX_CTX_CHECK(ctx);
  size_t f = 0, t = 0;
  TGX_CHECK_HIP(hipMemGetInfo(&f, &t));
  if (free_gib) *free_gib = (double)f / (1024.0 * 1024.0 * 1024.0);
  if (total_gib) *total_gib = (double)t / (1024.0 * 1024.0 * 1024.0);
  return TGX_OK;
}

extern "C" int tgx_device_name(tgx_ctx* ctx, char* buf, int buflen) {
  TGX_CTX_CHECK(ctx);
  hipDeviceProp_t p;
  TGX_CHECK_HIP(hipGetDeviceProperties(&p, ctx->device));
  snprintf(buf, (size_t)buflen, "%s (%s)", p.name, p.gcnArchName);
  return TGX_OK;
}

// ---- system / neighbours ----------------------------------------------------------
extern "C" int tgx_set_system(tgx_ctx* ctx, int n_sites, int n_sites_sc, const double* positions,
                              const int* species_sc, const double* box) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(n_sites > 0 && n_sites_sc >= n_sites && positions && species_sc && box,
              "tgx_set_system: bad arguments (n_sites=%d, n_sites_sc=%d)", n_sites, n_sites_sc);
  if (n_sites != ctx->n_sites) {
    // buffers sized for the previous system's atoms: no MD state, fixed-atom
    // mask or results carry over to a system of another size
    ctx->md_n = 0;
    ctx->md_fixed = false;
    ctx->results_n = 0;
  }
  ctx->n_sites = n_sites;
  ctx->n_sites_sc = n_sites_sc;
  for (int k = 0; k < 9; ++k) ctx->box[k] = box[k];
  ctx->species_sc_h.assign(species_sc, species_sc + n_sites_sc);
  TGX_TRY(ctx->pos.ensure(sizeof(double) * 3 * (size_t)n_sites_sc));
  TGX_TRY(ctx->species_sc.ensure(sizeof(int) * (size_t)n_sites_sc));
  TGX_CHECK_HIP(hipMemcpyAsync(ctx->pos.p, positions, sizeof(double) * 3 * (size_t)n_sites_sc, hipMemcpyHostToDevice,
                               ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(ctx->species_sc.p, species_sc, sizeof(int) * (size_t)n_sites_sc, hipMemcpyHostToDevice,
                               ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_list = false;
  ctx->md_box0_set = false;
  // a new system starts a new descent (the accepted step of the last one is kept)
  ctx->gd_started = false;
  ctx->gd_backtracking = false;
  return TGX_OK;
}

extern "C" int tgx_set_positions(tgx_ctx* ctx, const double* positions) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(ctx->n_sites_sc > 0 && positions, "tgx_set_positions: no system or null positions");
  TGX_CHECK_HIP(hipMemcpyAsync(ctx->pos.p, positions, sizeof(double) * 3 * (size_t)ctx->n_sites_sc,
                               hipMemcpyHostToDevice, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

extern "C" int tgx_set_positions_device(tgx_ctx* ctx, const double* d_positions) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(ctx->n_sites_sc > 0 && d_positions, "tgx_set_positions_device: no system or null positions");
  TGX_CHECK_HIP(hipMemcpyAsync(ctx->pos.p, d_positions, sizeof(double) * 3 * (size_t)ctx->n_sites_sc,
                               hipMemcpyDeviceToDevice, ctx->stream));
  return TGX_OK;
}

extern "C" int tgx_build_neighbors_list(tgx_ctx* ctx, double rcut_max, const int* indices, const int* do_list,
                                        int64_t* n_atom_pairs) {
  TGX_CTX_CHECK(ctx);
  TGX_TRY(tgx_neighbors_run(ctx, rcut_max, indices, do_list));
  if (n_atom_pairs) *n_atom_pairs = ctx->n_pairs;
  return TGX_OK;
}

extern "C" int tgx_refresh_neighbors_list(tgx_ctx* ctx) {
  TGX_CTX_CHECK(ctx);
  return tgx_neighbors_refresh(ctx);
}

extern "C" int tgx_get_neighbors_list(tgx_ctx* ctx, int* n_neigh, int* neighbors_list, double* rjs, double* thetas,
                                      double* phis, double* xyz, int* neighbor_species) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(ctx->have_list, "tgx_get_neighbors_list: no list built");
  const size_t P = (size_t)ctx->n_pairs;
  if (thetas || phis) TGX_TRY(tgx_neighbor_angles(ctx));
  auto cp = [&](void* dst, const dbuf& src, size_t bytes) -> int {
    if (dst && bytes) TGX_CHECK_HIP(hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
    return TGX_OK;
  };
  TGX_TRY(cp(n_neigh, ctx->n_neigh, sizeof(int) * (size_t)ctx->n_sites));
  TGX_TRY(cp(neighbors_list, ctx->nlist, sizeof(int) * P));
  TGX_TRY(cp(rjs, ctx->rjs, sizeof(double) * P));
  TGX_TRY(cp(thetas, ctx->thetas, sizeof(double) * P));
  TGX_TRY(cp(phis, ctx->phis, sizeof(double) * P));
  TGX_TRY(cp(xyz, ctx->xyz, sizeof(double) * 3 * P));
  TGX_TRY(cp(neighbor_species, ctx->nspec, sizeof(int) * P));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

// ---- models -----------------------------------------------------------------------
static int copy_h(const double* src, int n, std::vector<double>& dst, const char* what) {
  TGX_REQUIRE(src, "tgx_soap_model_create: %s is NULL", what);
  dst.assign(src, src + n);
  return TGX_OK;
}

// A soap_turbo block. comp_nnz < 0: no compression, dim is the full power
// spectrum size; otherwise dim is the compressed size and (P_i, P_j, P_el) the
// comp_nnz entries (1-based) of the compression matrix.
static int create_model(tgx_ctx* ctx, const tgx_soap_hypers* h, const int* species_types, int central_species,
                        double delta, double zeta, double e0, int n_sparse, int dim, const double* Qs,
                        const double* alphas, int comp_nnz, const int* P_i, const int* P_j, const double* P_el,
                        int* model_id) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(h && species_types && model_id, "tgx_soap_model_create: null argument");
  TGX_REQUIRE(h->n_species >= 1 && h->n_species <= TGX_MAX_SPECIES, "n_species=%d outside 1..%d", h->n_species,
              TGX_MAX_SPECIES);
  TGX_REQUIRE(h->l_max >= 0 && h->l_max <= TGX_MAX_L, "l_max=%d outside 0..%d", h->l_max, TGX_MAX_L);
  TGX_REQUIRE(h->basis == 0 || h->basis == 1, "basis must be 0 (poly3) or 1 (poly3gauss)");
  TGX_REQUIRE(h->radial_enhancement >= 0 && h->radial_enhancement <= 2, "radial_enhancement must be 0, 1 or 2");
  TGX_REQUIRE(h->n_quad_soft >= 2 && h->n_quad_buffer >= 0 && h->n_quad_soft + h->n_quad_buffer <= TGX_MAX_Q,
              "quadrature nodes %d + %d outside 2..%d", h->n_quad_soft, h->n_quad_buffer, TGX_MAX_Q);
  TGX_REQUIRE(central_species >= 0 && central_species <= h->n_species, "central_species=%d outside 0..%d",
              central_species, h->n_species);
  for (int s = 0; s < h->n_species; ++s)
    for (int t = s + 1; t < h->n_species; ++t)
      TGX_REQUIRE(species_types[s] != species_types[t],
                  "species %d listed twice: species multiplicity above 1 is not supported", species_types[s]);
  tgx_model_t* m = new tgx_model_t();
  auto fail = [&](int code) { free_model(m); return code; };
  m->n_species = h->n_species;
  m->l_max = h->l_max;
  m->basis = h->basis;
  m->radial_enhancement = h->radial_enhancement;
  m->nq_soft = h->n_quad_soft;
  m->nq_buffer = h->n_quad_buffer;
  const int S = h->n_species;
  m->alpha_max.assign(h->alpha_max, h->alpha_max + S);
  m->species_types.assign(species_types, species_types + S);
  int st;
  if ((st = copy_h(h->rcut_hard, S, m->rcut_hard, "rcut_hard")) ||
      (st = copy_h(h->rcut_soft, S, m->rcut_soft, "rcut_soft")) || (st = copy_h(h->nf, S, m->nf, "nf")) ||
      (st = copy_h(h->global_scaling, S, m->global_scaling, "global_scaling")) ||
      (st = copy_h(h->atom_sigma_r, S, m->sig_r, "atom_sigma_r")) ||
      (st = copy_h(h->atom_sigma_r_scaling, S, m->sig_r_sc, "atom_sigma_r_scaling")) ||
      (st = copy_h(h->atom_sigma_t, S, m->sig_t, "atom_sigma_t")) ||
      (st = copy_h(h->atom_sigma_t_scaling, S, m->sig_t_sc, "atom_sigma_t_scaling")) ||
      (st = copy_h(h->amplitude_scaling, S, m->amp_a, "amplitude_scaling")) ||
      (st = copy_h(h->central_weight, S, m->cw, "central_weight")))
    return fail(st);
  int N = 0;
  for (int s = 0; s < S; ++s) {
    if (m->alpha_max[s] < 1 || m->alpha_max[s] > TGX_MAX_NB || (m->basis == 1 && m->alpha_max[s] < 2)) {
      tgx_set_error("alpha_max[%d]=%d outside %d..%d", s, m->alpha_max[s], m->basis == 1 ? 2 : 1, TGX_MAX_NB);
      return fail(TGX_ERR_ARG);
    }
    if (!(m->rcut_hard[s] > 0.0) || !(m->rcut_soft[s] > 0.0) || m->rcut_soft[s] > m->rcut_hard[s] ||
        !(m->sig_r[s] > 0.0) || !(m->sig_t[s] >= 0.0)) {
      tgx_set_error("species %d: need 0 < rcut_soft <= rcut_hard and atom_sigma_r > 0", s);
      return fail(TGX_ERR_ARG);
    }
    N += m->alpha_max[s];
  }
  if (N > TGX_MAX_NRAD) {
    tgx_set_error("sum of alpha_max = %d exceeds %d", N, TGX_MAX_NRAD);
    return fail(TGX_ERR_ARG);
  }
  m->N = N;
  const int d_full = N * (N + 1) / 2 * (m->l_max + 1);
  m->D = comp_nnz < 0 ? d_full : dim;
  if (dim != m->D) {
    tgx_set_error("sparse set dimension %d does not match the descriptor size %d", dim, m->D);
    return fail(TGX_ERR_ARG);
  }
  std::vector<int> c_row, c_col, ct_col, ct_row;
  std::vector<double> c_el, ct_el;
  if (comp_nnz >= 0) {
    // q_c(P_i) += P_el q(P_j) for every entry, as CSR over the compressed rows
    // (forward map) and CSR over the full components (its transpose); entries
    // keep their file order inside a row, so every sum has a fixed order
    if (dim < 1 || comp_nnz < 1 || !P_i || !P_j || !P_el) {
      tgx_set_error("compression: need dim >= 1 and at least one (P_i, P_j, P_el) entry");
      return fail(TGX_ERR_ARG);
    }
    c_row.assign((size_t)dim + 1, 0);
    ct_col.assign((size_t)d_full + 1, 0);
    for (int k = 0; k < comp_nnz; ++k) {
      if (P_i[k] < 1 || P_i[k] > dim || P_j[k] < 1 || P_j[k] > d_full) {
        tgx_set_error("compression entry %d: (P_i, P_j) = (%d, %d) outside 1..%d x 1..%d", k + 1, P_i[k], P_j[k], dim,
                      d_full);
        return fail(TGX_ERR_ARG);
      }
      ++c_row[(size_t)P_i[k]];
      ++ct_col[(size_t)P_j[k]];
    }
    for (int i = 0; i < dim; ++i) c_row[(size_t)i + 1] += c_row[(size_t)i];
    for (int j = 0; j < d_full; ++j) ct_col[(size_t)j + 1] += ct_col[(size_t)j];
    c_col.resize((size_t)comp_nnz);
    c_el.resize((size_t)comp_nnz);
    ct_row.resize((size_t)comp_nnz);
    ct_el.resize((size_t)comp_nnz);
    std::vector<int> fr(c_row.begin(), c_row.end() - 1), fc(ct_col.begin(), ct_col.end() - 1);
    for (int k = 0; k < comp_nnz; ++k) {
      const int i = P_i[k] - 1, j = P_j[k] - 1;
      c_col[(size_t)fr[(size_t)i]] = j;
      c_el[(size_t)fr[(size_t)i]++] = P_el[k];
      ct_row[(size_t)fc[(size_t)j]] = i;
      ct_el[(size_t)fc[(size_t)j]++] = P_el[k];
    }
    m->compress = true;
  }
  if (n_sparse < 1 || !Qs || !alphas) {
    tgx_set_error("empty sparse set");
    return fail(TGX_ERR_ARG);
  }
  m->central_species = central_species;
  m->delta = delta;
  m->zeta = zeta;
  m->e0 = e0;
  m->S = n_sparse;
  m->Dpad = (int)tgx_round_up(m->D, 128);
  m->Spad = (int)tgx_round_up(n_sparse, 128);
  m->rcut_max = *std::max_element(m->rcut_hard.begin(), m->rcut_hard.end());
  // kernel parameters
  SoapParams& P = m->P;
  P.n_species = S;
  P.L = m->l_max;
  P.NLM = (m->l_max + 1) * (m->l_max + 1);
  P.N = N;
  P.D = d_full;
  P.Dpad = (int)tgx_round_up(d_full, 128);
  int off = 0;
  for (int s = 0; s < S; ++s) {
    SpeciesParams& sp = P.sp[s];
    sp.rc = m->rcut_hard[s];
    sp.inv_rc = 1.0 / sp.rc;
    sp.sig_r = m->sig_r[s];
    sp.sig_r_sc = m->sig_r_sc[s];
    sp.sig_t = m->sig_t[s];
    sp.sig_t_sc = m->sig_t_sc[s];
    sp.amp_a = m->amp_a[s];
    sp.gs = m->global_scaling[s];
    sp.nb = m->alpha_max[s];
    sp.off = off;
    off += sp.nb;
  }
  std::vector<double> B, xq, cen;
  tgx_build_radial_tables(m, B, xq, cen);
  // constants of the real spherical harmonics, lm = l*l + l + m
  std::vector<double> ynorm((size_t)(m->l_max + 1) * (m->l_max + 1));
  for (int l = 0; l <= m->l_max; ++l)
    for (int mm = -l; mm <= l; ++mm) {
      const int am = mm < 0 ? -mm : mm;
      double ratio = 1.0;
      for (int k = l - am + 1; k <= l + am; ++k) ratio /= (double)k;
      double nrm = std::sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * ratio);
      double df = 1.0;
      for (int k = 1; k <= 2 * am - 1; k += 2) df *= (double)k;
      nrm *= df * ((am & 1) ? -1.0 : 1.0);
      if (am != 0) nrm *= M_SQRT2;
      ynorm[(size_t)l * l + l + mm] = nrm;
    }
  std::vector<int> dtab;
  for (int n = 0; n < N; ++n)
    for (int np = n; np < N; ++np)
      for (int l = 0; l <= m->l_max; ++l) dtab.push_back(n | (np << 8) | (l << 16));
  auto up = [&](dbuf& b, const void* src, size_t bytes) -> int {
    TGX_TRY(b.ensure(bytes));
    TGX_CHECK_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return TGX_OK;
  };
  // sparse set: Qs(dim, n_sparse) column-major is [s][d] row-major
  const int Dp = m->Dpad, Sp = m->Spad;
  std::vector<double> qs((size_t)Sp * Dp, 0.0), qst((size_t)Dp * Sp, 0.0), al(Sp, 0.0);
  std::vector<__bf16> qsb((size_t)Sp * Dp), qstb((size_t)Dp * Sp);
  std::vector<float> alf(Sp, 0.0f);
  for (int s = 0; s < n_sparse; ++s) {
    al[s] = alphas[s];
    alf[s] = (float)alphas[s];
    for (int d = 0; d < m->D; ++d) {
      const double v = Qs[(size_t)s * m->D + d];
      qs[(size_t)s * Dp + d] = v;
      qst[(size_t)d * Sp + s] = v;
    }
  }
  for (size_t k = 0; k < qs.size(); ++k) qsb[k] = (__bf16)(float)qs[k];
  for (size_t k = 0; k < qst.size(); ++k) qstb[k] = (__bf16)(float)qst[k];
  if ((st = up(m->B, B.data(), sizeof(double) * B.size())) || (st = up(m->xq, xq.data(), sizeof(double) * xq.size())) ||
      (st = up(m->cen, cen.data(), sizeof(double) * cen.size())) ||
      (st = up(m->dtab, dtab.data(), sizeof(int) * dtab.size())) ||
      (st = up(m->ynorm, ynorm.data(), sizeof(double) * ynorm.size())) ||
      (st = up(m->Qs, qs.data(), sizeof(double) * qs.size())) ||
      (st = up(m->QsT, qst.data(), sizeof(double) * qst.size())) ||
      (st = up(m->alpha, al.data(), sizeof(double) * al.size())) ||
      (st = up(m->Qs_bf, qsb.data(), sizeof(__bf16) * qsb.size())) ||
      (st = up(m->QsT_bf, qstb.data(), sizeof(__bf16) * qstb.size())) ||
      (st = up(m->alpha_f, alf.data(), sizeof(float) * alf.size())))
    return fail(st);
  if (m->compress &&
      ((st = up(m->cp_row, c_row.data(), sizeof(int) * c_row.size())) ||
       (st = up(m->cp_col, c_col.data(), sizeof(int) * c_col.size())) ||
       (st = up(m->cp_el, c_el.data(), sizeof(double) * c_el.size())) ||
       (st = up(m->cpt_col, ct_col.data(), sizeof(int) * ct_col.size())) ||
       (st = up(m->cpt_row, ct_row.data(), sizeof(int) * ct_row.size())) ||
       (st = up(m->cpt_el, ct_el.data(), sizeof(double) * ct_el.size()))))
    return fail(st);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    tgx_set_error("upload of the sparse set failed");
    return fail(TGX_ERR_HIP);
  }
  P.B = m->B.as<double>();
  P.xq = m->xq.as<double>();
  P.cen = m->cen.as<double>();
  P.dtab = m->dtab.as<int>();
  P.ynorm = m->ynorm.as<double>();
  if ((st = tgx_radial_table_build(ctx, m)) != TGX_OK) return fail(st);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    tgx_set_error("radial table build failed");
    return fail(TGX_ERR_HIP);
  }
  ctx->models.push_back(m);
  *model_id = (int)ctx->models.size() - 1;
  return TGX_OK;
}

extern "C" int tgx_soap_model_create(tgx_ctx* ctx, const tgx_soap_hypers* h, const int* species_types,
                                     int central_species, double delta, double zeta, double e0, int n_sparse,
                                     int dim, const double* Qs, const double* alphas, int* model_id) {
  return create_model(ctx, h, species_types, central_species, delta, zeta, e0, n_sparse, dim, Qs, alphas, -1, nullptr,
                      nullptr, nullptr, model_id);
}

extern "C" int tgx_soap_model_create_compressed(tgx_ctx* ctx, const tgx_soap_hypers* h, const int* species_types,
                                                int central_species, double delta, double zeta, double e0,
                                                int n_sparse, int dim, const double* Qs, const double* alphas,
                                                int compress_P_nonzero, const int* compress_P_i,
                                                const int* compress_P_j, const double* compress_P_el,
                                                int* model_id) {
  TGX_REQUIRE(compress_P_nonzero >= 1, "tgx_soap_model_create_compressed: compress_P_nonzero=%d", compress_P_nonzero);
  return create_model(ctx, h, species_types, central_species, delta, zeta, e0, n_sparse, dim, Qs, alphas,
                      compress_P_nonzero, compress_P_i, compress_P_j, compress_P_el, model_id);
}

static int get_model(tgx_ctx* ctx, int id, tgx_model_t** m) {
  TGX_REQUIRE(id >= 0 && id < (int)ctx->models.size(), "model id %d not defined", id);
  *m = ctx->models[id];
  return TGX_OK;
}

extern "C" int tgx_soap_model_dim(tgx_ctx* ctx, int model_id, int* dim) {
  TGX_CTX_CHECK(ctx);
  tgx_model_t* m;
  TGX_TRY(get_model(ctx, model_id, &m));
  if (dim) *dim = m->D;
  return TGX_OK;
}

extern "C" int tgx_soap_backend_begin(tgx_ctx* ctx, int model_id) {
  TGX_CTX_CHECK(ctx);
  tgx_model_t* m;
  TGX_TRY(get_model(ctx, model_id, &m));
  ctx->current = model_id;
  return TGX_OK;
}

extern "C" int tgx_soap_backend_end(tgx_ctx* ctx) {
  TGX_CTX_CHECK(ctx);
  ctx->current = -1;
  return TGX_OK;
}

static int current_model(tgx_ctx* ctx, tgx_model_t** m) {
  TGX_REQUIRE(ctx->current >= 0, "no model selected: call tgx_soap_backend_begin first");
  *m = ctx->models[ctx->current];
  return TGX_OK;
}

// One block's energies and pair forces (pair forces accumulated into fglob).
// Descriptors, contraction and pair forces of a block whose selection is done.
static int run_selected_block(tgx_ctx* ctx, tgx_model_t* m, int path, bool do_forces) {
  const bool plain = path == TGX_PATH_FP64_PLAIN;
  TGX_TRY(tgx_soap_descriptors(ctx, m, path, plain && do_forces));
  m->der_whole = plain && do_forces && m->sel_lo <= 0 && (m->sel_hi < 0 || m->sel_hi >= ctx->n_sites);
  TGX_TRY(tgx_gap_contract(ctx, m, path, do_forces));
  if (do_forces) {
    if (plain) TGX_TRY(tgx_plain_pair_forces(ctx, m));
    else TGX_TRY(tgx_adjoint_pair_forces(ctx, m, path));
  }
  return TGX_OK;
}

static int begin_step(tgx_ctx* ctx) {
  TGX_REQUIRE(ctx->have_list, "no neighbour list: call tgx_build_neighbors_list first");
  const int n = ctx->n_sites;
  const size_t P = (size_t)(ctx->n_pairs > 0 ? ctx->n_pairs : 1);
  TGX_TRY(ctx->energies.ensure(sizeof(double) * (size_t)n));
  TGX_TRY(ctx->forces.ensure(sizeof(double) * 3 * (size_t)n));
  TGX_TRY(ctx->fglob.ensure(sizeof(double) * 3 * P));
  TGX_TRY(tgx_zero(ctx, ctx->energies.p, sizeof(double) * (size_t)n));
  TGX_TRY(tgx_zero(ctx, ctx->fglob.p, sizeof(double) * 3 * P));
  ctx->results_n = n;
  return TGX_OK;
}

// ---- memory-bounded batches of a block (src/neighbors.f90:637-778) -------------
// Device bytes the FP64 plain path holds for ns centres and np pairs: the
// materialised dq/dr (3 x D FP64 per pair) and the pair data, and per centre
// q, Qss and K^(zeta-1) alpha rows; a compressed block holds the full-size
// dq/dr and q besides the compressed ones.
static double plain_block_bytes(const tgx_model_t* m, int64_t ns, int64_t np) {
  const double full_np = m->compress ? 24.0 * m->P.D : 0.0, full_ns = m->compress ? 8.0 * m->P.Dpad + 8.0 : 0.0;
  return (24.0 * m->D + full_np + 64.0) * (double)np + (8.0 * (2.0 * m->Dpad + m->Spad) + full_ns + 64.0) * (double)ns;
}

static int memory_budget_bytes(tgx_ctx* ctx, double* bytes) {
  if (ctx->mem_budget > 0.0) {
    *bytes = ctx->mem_budget;
    return TGX_OK;
  }
  size_t f = 0, t = 0;
  TGX_CHECK_HIP(hipMemGetInfo(&f, &t));
  *bytes = 0.8 * (double)t;
  return TGX_OK;
}

// After the count half of a block whose centres come from atoms [lo, hi): the
// atom ranges, in order, whose centres' plain-path work space fits the budget
// (get_number_of_atom_pairs's split of the centres into batches). One range
// when the whole block fits. A single centre larger than the budget is a
// range of its own.
static int plain_batches(tgx_ctx* ctx, tgx_model_t* m, int lo, int hi, std::vector<std::pair<int, int>>& out) {
  out.clear();
  double budget = 0.0;
  TGX_TRY(memory_budget_bytes(ctx, &budget));
  if (plain_block_bytes(m, m->sel_host[0], m->sel_host[1]) <= budget) {
    out.emplace_back(lo, hi);
    return TGX_OK;
  }
  std::vector<int> flag;
  std::vector<int64_t> off;
  TGX_TRY(tgx_block_select_counts_host(ctx, m, flag, off));
  int b = lo;
  int64_t ns = 0, np = 0;
  for (int i = lo; i < hi; ++i) {
    if (!flag[(size_t)i]) continue;
    const int64_t pi = off[(size_t)i + 1] - off[(size_t)i];
    if (ns > 0 && plain_block_bytes(m, ns + 1, np + pi) > budget) {
      out.emplace_back(b, i);
      b = i;
      ns = 0;
      np = 0;
    }
    ++ns;
    np += pi;
  }
  out.emplace_back(b, hi);
  return TGX_OK;
}

// A block's energies and pair forces over the centres of atoms [lo, hi),
// batched by the memory budget on the FP64 plain path. Energies are added into
// the context's per-atom energies and pair forces into fglob. When `sites` is
// given, the centres' atoms and energies are appended to it and `seen` marks
// (device) the atoms the pairs reach. Runs on ctx->stream.
struct BlockOut {
  std::vector<int>* atoms = nullptr;
  std::vector<double>* e = nullptr;
  int* d_seen = nullptr;
};

static int run_block_range(tgx_ctx* ctx, tgx_model_t* m, int path, bool do_forces, int lo, int hi, BlockOut* out,
                           int* n_sites_tot, int64_t* n_pairs_tot) {
  const int save_lo = m->sel_lo, save_hi = m->sel_hi;
  struct Restore {
    tgx_model_t* m;
    int lo, hi;
    ~Restore() { m->sel_lo = lo; m->sel_hi = hi; }
  } restore{m, save_lo, save_hi};
  m->sel_lo = lo;
  m->sel_hi = hi;
  TGX_TRY(tgx_block_select_count(ctx, m));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<std::pair<int, int>> batches;
  if (path == TGX_PATH_FP64_PLAIN && do_forces) TGX_TRY(plain_batches(ctx, m, lo, hi, batches));
  else batches.emplace_back(lo, hi);
  int ns_tot = 0;
  int64_t np_tot = 0;
  for (size_t k = 0; k < batches.size(); ++k) {
    if (batches.size() > 1) {  // this batch's selection
      m->sel_lo = batches[k].first;
      m->sel_hi = batches[k].second;
      TGX_TRY(tgx_block_select_count(ctx, m));
      TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    }
    TGX_TRY(tgx_block_select_fill(ctx, m));
    TGX_TRY(run_selected_block(ctx, m, path, do_forces));
    TGX_TRY(tgx_energy_scatter(ctx, m));
    ns_tot += m->n_sites_b;
    np_tot += m->n_pairs_b;
    if (out && out->atoms && m->n_sites_b > 0) {
      const size_t b0 = out->atoms->size();
      out->atoms->resize(b0 + (size_t)m->n_sites_b);
      out->e->resize(b0 + (size_t)m->n_sites_b);
      TGX_CHECK_HIP(hipMemcpyAsync(out->atoms->data() + b0, m->site_atom.p, sizeof(int) * (size_t)m->n_sites_b,
                                   hipMemcpyDeviceToHost, ctx->stream));
      TGX_CHECK_HIP(hipMemcpyAsync(out->e->data() + b0, m->e_site.p, sizeof(double) * (size_t)m->n_sites_b,
                                   hipMemcpyDeviceToHost, ctx->stream));
      TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (out && out->d_seen) TGX_TRY(tgx_block_seen(ctx, m, out->d_seen, false));
  }
  if (n_sites_tot) *n_sites_tot = ns_tot;
  if (n_pairs_tot) *n_pairs_tot = np_tot;
  return TGX_OK;
}

extern "C" int tgx_set_memory_budget(tgx_ctx* ctx, double gib) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(gib >= 0.0, "tgx_set_memory_budget: negative budget");
  ctx->mem_budget = gib * 1024.0 * 1024.0 * 1024.0;
  return TGX_OK;
}

// get_gap_soap over the centres of atoms [lo, hi) (0-based, hi exclusive).
static int get_gap_soap_impl(tgx_ctx* ctx, int path, int do_forces, int lo, int hi, double* energies,
                             double* forces, double* virial, int* n_sites_block, int64_t* n_pairs_block) {
  TGX_REQUIRE(path == TGX_PATH_FP64_PLAIN || path == TGX_PATH_FP64_ADJOINT || path == TGX_PATH_MIXED,
              "tgx_get_gap_soap: path %d not available here", path);
  tgx_model_t* m;
  TGX_TRY(current_model(ctx, &m));
  TGX_TRY(begin_step(ctx));
  const int n = ctx->n_sites;
  TGX_REQUIRE(lo >= 0 && lo <= hi && hi <= n, "tgx_get_gap_soap: sites [%d, %d) outside [0, %d)", lo, hi, n);
  std::vector<int> sa;
  std::vector<double> es;
  BlockOut out;
  out.atoms = &sa;
  out.e = &es;
  if (do_forces) {
    TGX_TRY(ctx->gate_mask.ensure(sizeof(int) * (size_t)n));
    TGX_TRY(tgx_zero(ctx, ctx->gate_mask.p, sizeof(int) * (size_t)n));
    out.d_seen = ctx->gate_mask.as<int>();
  }
  int ns_b = 0;
  int64_t np_b = 0;
  TGX_TRY(run_block_range(ctx, m, path, do_forces != 0, lo, hi, &out, &ns_b, &np_b));
  double vir[9] = {0};
  std::vector<int> seen;
  std::vector<double> F;
  if (do_forces) {
    TGX_TRY(tgx_gather_forces(ctx, nullptr, vir));
    seen.resize((size_t)n);
    F.resize(3 * (size_t)n);
    TGX_CHECK_HIP(hipMemcpyAsync(seen.data(), ctx->gate_mask.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost,
                                 ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(F.data(), ctx->forces.p, sizeof(double) * F.size(), hipMemcpyDeviceToHost, ctx->stream));
    TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  }
  // fold back like src/gap_interface.f90:422-431: assign, for touched atoms only
  if (energies)
    for (size_t s = 0; s < sa.size(); ++s) energies[sa[s]] = es[s];
  if (do_forces) {
    if (forces)
      for (int i = 0; i < n; ++i)
        if (seen[(size_t)i])
          for (int k = 0; k < 3; ++k) forces[3 * (size_t)i + k] = F[3 * (size_t)i + k];
    if (virial)
      for (int k = 0; k < 9; ++k) virial[k] = vir[k];
  } else if (virial) {
    for (int k = 0; k < 9; ++k) virial[k] = 0.0;
  }
  if (n_sites_block) *n_sites_block = ns_b;
  if (n_pairs_block) *n_pairs_block = np_b;
  return TGX_OK;
}

extern "C" int tgx_get_gap_soap(tgx_ctx* ctx, int path, int do_forces, double* energies, double* forces,
                                double* virial, int* n_sites_block, int64_t* n_pairs_block) {
  TGX_CTX_CHECK(ctx);
  return get_gap_soap_impl(ctx, path, do_forces, 0, ctx->n_sites, energies, forces, virial, n_sites_block,
                           n_pairs_block);
}

extern "C" int tgx_get_gap_soap_sites(tgx_ctx* ctx, int path, int do_forces, int i_beg, int i_end, double* energies,
                                      double* forces, double* virial, int* n_sites_block, int64_t* n_pairs_block) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(i_beg >= 1 && i_end >= i_beg - 1 && i_end <= ctx->n_sites,
              "tgx_get_gap_soap_sites: sites %d..%d outside 1..%d", i_beg, i_end, ctx->n_sites);
  return get_gap_soap_impl(ctx, path, do_forces, i_beg - 1, i_end, energies, forces, virial, n_sites_block,
                           n_pairs_block);
}

extern "C" int tgx_local_property_predict(tgx_ctx* ctx, int do_derivatives, double* V, double* V_der,
                                          int* n_sites_block, int64_t* n_pairs_block) {
  TGX_CTX_CHECK(ctx);
  tgx_model_t* m;
  TGX_TRY(current_model(ctx, &m));
  TGX_REQUIRE(ctx->have_list, "no neighbour list: call tgx_build_neighbors_list first");
  TGX_TRY(tgx_block_select(ctx, m));
  TGX_TRY(tgx_soap_descriptors(ctx, m, TGX_PATH_FP64_PLAIN, do_derivatives != 0));
  TGX_TRY(tgx_gap_contract(ctx, m, TGX_PATH_FP64_PLAIN, do_derivatives != 0));
  double *dV = nullptr, *dD = nullptr;
  TGX_TRY(tgx_local_property_finish(ctx, m, do_derivatives != 0, &dV, &dD));
  if (V && m->n_sites_b)
    TGX_CHECK_HIP(hipMemcpyAsync(V, dV, sizeof(double) * m->n_sites_b, hipMemcpyDeviceToHost, ctx->stream));
  if (V_der && do_derivatives && m->n_pairs_b)
    TGX_CHECK_HIP(hipMemcpyAsync(V_der, dD, sizeof(double) * 3 * (size_t)m->n_pairs_b, hipMemcpyDeviceToHost,
                                 ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (n_sites_block) *n_sites_block = m->n_sites_b;
  if (n_pairs_block) *n_pairs_block = m->n_pairs_b;
  return TGX_OK;
}

extern "C" int tgx_soap_block_size(tgx_ctx* ctx, int* n_sites_block, int64_t* n_pairs_block) {
  TGX_CTX_CHECK(ctx);
  tgx_model_t* m;
  TGX_TRY(current_model(ctx, &m));
  TGX_TRY(tgx_block_select(ctx, m));
  if (n_sites_block) *n_sites_block = m->n_sites_b;
  if (n_pairs_block) *n_pairs_block = m->n_pairs_b;
  return TGX_OK;
}

extern "C" int tgx_get_soap(tgx_ctx* ctx, double* soap, double* soap_cart_der) {
  TGX_CTX_CHECK(ctx);
  tgx_model_t* m;
  TGX_TRY(current_model(ctx, &m));
  TGX_TRY(tgx_block_select(ctx, m));
  TGX_TRY(tgx_soap_descriptors(ctx, m, TGX_PATH_FP64_PLAIN, soap_cart_der != nullptr));
  const int ns = m->n_sites_b;
  if (soap && ns)
    TGX_CHECK_HIP(hipMemcpy2DAsync(soap, sizeof(double) * m->D, m->q64.p, sizeof(double) * m->Dpad, sizeof(double) * m->D,
                                   ns, hipMemcpyDeviceToHost, ctx->stream));
  if (soap_cart_der && m->n_pairs_b)
    TGX_CHECK_HIP(hipMemcpyAsync(soap_cart_der, m->soap_der.p, sizeof(double) * 3 * m->D * (size_t)m->n_pairs_b,
                                 hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

// get_soap on a list the caller marshalled (soap_turbo's call shape,
// src/gap_interface.f90:320-325): the current block's hypers, centres with
// their pairs in order (self pair first), each pair's species as a 0-based
// entry of the block's species list (-1: not seen).
extern "C" int tgx_get_soap_list(tgx_ctx* ctx, int n_sites, const int* n_neigh, int64_t n_pairs,
                                 const int* pair_species, const double* rjs, const double* xyz, int do_derivatives,
                                 double* soap, double* soap_cart_der) {
  TGX_CTX_CHECK(ctx);
  tgx_model_t* m;
  TGX_TRY(current_model(ctx, &m));
  TGX_REQUIRE(n_sites >= 0 && n_pairs >= 0 && (n_sites == 0 || (n_neigh && soap)) &&
                  (n_pairs == 0 || (pair_species && rjs && xyz)) && (!do_derivatives || n_pairs == 0 || soap_cart_der),
              "tgx_get_soap_list: bad arguments");
  if (n_sites == 0) return TGX_OK;
  std::vector<int64_t> beg((size_t)n_sites + 1, 0), gidx((size_t)(n_pairs > 0 ? n_pairs : 1));
  std::vector<int> atom((size_t)n_sites), zero((size_t)(n_pairs > 0 ? n_pairs : 1), 0);
  for (int i = 0; i < n_sites; ++i) {
    TGX_REQUIRE(n_neigh[i] >= 1, "tgx_get_soap_list: site %d has no self entry", i + 1);
    beg[(size_t)i + 1] = beg[(size_t)i] + n_neigh[i];
    atom[(size_t)i] = i;
  }
  TGX_REQUIRE(beg[(size_t)n_sites] == n_pairs, "tgx_get_soap_list: sum of n_neigh (%lld) differs from n_pairs (%lld)",
              (long long)beg[(size_t)n_sites], (long long)n_pairs);
  for (int64_t p = 0; p < n_pairs; ++p) {
    TGX_REQUIRE(pair_species[p] >= -1 && pair_species[p] < m->n_species,
                "tgx_get_soap_list: pair %lld has species index %d", (long long)p + 1, pair_species[p]);
    gidx[(size_t)p] = p;
  }
  const size_t np = (size_t)(n_pairs > 0 ? n_pairs : 1);
  TGX_TRY(m->site_atom.ensure(sizeof(int) * (size_t)n_sites));
  TGX_TRY(m->site_beg.ensure(sizeof(int64_t) * beg.size()));
  TGX_TRY(m->pr_gidx.ensure(sizeof(int64_t) * np));
  TGX_TRY(m->pr_nbr.ensure(sizeof(int) * np));
  TGX_TRY(m->pr_spec.ensure(sizeof(int) * np));
  TGX_TRY(m->pr_r.ensure(sizeof(double) * np));
  TGX_TRY(m->pr_xyz.ensure(sizeof(double) * 3 * np));
  TGX_CHECK_HIP(hipMemcpyAsync(m->site_atom.p, atom.data(), sizeof(int) * atom.size(), hipMemcpyHostToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(m->site_beg.p, beg.data(), sizeof(int64_t) * beg.size(), hipMemcpyHostToDevice, ctx->stream));
  if (n_pairs > 0) {
    TGX_CHECK_HIP(hipMemcpyAsync(m->pr_gidx.p, gidx.data(), sizeof(int64_t) * (size_t)n_pairs, hipMemcpyHostToDevice,
                                 ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(m->pr_nbr.p, zero.data(), sizeof(int) * (size_t)n_pairs, hipMemcpyHostToDevice,
                                 ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(m->pr_spec.p, pair_species, sizeof(int) * (size_t)n_pairs, hipMemcpyHostToDevice,
                                 ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(m->pr_r.p, rjs, sizeof(double) * (size_t)n_pairs, hipMemcpyHostToDevice, ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(m->pr_xyz.p, xyz, sizeof(double) * 3 * (size_t)n_pairs, hipMemcpyHostToDevice,
                                 ctx->stream));
  }
  m->n_sites_b = n_sites;
  m->n_pairs_b = n_pairs;
  TGX_TRY(tgx_soap_descriptors(ctx, m, TGX_PATH_FP64_PLAIN, do_derivatives != 0 && n_pairs > 0));
  TGX_CHECK_HIP(hipMemcpy2DAsync(soap, sizeof(double) * m->D, m->q64.p, sizeof(double) * m->Dpad, sizeof(double) * m->D,
                                 (size_t)n_sites, hipMemcpyDeviceToHost, ctx->stream));
  if (do_derivatives && n_pairs > 0)
    TGX_CHECK_HIP(hipMemcpyAsync(soap_cart_der, m->soap_der.p, sizeof(double) * 3 * m->D * (size_t)n_pairs,
                                 hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

extern "C" int tgx_device_count(int* n) {
  TGX_REQUIRE(n, "tgx_device_count: null output");
  TGX_CHECK_HIP(hipGetDeviceCount(n));
  return TGX_OK;
}

extern "C" int tgx_get_soap_energy_and_forces(tgx_ctx* ctx, int n_sites, int n_soap, int64_t n_pairs,
                                              const double* soap, const double* soap_der, double delta,
                                              double zeta0, double e0, const int* n_neigh,
                                              const int* neighbors_list, const double* xyz, int n_sites0,
                                              int do_forces, double* energies, double* forces, double* virial) {
  TGX_CTX_CHECK(ctx);
  tgx_model_t* m;
  TGX_TRY(current_model(ctx, &m));
  TGX_REQUIRE(n_soap == m->D, "n_soap=%d does not match the model's descriptor size %d", n_soap, m->D);
  // delta and zeta are fixed when the block is created (its sparse set and
  // kernel are prepared with them): the caller's values must be the block's
  TGX_REQUIRE(std::fabs(delta - m->delta) <= 1e-12 * std::fabs(m->delta) &&
                  std::fabs(zeta0 - m->zeta) <= 1e-12 * std::fabs(m->zeta),
              "tgx_get_soap_energy_and_forces: delta=%g, zeta0=%g differ from the block's delta=%g, zeta=%g", delta,
              zeta0, m->delta, m->zeta);
  TGX_REQUIRE(n_sites >= 0 && soap && energies, "tgx_get_soap_energy_and_forces: bad arguments");
  if (do_forces)
    TGX_REQUIRE(soap_der && n_neigh && neighbors_list && xyz && forces && virial && n_sites0 > 0,
                "tgx_get_soap_energy_and_forces: forces need soap_der, lists, xyz and outputs");
  if (n_sites == 0) return TGX_OK;
  // stage the inputs as a block of the current model
  std::vector<int64_t> beg(n_sites + 1, 0), gidx;
  if (do_forces) {
    for (int i = 0; i < n_sites; ++i) beg[i + 1] = beg[i] + n_neigh[i];
    TGX_REQUIRE(beg[n_sites] == n_pairs, "sum of n_neigh (%lld) differs from n_pairs (%lld)", (long long)beg[n_sites],
                (long long)n_pairs);
    gidx.resize((size_t)n_pairs);
    for (int64_t p = 0; p < n_pairs; ++p) gidx[(size_t)p] = p;
  }
  m->n_sites_b = n_sites;
  m->n_pairs_b = do_forces ? n_pairs : 0;
  const size_t nsp = (size_t)tgx_round_up(n_sites, 128);
  TGX_TRY(m->q64.ensure(sizeof(double) * nsp * m->Dpad));
  TGX_TRY(tgx_zero(ctx, m->q64.p, sizeof(double) * nsp * m->Dpad));
  TGX_CHECK_HIP(hipMemcpy2DAsync(m->q64.p, sizeof(double) * m->Dpad, soap, sizeof(double) * m->D, sizeof(double) * m->D,
                                 n_sites, hipMemcpyHostToDevice, ctx->stream));
  const double e0_save = m->e0;
  m->e0 = e0;
  int st = tgx_gap_contract(ctx, m, TGX_PATH_FP64_PLAIN, do_forces != 0);
  m->e0 = e0_save;
  TGX_TRY(st);
  std::vector<double> fp;
  if (do_forces && n_pairs > 0) {
    TGX_TRY(m->site_beg.ensure(sizeof(int64_t) * beg.size()));
    TGX_TRY(m->pr_gidx.ensure(sizeof(int64_t) * gidx.size()));
    TGX_TRY(m->soap_der.ensure(sizeof(double) * 3 * (size_t)m->D * (size_t)n_pairs));
    TGX_TRY(ctx->fglob.ensure(sizeof(double) * 3 * (size_t)n_pairs));
    TGX_TRY(tgx_zero(ctx, ctx->fglob.p, sizeof(double) * 3 * (size_t)n_pairs));
    TGX_CHECK_HIP(hipMemcpyAsync(m->site_beg.p, beg.data(), sizeof(int64_t) * beg.size(), hipMemcpyHostToDevice, ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(m->pr_gidx.p, gidx.data(), sizeof(int64_t) * gidx.size(), hipMemcpyHostToDevice, ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(m->soap_der.p, soap_der, sizeof(double) * 3 * (size_t)m->D * (size_t)n_pairs,
                                 hipMemcpyHostToDevice, ctx->stream));
    TGX_TRY(tgx_plain_pair_forces(ctx, m));
    fp.resize(3 * (size_t)n_pairs);
    TGX_CHECK_HIP(hipMemcpyAsync(fp.data(), ctx->fglob.p, sizeof(double) * fp.size(), hipMemcpyDeviceToHost, ctx->stream));
  }
  TGX_CHECK_HIP(hipMemcpyAsync(energies, m->e_site.p, sizeof(double) * (size_t)n_sites, hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (do_forces) {
    // src/gap.f90:204-224: accumulate in the reference's pair order
    for (int64_t k = 0; k < 3 * (int64_t)n_sites0; ++k) forces[k] = 0.0;
    for (int k = 0; k < 9; ++k) virial[k] = 0.0;
    for (int64_t l = 0; l < n_pairs; ++l) {
      const int j2 = (neighbors_list[l] - 1) % n_sites0;
      const double* f = &fp[3 * (size_t)l];
      for (int k = 0; k < 3; ++k) forces[3 * (size_t)j2 + k] += f[k];
      for (int k1 = 0; k1 < 3; ++k1)
        for (int k2 = 0; k2 < 3; ++k2) virial[3 * k2 + k1] += 0.5 * (f[k1] * xyz[3 * l + k2] + f[k2] * xyz[3 * l + k1]);
    }
  }
  return TGX_OK;
}

// ---- whole step ------------------------------------------------------------------------
int tgx_surrogate_eval(tgx_ctx* ctx, tgx_model_t* m, bool selected);
int tgx_surrogate_finish(tgx_ctx* ctx, tgx_model_t* m, int n_gated);

extern "C" int tgx_compute(tgx_ctx* ctx, int path, double* energy, double* virial) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(path >= TGX_PATH_FP64_PLAIN && path <= TGX_PATH_SURROGATE, "tgx_compute: unknown path %d", path);
  TGX_REQUIRE(!ctx->models.empty(), "tgx_compute: no model defined");
  TGX_TRY(begin_step(ctx));
  ctx->gate_sites = ctx->gate_gated = 0;
  ctx->gated_fraction = ctx->max_spread = 0.0;
  const int nm = (int)ctx->models.size();
  // Blocks of distinct central species touch disjoint centres, pair entries
  // and atoms' energies, and each keeps its scratch in its own buffers: each
  // then runs on a stream of its own, so that one block's kernels fill the gaps
  // and tails of the other's. Blocks that share a central species (or take
  // every site) run in turn on the context stream, and so do all blocks while
  // kernels are being timed (each kernel's events then time it alone). The
  // mixed path runs its blocks in turn too: its BF16 GEMMs and the FP32
  // adjoint slow each other down more than the overlap gains (65,536 atoms:
  // 3.86 ms per step in turn against 4.14 ms concurrent; the surrogate and
  // FP64 adjoint paths gain 1-2% from the overlap).
  bool concurrent = nm > 1 && ctx->block_streams > 0 && !ctx->profiling &&
                    (path != TGX_PATH_MIXED || ctx->block_streams == 2);
  for (int a = 0; a < nm && concurrent; ++a) {
    const int ca = ctx->models[a]->central_species;
    if (ca == 0) concurrent = false;
    for (int b = 0; b < a && concurrent; ++b)
      if (ctx->models[b]->central_species == ca) concurrent = false;
  }
  hipStream_t main_stream = ctx->stream;
  // on every exit: the context stream back; on an error exit, no block work
  // left running
  struct StreamRestore {
    tgx_ctx* c;
    hipStream_t s;
    bool blocks;
    ~StreamRestore() {
      c->stream = s;
      if (blocks)
        for (auto* m : c->models)
          if (m->stream) (void)hipStreamSynchronize(m->stream);
    }
  } restore{ctx, main_stream, concurrent};
  if (concurrent) {
    if (!ctx->ev_blocks) TGX_CHECK_HIP(hipEventCreateWithFlags(&ctx->ev_blocks, hipEventDisableTiming));
    TGX_CHECK_HIP(hipEventRecord(ctx->ev_blocks, main_stream));
    for (auto* m : ctx->models) {
      if (!m->stream) TGX_CHECK_HIP(hipStreamCreate(&m->stream));
      if (!m->done) TGX_CHECK_HIP(hipEventCreateWithFlags(&m->done, hipEventDisableTiming));
      TGX_CHECK_HIP(hipStreamWaitEvent(m->stream, ctx->ev_blocks, 0));
    }
  }
  auto block_stream = [&](tgx_model_t* m) { return concurrent ? m->stream : main_stream; };
  // every block's selection counts in flight before the first wait
  for (auto* m : ctx->models) {
    ctx->stream = block_stream(m);
    TGX_TRY(tgx_block_select_count(ctx, m));
  }
  if (path == TGX_PATH_SURROGATE) {
    // every block's network first; each block's gate counters come back in a
    // pinned copy
    if (ctx->gate_host_n < 2 * nm) {
      if (ctx->gate_host) (void)hipHostFree(ctx->gate_host);
      ctx->gate_host = nullptr;
      ctx->gate_host_n = 0;
      TGX_CHECK_HIP(hipHostMalloc((void**)&ctx->gate_host, sizeof(int) * 2 * nm, hipHostMallocDefault));
      ctx->gate_host_n = 2 * nm;
    }
    for (int k = 0; k < nm; ++k) {
      tgx_model_t* m = ctx->models[k];
      ctx->stream = block_stream(m);
      TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
      TGX_TRY(tgx_block_select_fill(ctx, m));
      TGX_TRY(tgx_surrogate_eval(ctx, m, true));
      ctx->gate_host[2 * k] = ctx->gate_host[2 * k + 1] = 0;
      if (m->n_sites_b > 0 && ctx->gate_members > 1)
        TGX_CHECK_HIP(hipMemcpyAsync(ctx->gate_host + 2 * k, m->gate_stat.p, sizeof(int) * 2, hipMemcpyDeviceToHost,
                                     ctx->stream));
    }
    // each block's second phase as soon as its own gate counters are back, so
    // that it overlaps the other block's network
    for (int k = 0; k < nm; ++k) {
      tgx_model_t* m = ctx->models[k];
      ctx->stream = block_stream(m);
      TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
      const int ng = ctx->gate_host[2 * k];
      float sp;
      std::memcpy(&sp, &ctx->gate_host[2 * k + 1], sizeof(float));
      ctx->gate_sites += m->n_sites_b;
      ctx->gate_gated += ng;
      ctx->max_spread = std::max(ctx->max_spread, (double)sp);
      TGX_TRY(tgx_surrogate_finish(ctx, m, ng));
      TGX_TRY(tgx_energy_scatter(ctx, m));
      if (concurrent) TGX_CHECK_HIP(hipEventRecord(m->done, m->stream));
    }
    ctx->gated_fraction = ctx->gate_sites > 0 ? (double)ctx->gate_gated / (double)ctx->gate_sites : 0.0;
  } else {
    double budget = 0.0;
    TGX_TRY(memory_budget_bytes(ctx, &budget));
    for (auto* m : ctx->models) {
      ctx->stream = block_stream(m);
      TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
      if (path == TGX_PATH_FP64_PLAIN && plain_block_bytes(m, m->sel_host[0], m->sel_host[1]) > budget) {
        // dq/dr of the whole block does not fit: its centres in memory-bounded batches
        TGX_TRY(run_block_range(ctx, m, path, true, 0, ctx->n_sites, nullptr, nullptr, nullptr));
      } else {
        TGX_TRY(tgx_block_select_fill(ctx, m));
        TGX_TRY(run_selected_block(ctx, m, path, true));
        TGX_TRY(tgx_energy_scatter(ctx, m));
      }
      if (concurrent) TGX_CHECK_HIP(hipEventRecord(m->done, m->stream));
    }
  }
  ctx->stream = main_stream;
  if (concurrent)
    for (auto* m : ctx->models) TGX_CHECK_HIP(hipStreamWaitEvent(main_stream, m->done, 0));
  restore.blocks = false;  // the context stream now orders everything after the blocks
  TGX_TRY(tgx_gather_forces(ctx, energy, virial));
  return TGX_OK;
}

extern "C" int tgx_get_results(tgx_ctx* ctx, double* energies, double* forces) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(ctx->results_n > 0 && ctx->results_n == ctx->n_sites, "tgx_get_results: no results for this system");
  const size_t n = (size_t)ctx->n_sites;
  if (energies) TGX_CHECK_HIP(hipMemcpyAsync(energies, ctx->energies.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  if (forces) TGX_CHECK_HIP(hipMemcpyAsync(forces, ctx->forces.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

extern "C" int tgx_results_device(tgx_ctx* ctx, double** d_forces, double** d_energies) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(ctx->results_n > 0 && ctx->results_n == ctx->n_sites, "tgx_results_device: no results for this system");
  if (d_forces) *d_forces = ctx->forces.as<double>();
  if (d_energies) *d_energies = ctx->energies.as<double>();
  return TGX_OK;
}
