// C ABI (include/sus2sh.h): status codes instead of exit(1), messages kept per handle.
#include <cstring>
#include <memory>
#include <string>

#include "../../include/sus2sh.h"
#include "block_program.hpp"
#include "engine.hpp"
#include "lane_program.hpp"
#include "model.hpp"
#include "zbl.hpp"

struct sus2sh_handle {
  std::unique_ptr<sus2sh::Engine> engine;
  mutable std::string error;  // also written by the const (read-only) entry points
};

struct sus2sh_graph {
  sus2sh::StandardGraph g;
};

namespace {

thread_local std::string g_create_error;

template <typename F>
int guarded(std::string* err, F&& fn)
{
  try {
    fn();
    return SUS2SH_OK;
  } catch (const sus2sh::InputError& e) {
    *err = e.what();
    return SUS2SH_ERR_INPUT;
  } catch (const sus2sh::ArgumentError& e) {
    *err = e.what();
    return SUS2SH_ERR_ARGUMENT;
  } catch (const sus2sh::UnsupportedError& e) {
    *err = e.what();
    return SUS2SH_ERR_UNSUPPORTED;
  } catch (const sus2sh::DeviceError& e) {
    *err = e.what();
    return SUS2SH_ERR_DEVICE;
  } catch (const sus2sh::CapacityError& e) {
    *err = e.what();
    return SUS2SH_ERR_CAPACITY;
  } catch (const std::bad_alloc&) {
    *err = "host allocation failed";
    return SUS2SH_ERR_DEVICE;
  } catch (const std::exception& e) {
    *err = e.what();
    return SUS2SH_ERR_INPUT;
  }
}

}  // namespace

extern "C" {

int sus2sh_probe(const char* path)
{
  if (!path) return 0;
  try {
    return sus2sh::probe_model_file(path) ? 1 : 0;
  } catch (...) {
    return 0;
  }
}

int sus2sh_create(const char* mtp_path, int num_atoms, int nopts, const char** opts,
                  sus2sh_handle** out)
{
  if (!out) {
    g_create_error = "sus2sh_create: out is NULL";
    return SUS2SH_ERR_ARGUMENT;
  }
  *out = nullptr;
  if (!mtp_path || num_atoms < 0 || nopts < 0 || (nopts > 0 && !opts)) {
    g_create_error = "sus2sh_create: invalid arguments";
    return SUS2SH_ERR_ARGUMENT;
  }
  std::unique_ptr<sus2sh_handle> h(new sus2sh_handle());
  const int rc = guarded(&g_create_error, [&] {
    h->engine.reset(new sus2sh::Engine(mtp_path, num_atoms, nopts, opts));
  });
  if (rc != SUS2SH_OK) return rc;
  g_create_error.clear();
  *out = h.release();
  return SUS2SH_OK;
}

void sus2sh_destroy(sus2sh_handle* handle) { delete handle; }

int sus2sh_compute(sus2sh_handle* handle, const sus2sh_box* box, int num_atoms,
                   const int* d_type, const double* d_position, double* d_potential,
                   double* d_force, double* d_virial, void* stream)
{
  if (!handle) return SUS2SH_ERR_ARGUMENT;
  if (!box) {
    handle->error = "sus2sh_compute: box is NULL";
    return SUS2SH_ERR_ARGUMENT;
  }
  return guarded(&handle->error, [&] {
    handle->engine->compute(*box, num_atoms, d_type, d_position, d_potential, d_force, d_virial,
                            static_cast<hipStream_t>(stream), nullptr);
  });
}

int sus2sh_compute_noalloc(sus2sh_handle* handle, const sus2sh_box* box, int num_atoms,
                           const int* d_type, const double* d_position, double* d_potential,
                           double* d_force, double* d_virial, void* stream)
{
  if (!handle) return SUS2SH_ERR_ARGUMENT;
  if (!box) {
    handle->error = "sus2sh_compute_noalloc: box is NULL";
    return SUS2SH_ERR_ARGUMENT;
  }
  return guarded(&handle->error, [&] {
    handle->engine->compute_noalloc(*box, num_atoms, d_type, d_position, d_potential, d_force,
                                    d_virial, static_cast<hipStream_t>(stream));
  });
}

int sus2sh_compute_param_grad(sus2sh_handle* handle, const sus2sh_box* box, int num_atoms,
                              const int* d_type, const double* d_position,
                              double* d_potential, double* d_force, double* d_virial,
                              double* d_param_grad, void* stream)
{
  if (!handle) return SUS2SH_ERR_ARGUMENT;
  if (!box || !d_param_grad) {
    handle->error = "sus2sh_compute_param_grad: box or d_param_grad is NULL";
    return SUS2SH_ERR_ARGUMENT;
  }
  return guarded(&handle->error, [&] {
    handle->engine->compute(*box, num_atoms, d_type, d_position, d_potential, d_force, d_virial,
                            static_cast<hipStream_t>(stream), d_param_grad);
  });
}

int sus2sh_compute_batch(sus2sh_handle* handle, int num_structures, const sus2sh_box* boxes,
                         const int* atom_offsets, const int* d_type, const double* d_position,
                         double* d_potential, double* d_force, double* d_virial,
                         double* d_param_grad, void* stream)
{
  if (!handle) return SUS2SH_ERR_ARGUMENT;
  return guarded(&handle->error, [&] {
    handle->engine->compute_batch(num_structures, boxes, atom_offsets, d_type, d_position,
                                  d_potential, d_force, d_virial,
                                  static_cast<hipStream_t>(stream), d_param_grad);
  });
}

int sus2sh_check(sus2sh_handle* handle, void* stream)
{
  if (!handle) return SUS2SH_ERR_ARGUMENT;
  return guarded(&handle->error,
                 [&] { handle->engine->check(static_cast<hipStream_t>(stream)); });
}

int sus2sh_md_nve(sus2sh_handle* handle, const sus2sh_box* box, int num_atoms,
                  const int* d_type, const double* masses, double* d_position,
                  double* d_velocity, double dt_fs, int nsteps, double* d_potential,
                  double* d_force, double* d_virial, double* energy_trace, void* stream)
{
  if (!handle) return SUS2SH_ERR_ARGUMENT;
  if (!box) {
    handle->error = "sus2sh_md_nve: box is NULL";
    return SUS2SH_ERR_ARGUMENT;
  }
  return guarded(&handle->error, [&] {
    handle->engine->md_nve(*box, num_atoms, d_type, masses, d_position, d_velocity, dt_fs, nsteps,
                           d_potential, d_force, d_virial, energy_trace,
                           static_cast<hipStream_t>(stream));
  });
}

int sus2sh_md_npt_mttk(sus2sh_handle* handle, sus2sh_box* box, int num_atoms, const int* d_type,
                       const double* masses, double* d_position, double* d_velocity,
                       double dt_fs, int nsteps, double t_start, double t_stop, double p_start,
                       double p_stop, double tperiod, double pperiod, int couple,
                       double* d_potential, double* d_force, double* d_virial, double* trace,
                       void* stream)
{
  if (!handle) return SUS2SH_ERR_ARGUMENT;
  if (!box) {
    handle->error = "sus2sh_md_npt_mttk: box is NULL";
    return SUS2SH_ERR_ARGUMENT;
  }
  return guarded(&handle->error, [&] {
    sus2sh::Engine::NptParams prm;
    prm.dt = dt_fs;
    prm.t_start = t_start;
    prm.t_stop = t_stop;
    prm.p_start = p_start;
    prm.p_stop = p_stop;
    prm.tperiod = tperiod;
    prm.pperiod = pperiod;
    prm.couple = couple;
    handle->engine->md_npt(*box, num_atoms, d_type, masses, d_position, d_velocity, prm, nsteps,
                           d_potential, d_force, d_virial, trace, static_cast<hipStream_t>(stream));
  });
}

int sus2sh_set_params(sus2sh_handle* handle, const double* theta, int n)
{
  if (!handle) return SUS2SH_ERR_ARGUMENT;
  return guarded(&handle->error, [&] { handle->engine->set_params(theta, n); });
}

int sus2sh_get_params(const sus2sh_handle* handle, double* theta, int n)
{
  if (!handle) return SUS2SH_ERR_ARGUMENT;
  return guarded(&handle->error, [&] { handle->engine->get_params(theta, n); });
}

const char* sus2sh_last_error(const sus2sh_handle* handle)
{
  if (!handle) return g_create_error.c_str();
  return handle->error.c_str();
}

double sus2sh_cutoff(const sus2sh_handle* handle)
{
  return handle ? handle->engine->model().max_dist : 0.0;
}

int sus2sh_get_model_info(const sus2sh_handle* handle, sus2sh_model_info* out)
{
  if (!handle || !out) return SUS2SH_ERR_ARGUMENT;
  const sus2sh::Model& m = handle->engine->model();
  const sus2sh::GraphProgram& g = handle->engine->program();
  std::memset(out, 0, sizeof(*out));
  out->species_count = m.species_count;
  out->l_max = m.l_max;
  out->k_max = m.k_max;
  out->body_order = m.body_order;
  out->radial_basis_size = m.rb_size;
  out->radial_funcs_count = m.radial_funcs_count;
  out->moments_count = m.moments_count;
  out->basic_count = m.basic_count;
  out->scalar_count = m.scalar_count;
  out->product_count = static_cast<int>(m.products.size());
  out->layer_count = m.layer_count;
  out->row_count = static_cast<int>(m.rows.size());
  out->back_row_count = m.back_row_count;
  out->back_term_count = m.back_term_count;
  out->standard_graph = m.standard_graph_matched ? 1 : 0;
  out->use_float = handle->engine->options().use_float ? 1 : 0;
  out->param_count = handle->engine->param_count();
  out->stored_slots = g.slot_count;
  out->graph_tile = handle->engine->graph_tile();
  out->graph_variant = handle->engine->graph_variant();
  out->neighbor_capacity = handle->engine->neighbor_capacity();
  out->max_dist = m.max_dist;
  const sus2sh::LaneProgram& lp = handle->engine->lane_program();
  out->lane_program = lp.ok ? 1 : 0;
  out->lane_slots = lp.ok ? lp.slots : 0;
  out->lane_instructions = lp.ok ? static_cast<int>(lp.ins.size()) : 0;
  out->lane_waves = lp.ok ? lp.waves : 0;
  out->lane_nslot = lp.ok ? lp.nslot : 0;
  out->lane_atoms = lp.ok ? lp.atoms : 0;
  return SUS2SH_OK;
}

int sus2sh_get_stage_times(const sus2sh_handle* handle, float* ms, int n)
{
  if (!handle || !ms || n < 0) return 0;
  const std::vector<float>& st = handle->engine->stage_ms();
  const int k = std::min<int>(n, static_cast<int>(st.size()));
  for (int i = 0; i < k; ++i) ms[i] = st[i];
  return k;
}

int sus2sh_get_tuning(const sus2sh_handle* handle, float* ms, int n)
{
  if (!handle || !ms || n < 0) return 0;
  const std::vector<float>& t = handle->engine->tune_ms();
  const int k = std::min<int>(n, static_cast<int>(t.size()));
  for (int i = 0; i < k; ++i) ms[i] = t[i];
  return k;
}

int sus2sh_get_edge_tuning(const sus2sh_handle* handle, int* shape, float* ms, int n)
{
  if (!handle) return 0;
  const sus2sh::Engine& e = *handle->engine;
  if (shape) {
    shape[0] = e.basic_shape().threads;
    shape[1] = e.basic_shape().npl;
    shape[2] = e.force_shape().threads;
    shape[3] = e.force_shape().npl;
    shape[4] = e.basic_shape().mfma;
  }
  if (!ms || n < 0) return 0;
  const std::vector<float>& t = e.edge_tune_ms();
  const int k = std::min<int>(n, static_cast<int>(t.size()));
  for (int i = 0; i < k; ++i) ms[i] = t[i];
  return k;
}

double sus2sh_mean_neighbors(const sus2sh_handle* handle)
{
  return handle ? handle->engine->mean_neighbors() : 0.0;
}

double sus2sh_clebsch_gordan(int j1, int m1, int j2, int m2, int j, int m)
{
  return sus2sh::clebsch_gordan(j1, m1, j2, m2, j, m);
}

double sus2sh_real_cg(int l1, int m1, int l2, int m2, int L, int M)
{
  try {
    return sus2sh::real_cg(l1, m1, l2, m2, L, M);
  } catch (...) {
    return 0.0;
  }
}

int sus2sh_standard_graph(int l_max, int k_max, int body_order, const int* body_l_max,
                          int n_body_l_max, sus2sh_graph** out)
{
  if (!out) return SUS2SH_ERR_ARGUMENT;
  *out = nullptr;
  std::unique_ptr<sus2sh_graph> g(new sus2sh_graph());
  const int rc = guarded(&g_create_error, [&] {
    std::vector<int> bl(7, l_max);
    if (body_l_max) {
      if (n_body_l_max < 0 || n_body_l_max > 5)
        throw sus2sh::ArgumentError("n_body_l_max must be in [0,5]");
      for (int i = 0; i < n_body_l_max; ++i) bl[2 + i] = body_l_max[i];
    }
    for (int b = 2; b <= body_order && b < 7; ++b)
      if (bl[b] < 0 || bl[b] > l_max) throw sus2sh::ArgumentError("body_l_max entry out of range");
    g->g = sus2sh::build_standard_graph(l_max, k_max, body_order, bl);
  });
  if (rc != SUS2SH_OK) return rc;
  *out = g.release();
  return SUS2SH_OK;
}

int sus2sh_graph_dims(const sus2sh_graph* graph, int* dims)
{
  if (!graph || !dims) return SUS2SH_ERR_ARGUMENT;
  dims[0] = graph->g.node_count;
  dims[1] = static_cast<int>(graph->g.basic_k.size());
  dims[2] = static_cast<int>(graph->g.products.size());
  dims[3] = static_cast<int>(graph->g.scalars.size());
  dims[4] = graph->g.max_layer;
  return SUS2SH_OK;
}

int sus2sh_graph_export(const sus2sh_graph* graph, int* basic_klm, int* product_idx,
                        double* product_coeff, int* scalars, int* scalar_body)
{
  if (!graph) return SUS2SH_ERR_ARGUMENT;
  const sus2sh::StandardGraph& g = graph->g;
  if (basic_klm)
    for (size_t b = 0; b < g.basic_k.size(); ++b) {
      basic_klm[3 * b] = g.basic_k[b];
      basic_klm[3 * b + 1] = g.basic_l[b];
      basic_klm[3 * b + 2] = g.basic_m[b];
    }
  for (size_t p = 0; p < g.products.size(); ++p) {
    if (product_idx) {
      product_idx[3 * p] = g.products[p].left;
      product_idx[3 * p + 1] = g.products[p].right;
      product_idx[3 * p + 2] = g.products[p].target;
    }
    if (product_coeff) product_coeff[p] = g.products[p].coeff;
  }
  for (size_t s = 0; s < g.scalars.size(); ++s) {
    if (scalars) scalars[s] = g.scalars[s];
    if (scalar_body) scalar_body[s] = g.scalar_body[s];
  }
  return SUS2SH_OK;
}

void sus2sh_graph_free(sus2sh_graph* graph) { delete graph; }

int sus2sh_lane_check(const char* mtp_path, int max_slots, int waves, int nslot, int real_bytes,
                      int atoms, int n, const double* basics, double* e_prime, double* grad,
                      int* info)
{
  if (nslot == 0) nslot = waves >= 8 ? 2 : 4;
  if (!mtp_path || n < 0 || (n > 0 && (!basics || !e_prime || !grad)) ||
      (waves != 1 && waves != 2 && waves != 4 && waves != 8 && waves != 16) ||
      (nslot != 1 && nslot != 2 && nslot != 4 && nslot != 8) ||
      (real_bytes != 0 && real_bytes != 4 && real_bytes != 8) || (atoms != 1 && atoms != 2))
    return SUS2SH_ERR_ARGUMENT;
  return guarded(&g_create_error, [&] {
    const sus2sh::Model m = sus2sh::load_model(mtp_path);
    const sus2sh::LaneProgram p = sus2sh::compile_lane_program(m, max_slots, nslot, waves);
    if (info) {
      info[0] = p.ok ? 1 : 0;
      info[1] = p.slots;
      info[2] = static_cast<int>(p.ins.size());
      info[3] = p.components;
      info[4] = p.batches;
      info[5] = p.largest_component;
    }
    if (!p.ok) return;
    const int B = m.basic_count;
    if (real_bytes == 0) {
      for (int i = 0; i < n; ++i)
        e_prime[i] = sus2sh::run_lane_program_host(p, basics + static_cast<size_t>(i) * B,
                                                   grad + static_cast<size_t>(i) * B);
      return;
    }
    const sus2sh::LaneRecords rec = sus2sh::encode_lane_records(p, real_bytes, 64 * atoms);
    for (int i = 0; i < n; ++i)
      e_prime[i] = sus2sh::run_lane_records_host(p, rec, basics + static_cast<size_t>(i) * B,
                                                 grad + static_cast<size_t>(i) * B);
  });
}

int sus2sh_block_check(const char* mtp_path, int max_slots, int target_slots, int waves,
                       int real_bytes, int n, const double* basics, double* e_prime,
                       double* grad, int* info)
{
  if (!mtp_path || n < 0 || (n > 0 && (!basics || !e_prime || !grad)) ||
      (waves != 1 && waves != 2 && waves != 4 && waves != 8 && waves != 16) ||
      (real_bytes != 4 && real_bytes != 8))
    return SUS2SH_ERR_ARGUMENT;
  return guarded(&g_create_error, [&] {
    const sus2sh::Model m = sus2sh::load_model(mtp_path);
    const sus2sh::BlockProgram p =
      sus2sh::compile_block_program(m, waves, real_bytes, max_slots, target_slots);
    if (info) {
      info[0] = p.ok ? 1 : 0;
      info[1] = p.slots;
      info[2] = p.batches;
      info[3] = p.tensors;
      info[4] = p.fwd_blocks;
      info[5] = static_cast<int>(p.fwd_terms);
      info[6] = static_cast<int>(p.adj_terms);
      info[7] = static_cast<int>(p.bwd_terms);
      info[8] = p.units;
    }
    if (!p.ok) return;
    const int B = m.basic_count;
    for (int i = 0; i < n; ++i)
      e_prime[i] = sus2sh::run_block_program_host(p, basics + static_cast<size_t>(i) * B,
                                                  grad + static_cast<size_t>(i) * B);
  });
}

int sus2sh_zbl_tables(const char* mtp_path, int nsym, const char** symbols, int* atomic_numbers,
                      double* inner, double* outer)
{
  if (!mtp_path || nsym < 0 || (nsym > 0 && !symbols)) return SUS2SH_ERR_ARGUMENT;
  return guarded(&g_create_error, [&] {
    const sus2sh::Model m = sus2sh::load_model(mtp_path);
    if (!m.zbl_enabled) throw sus2sh::ArgumentError("model has no ZBL term (zbl_enabled)");
    const sus2sh::ZblTables t = sus2sh::build_zbl_tables(m, nsym, symbols);
    const int S = m.species_count;
    if (atomic_numbers) std::memcpy(atomic_numbers, t.atomic_number.data(), S * sizeof(int));
    if (inner) std::memcpy(inner, t.inner.data(), static_cast<size_t>(S) * S * sizeof(double));
    if (outer) std::memcpy(outer, t.outer.data(), static_cast<size_t>(S) * S * sizeof(double));
  });
}

int sus2sh_model_check(const char* mtp_path, int* dims)
{
  if (!mtp_path) return SUS2SH_ERR_ARGUMENT;
  return guarded(&g_create_error, [&] {
    const sus2sh::Model m = sus2sh::load_model(mtp_path);
    if (dims) {
      dims[0] = m.species_count;
      dims[1] = m.l_max;
      dims[2] = m.k_max;
      dims[3] = m.body_order;
      dims[4] = m.rb_size;
      dims[5] = m.radial_funcs_count;
      dims[6] = m.moments_count;
      dims[7] = m.basic_count;
      dims[8] = m.scalar_count;
      dims[9] = static_cast<int>(m.products.size());
      dims[10] = m.layer_count;
      dims[11] = static_cast<int>(m.rows.size());
      dims[12] = static_cast<int>(m.row_terms.size());
      dims[13] = m.back_row_count;
      dims[14] = m.back_term_count;
      dims[15] = m.standard_graph_matched ? 1 : 0;
    }
  });
}

}  // extern "C"
