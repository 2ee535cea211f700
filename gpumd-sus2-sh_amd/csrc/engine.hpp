// Host engine behind the C ABI: owns the model, its device tables and the per-atom work
// buffers, and runs one evaluation as a fixed kernel sequence on the caller's stream:
//   cell sort -> neighbours -> basic moments -> CG graph -> forces/virial -> scatter to
//   caller order.
// Mirrors SUS2_SH (src/force/sus2_sh.cuh:7-203): constructor = model load + table upload
// (src/force/sus2_sh.cu:7782-9467), compute = src/force/sus2_sh.cu:9620-10675.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/sus2sh.h"
#include "device_common.hpp"
#include "kernels.hpp"
#include "model.hpp"
#include "block_program.hpp"
#include "lane_program.hpp"
#include "program.hpp"
#include "zbl.hpp"

namespace sus2sh {

struct DeviceError : public std::runtime_error {
  explicit DeviceError(const std::string& msg) : std::runtime_error(msg) {}
};

struct UnsupportedError : public std::runtime_error {
  explicit UnsupportedError(const std::string& msg) : std::runtime_error(msg) {}
};

struct ArgumentError : public std::runtime_error {
  explicit ArgumentError(const std::string& msg) : std::runtime_error(msg) {}
};

// Owning device allocation.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  ~DeviceBuffer();
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  void resize(size_t bytes);  // keeps nothing; grows only
  void upload(const void* host, size_t bytes);
  template <typename T>
  T* as() const
  {
    return static_cast<T*>(ptr_);
  }
  size_t bytes() const { return bytes_; }

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

struct EngineOptions {
  bool use_float = true;
  int graph_tile = 0;  // 0: automatic
  int graph_variant = -1;  // -1: automatic; 0 rows, 1 balanced streams, 2 round-robin,
                          // 3 lane program, 4 block program
  int block_waves = 0;   // block kernel: waves per 64-atom tile (0: autotuned, 4, 8 or 16)
  bool profile = false;
  bool autotune = true;  // time every feasible graph tile on the first call
  int lane_waves = 0;    // lane kernel: waves per 64-atom tile (0: autotuned shape)
  int lane_nslot = 0;    // lane kernel: accumulators per wave (0: 2 with 8 waves, else 4)
  int lane_min_slots = 160;  // lane kernel: smallest per-lane slot budget tried
  int lane_atoms = 1;        // lane kernel: atoms per lane
  // no host synchronisation in steady calls (capturable); the neighbour-capacity and type
  // checks are latched on the device and reported by check()
  bool deferred_check = false;
  bool md_graph = true;  // md_nve replays steady steps from a captured HIP graph
  // bit-reproducible energies, forces and virials: fixed-point force sums, no autotuner,
  // graph programs without atomic back-chunk merges
  bool deterministic = false;
  // edge-kernel launch shapes (threads per workgroup x lanes per atom) of the basic-moment
  // and force kernels; 0: the autotuner's choice (or the size rule when it does not run)
  int basic_threads = 0, basic_npl = 0;
  int force_threads = 0, force_npl = 0;
  // basic moments on the matrix cores: -1 the autotuner's choice between the vector kernel
  // and FP32 MFMAs, 0 never, 1 FP32 MFMAs, 2 BF16 operands with FP32 accumulation
  int basic_mfma = -1;
};

// Edge-kernel launch shape: threads per workgroup and lanes per atom.
struct EdgeShape {
  int threads = 256;
  int npl = 1;
  int mfma = 0;  // basic moments only: 0 vector kernel, 1 FP32 MFMA, 2 BF16 MFMA
};

struct CapacityError : public std::runtime_error {
  explicit CapacityError(const std::string& msg) : std::runtime_error(msg) {}
};

// Device work buffers of one single-structure evaluation: views of the engine's own
// allocations, or of the fixed scratch array in the allocation-free path.
struct Workspace {
  int *key = nullptr, *slot = nullptr, *arrival = nullptr, *skey = nullptr, *satom = nullptr,
      *rank = nullptr, *type_s = nullptr, *count = nullptr;
  double *wrapped = nullptr, *pos_s = nullptr, *e_s = nullptr, *f_s = nullptr, *v_s = nullptr;
  PackedAtom* packed = nullptr;
  void* nl = nullptr;      // Edge<Real>[cap][N]
  void* basics = nullptr;  // Real[B][N]
  void* grads = nullptr;   // Real[B][N]
  int *cell_count = nullptr, *cell_begin = nullptr, *scan_totals = nullptr,
      *max_count = nullptr, *win = nullptr;
  // two-layer gate models: gate moments and their adjoints Real[B][N]; per atom and radial
  // channel Real[N][Mr]: multiplier G, dG/dh and dE/dh
  void *gbasics = nullptr, *ggrads = nullptr, *gate_m = nullptr, *gate_d = nullptr,
       *gate_adj = nullptr;
  int cap = 0;  // neighbour-list capacity per atom
};

class Engine {
 public:
  Engine(const std::string& path, int num_atoms, int nopts, const char** opts);
  ~Engine();

  void compute(const sus2sh_box& box, int N, const int* type, const double* pos, double* pot,
               double* force, double* virial, hipStream_t stream, double* param_grad);
  // The same evaluation with every work buffer carved from the device's fixed scratch array
  // (kScratchBytes): no device allocation, no autotuning, a synchronised neighbour count.
  void compute_noalloc(const sus2sh_box& box, int N, const int* type, const double* pos,
                       double* pot, double* force, double* virial, hipStream_t stream);
  // Batch of independent periodic structures: structure s owns atoms
  // [atom_offsets[s], atom_offsets[s+1]) and has box boxes[s] (host arrays).
  void compute_batch(int nstruct, const sus2sh_box* boxes, const int* atom_offsets,
                     const int* type, const double* pos, double* pot, double* force,
                     double* virial, hipStream_t stream, double* param_grad);

  // Velocity-Verlet NVE: nsteps steps of dt_fs; energy_trace (host, may be null) receives
  // (sum potential, sum kinetic) in eV for the initial state and after every step.
  void md_nve(const sus2sh_box& box, int N, const int* type, const double* masses, double* pos,
              double* vel, double dt_fs, int nsteps, double* pot, double* force, double* virial,
              double* energy_trace, hipStream_t stream);

  // NPT (MTTK barostat + Nose-Hoover chains, npt.cpp): nsteps steps of dt fs; `box` is
  // updated in place; trace (host, may be null) receives double[nsteps + 1][8] = (potential,
  // kinetic, conserved extended energy (eV), volume (A^3), temperature (K), pressure (GPa),
  // cube root of the volume (A), target temperature (K)).
  struct NptParams {
    double dt = 1.0;
    double t_start = 300.0, t_stop = 300.0;  // K
    double p_start = 0.0, p_stop = 0.0;      // GPa
    double tperiod = 100.0, pperiod = 1000.0;  // in time steps
    int couple = 1;  // 0 iso, 1 aniso
  };
  void md_npt(sus2sh_box& box, int N, const int* type, const double* masses, double* pos,
              double* vel, const NptParams& prm, int nsteps, double* pot, double* force,
              double* virial, double* trace, hipStream_t stream);

  // Synchronise `stream` and report the checks latched by deferred-check evaluations since
  // the previous check (neighbour capacity: CapacityError, with the capacity raised for the
  // next call; atom type out of range: ArgumentError).
  void check(hipStream_t stream);

  // Replace the model coefficients (dE/dtheta order) and re-upload the device tables.
  void set_params(const double* theta, int n);
  void get_params(double* theta, int n) const;

  const Model& model() const { return model_; }
  const GraphProgram& program() const { return prog_; }
  const LaneProgram& lane_program() const { return lane_; }
  const BlockProgram& block_program() const { return block_; }
  const ZblTables& zbl() const { return zbl_; }
  const EngineOptions& options() const { return opt_; }
  int graph_tile() const { return tile_; }
  int graph_variant() const { return variant_; }
  const std::vector<float>& tune_ms() const { return tune_ms_; }
  // edge-kernel shapes in use and the autotuner's times per candidate
  // (edge_candidates()[c] at edge_tune_ms()[c] for the basic kernel, [C + c] for forces)
  EdgeShape basic_shape() const { return basic_shape_; }
  EdgeShape force_shape() const { return force_shape_; }
  static const std::vector<EdgeShape>& edge_candidates();
  const std::vector<float>& edge_tune_ms() const { return edge_tune_ms_; }
  int neighbor_capacity() const { return cap_; }
  int param_count() const;
  const std::vector<float>& stage_ms() const { return stage_ms_; }
  double mean_neighbors() const { return mean_neighbors_; }

 private:
  // latch_ word set by a fixed-point force contribution out of range (deterministic mode)
  static constexpr int kLatchFix = 5;
  // back-chunk length of the row programs: deterministic mode merges no chunk atomically
  int graph_chunk_terms() const { return opt_.deterministic ? (1 << 30) : 8; }
  static std::string fix_range_message();
  // start of k chunk c's block in the radial parameter-gradient accumulators
  size_t pg_chunk_offset(int c) const;
  template <typename Real>
  void upload_tables();
  template <typename Real>
  void upload_gate_tables();
  struct LaneShape {
    int waves;      // waves per tile
    int nslot;      // accumulators per wave
    int min_slots;  // smallest per-lane slot budget tried
    int atoms;      // atoms per lane
  };
  static const std::vector<LaneShape>& lane_shapes();
  LaneProgram build_lane(const LaneShape& shape) const;
  void compile_lane();
  void upload_lane();
  // block program: waves per tile the autotuner tries (the first is the default)
  static const std::vector<int>& block_waves_list();
  BlockProgram build_block(int waves) const;
  void upload_block();
  template <typename Real>
  bool launch_graph_variant(int variant, int tile, int N, const GraphTable<Real>& gt,
                            hipStream_t s);
  template <typename Real>
  void run(int nstruct, const sus2sh_box* boxes, const int* atom_offsets, int N, const int* type,
           const double* pos, double* pot, double* force, double* virial, hipStream_t stream,
           double* param_grad, bool noalloc);
  void ensure_atoms(int N, size_t edge_bytes);
  void bind_buffers(int N, int ncell, int n_totals);
  void carve_scratch(int N, int ncell, int n_totals, size_t edge_bytes, int win_words);
  void reset_latch(hipStream_t s);
  void read_latch(hipStream_t s, int* out);
  void finish_param_grad(double* d_param_grad, hipStream_t s);
  CellGrid make_grid(const sus2sh_box& box, int N, int* ncell) const;

  Model model_;
  GraphProgram prog_;
  // two-layer gate: the gate layer's graph program (gate scalars stored), its workgroup tile
  GraphProgram gate_prog_;
  int gate_tile_ = 1;
  EngineOptions opt_;
  int tile_ = 1;
  int variant_ = 1;  // 0: row mode, 1: balanced streams, 2: round-robin streams, 3: lane,
                     // 4: block
  bool tuned_ = false;
  // edge-kernel shapes: fixed by options, else tuned once on the first eligible call (the
  // size rule default_edge_shape() until then)
  EdgeShape basic_shape_, force_shape_;
  bool edge_tuned_ = false;
  std::vector<float> edge_tune_ms_;
  static EdgeShape default_edge_shape(int N);
  // graph-kernel ms per candidate, [variant * 5 + log2(tile)] for variants 0..2,
  // [15 + c] for lane-program shape c and [27 + c] for block-program shape c
  std::vector<float> tune_ms_;
  LaneProgram lane_;
  BlockProgram block_;
  int cap_ = 0;
  bool cap_known_ = false;  // cap_ was validated by a synchronised call with cap_atoms_ atoms
  int cap_atoms_ = -1;
  bool md_capture_ = false;  // md_nve is capturing a step: evaluations take the deferred path
  size_t max_lds_ = 65536;
  int device_ = 0;

  // model tables (precision-specific ones are uploaded for the active precision only)
  // The edge kernels are compiled for sh_k_max <= kMaxK; a larger k_max (up to the loader's
  // radial_funcs_count <= 48) runs them on chunks of consecutive k: chunk c covers
  // k in [kchunk_k0_[c], kchunk_k0_[c] + kchunk_K_[c]) with its own radial rows
  // (mu = k (l_max + 1) + l, contiguous) and (k,l,m) grid map.  Moments and forces are sums
  // over k, so the chunks' basics are disjoint and their forces add.
  static constexpr int kMaxKChunks = 8;
  int kchunks_ = 1;
  int kchunk_k0_[kMaxKChunks] = {};
  int kchunk_K_[kMaxKChunks] = {};
  DeviceBuffer radial_coeff_[kMaxKChunks], radial_scal_[kMaxKChunks], fmap_[kMaxKChunks];
  // matrix-core basic moments: shared radial coefficients [Mr][rb] and pair factors [S * S]
  DeviceBuffer mfma_rc_, pair_sigma_;
  // the uniform edge kernels' shared radial rows ([Mr_chunk][shared_nb_] per k chunk, and the
  // gate layer's): coeff_p = sigma_p shared
  DeviceBuffer shared_rc_[kMaxKChunks], gate_shared_;
  int shared_nb_ = 8;
  bool mfma_ok_ = false;  // the model fits the matrix-core basic-moment kernel
  DeviceBuffer fwd_rows_, fwd_w_, fwd_terms_, fwd_layer_;
  DeviceBuffer bwd_chunks_, bwd_terms_, bwd_layer_;
  DeviceBuffer e_slot_, e_w_, seed_, shift_, sigma_;
  DeviceBuffer fwd_ins_[2], fwd_sbegin_[2], fwd_slen_[2], bwd_ins_[2], bwd_sbegin_[2],
    bwd_slen_[2];
  DeviceBuffer row_sbegin_, row_slist_, stored_index_, stored_slot_;
  DeviceBuffer lane_segs_, lane_ins_;
  DeviceBuffer block_seg_, block_words_;
  DeviceBuffer win_;  // neighbour windows of the force tiles
  ZblTables zbl_;     // model.zbl_enabled: atomic numbers and pair cutoffs
  DeviceBuffer zbl_z_, zbl_inner_, zbl_outer_;
  // two-layer gate tables: radial coefficients, gate program, weights [Mr][Qg], slots,
  // additive coefficients
  DeviceBuffer gate_coeff_, g_fwd_rows_, g_fwd_terms_, g_fwd_layer_, g_bwd_chunks_, g_bwd_terms_,
    g_bwd_layer_, gate_weff_, gate_slot_, gate_add_;
  DeviceBuffer gbasics_, ggrads_, gate_m_, gate_d_, gate_adj_;
  int cu_count_ = 256;
  bool radial_uniform_ = false;

  // per-atom work buffers
  int atoms_alloc_ = 0;
  DeviceBuffer key_, slot_, arrival_, skey_, satom_, rank_, wrapped_, pos_s_, packed_, type_s_;
  DeviceBuffer count_;
  DeviceBuffer cell_count_, cell_begin_, scan_totals_, max_count_;
  DeviceBuffer latch_;  // deferred checks: max count, bad type, total edges (2), evaluations
  DeviceBuffer md_step_, md_x0_, md_v0_;
  hipStream_t md_stream_ = nullptr;  // md_nve graph mode (default priority)
  hipEvent_t md_event_ = nullptr;
  DeviceBuffer nl_, basics_, grads_, e_s_, f_s_, v_s_;
  DeviceBuffer pg_rad_, pg_misc_, md_mass_, md_energy_;
  DeviceBuffer batch_grids_, batch_cell_off_, batch_atom_off_, sid_, sid_s_;
  int* host_max_ = nullptr;  // pinned, 8 ints
  double* host_red_ = nullptr;  // pinned, 8 doubles (NPT reductions)
  DeviceBuffer md_red_;
  DeviceBuffer md_partial_;  // per-block partial sums of the MD reductions

  Workspace ws_;  // work buffers of the evaluation in progress
  std::vector<float> stage_ms_;
  std::vector<hipEvent_t> events_;
  double mean_neighbors_ = 0.0;
};

}  // namespace sus2sh
