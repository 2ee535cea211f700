// Kernel launch interface of the SUS2-SH engine.  Implementations live in
// neighbor.hip (cell list), edge_kernels.hip (basic moments, forces/virial),
// graph_kernel.hip (CG product graph forward / energy / reverse), lane_kernel.hip (the
// same graph as a wave-uniform program, one atom per lane), block_kernel.hip (the same graph
// as tensor blocks with compile-time couplings, one atom per lane).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"

namespace sus2sh {

// Raise a kernel's dynamic-LDS limit to the gfx950 maximum (160 KiB) once per kernel, on its
// first (uncaptured) launch, so that later launches -- also those recorded into a HIP
// graph -- make no attribute call.
constexpr int kMaxDynamicLds = 160 * 1024;
template <auto Kernel>
inline void allow_large_lds()
{
  static const bool done = [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(Kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, kMaxDynamicLds);
    return true;
  }();
  (void)done;
}

// Graph program tables on the device.
template <typename Real>
struct GraphTable {
  const int4* fwd_rows;   // (target slot or -1, term begin, term count, 0)
  const Real* fwd_w;      // energy weight of terminal rows
  const Term<Real>* fwd_terms;
  const int* fwd_layer;   // [layers + 1]
  const int4* bwd_chunks; // (src slot, term begin, term count, atomic)
  const Term<Real>* bwd_terms;
  const int* bwd_layer;   // [layers + 1]
  const int* e_slot;
  const Real* e_w;
  const Real* seed;       // [slots]
  // per-thread instruction streams (see GraphProgram)
  const Term<Real>* fwd_ins;
  const int* fwd_sbegin;  // [layers + 1]
  const int* fwd_slen;    // [layers]
  const Term<Real>* bwd_ins;
  const int* bwd_sbegin;
  const int* bwd_slen;
  int ones_slot;
  const double* shift;    // [S]
  const double* sigma;    // [S]
  int n_e;
  int slots;
  int basics;
  int layers;
};

// Parameter-gradient outputs of the graph kernel (all double, accumulated with atomics).
struct ParamGradOut {
  double* d_shift;          // [S]  += number of atoms of each type
  double* d_species;        // [S]  += sum over atoms of that type of (1 + E')
  double* d_moment;         // [Q]  += sum_i sigma_i M_i[map[s]]
  const int* row_sbegin;    // [rows + 1] scalars of each terminal row
  const int* row_slist;
  const int* stored_index;  // scalars held in a slot
  const int* stored_slot;
  int n_stored;
  int Q;
  int S;
  // 1: the forward pass and the sums above only (site energies and basic adjoints come from
  // another graph kernel); 0: also the site energies and the reverse pass
  int fwd_only;
};

// One cell-sorted atom for the neighbour search: wrapped position and species (as a
// double), 32 bytes, so a candidate is one cache-line-local pair of 16-byte loads.
struct alignas(32) PackedAtom {
  double x, y, z, t;
};

struct NeighborArgs {
  int N;
  CellGrid grid;              // single structure
  const CellGrid* grids;      // batch: per-structure grids (nullptr for a single structure)
  const int* cell_off;        // batch: first cell key of each structure
  const int* sid_s;           // batch: structure of each sorted atom
  const int* atom_off;        // batch: first atom of each structure (sorted order keeps them)
  double rc2;
  int cap;
};

// --- neighbour build (neighbor.hip) ---
// Cell sort: wrap into the box, cell key, counting sort by key (order inside a cell by atom
// index), gather of wrapped positions and types into the sorted order.  grids == nullptr:
// one structure with grid `grid`; otherwise atom_off [nstruct+1] (device) maps atoms to
// structures, sid receives each atom's structure and sid_s the same in sorted order.
struct CellSortArgs {
  int N;
  int S;             // species count; bad_type is set to 1 for a type outside [0, S)
  int ncell;
  const double* pos;
  const int* type;
  CellGrid grid;
  const CellGrid* grids;
  const int* cell_off;
  const int* atom_off;
  int nstruct;
  int* key;          // [N] cell of atom i
  int* slot;         // [N] arrival rank of atom i in its cell
  int* arrival;      // [N] atoms in cell order, arrival order inside a cell
  int* cell_count;   // [ncell + 1]
  int* cell_begin;   // [ncell + 1]: cell k holds sorted atoms [cell_begin[k], cell_begin[k+1])
  int* scan_totals;  // [cell_sort_totals(ncell)]
  double* wrapped;   // [3][N]
  int* sid;
  int* satom;        // [N] atom at sorted position p
  int* rank;         // [N] sorted position of atom i (inverse of satom)
  int* skey;         // [N] cell of sorted position p
  double* pos_s;     // [3][N]
  PackedAtom* packed;  // [N] position and species of sorted position p
  int* type_s;       // [N]
  int* bad_type;
  int* sid_s;
};
void launch_cell_sort(const CellSortArgs& a, hipStream_t s);
size_t cell_sort_totals(int ncell);
// latch (device int[8]) accumulates max_count[0..3] of a deferred-check call
void launch_latch(const int* max_count, int* latch, hipStream_t s);
void launch_fill_int(int* p, int value, size_t n, hipStream_t s);
void launch_fill_double(double* p, double value, size_t n, hipStream_t s);
template <typename Real>
void launch_neighbors(const NeighborArgs& a, const double* pos_s, const PackedAtom* packed,
                      const int* sorted_key, const int* cell_begin, const int* cell_end,
                      int* count, Edge<Real>* nl, int* max_count, hipStream_t s);
// Neighbour index windows of each force tile of kWinTile consecutive cell-sorted atoms: the
// neighbours of a tile lie in few contiguous ranges of the sorted order (the cells within
// +-R of the tile's cells, row by row).  win[tile * kWinStride + ...] = count (-1: no window
// set, the tile uses global atomics), total atoms, then count pairs (first atom, atoms).
// 128-atom tiles: their windows (~1,300-1,500 atoms in a dense box with rc-sized cells) fit
// 48 KiB of FP64 accumulators.
constexpr int kWinTile = 128;
constexpr int kWinMax = 64;
constexpr int kWinStride = 2 + 2 * kWinMax;
void launch_windows(int N, const CellGrid& g, const int* sorted_key, int* win, int cap,
                    hipStream_t s);
// The same for a batch (sid_s: structure of each sorted atom, atom_off [nstruct + 1] on the
// device): one window per tile, the sorted range of the structures the tile holds.
void launch_windows_batch(int N, const int* sid_s, const int* atom_off, int* win, int cap,
                          hipStream_t s);

// --- ZBL pair term (zbl.hip) ---
struct ZblTable {
  const int* z;          // [S] atomic numbers
  const double* inner;   // [S * S]
  const double* outer;   // [S * S]
  int S;
};
template <typename Real>
void launch_zbl(int N, const int* count, const Edge<Real>* nl, const int* type_s,
                const ZblTable& t, double* e_s, double* f_s, double* v_s, int* fix_flag,
                hipStream_t s);  // fix_flag != null: f_s holds fixed-point sums

// --- edge kernels (edge_kernels.hip) ---
template <typename Real>
struct EdgeLaunch {
  int N;
  int L;
  int K;
  int B;
  const int* count;
  const Edge<Real>* nl;
  const int* type_s;
  RadialTable<Real> rt;
  const int* fmap;  // [K (L+1)^2] full-grid index -> basic index, or -1
  // two-layer gate (null otherwise), per sorted atom and radial channel [N][Mr]:
  // gate_m = G multiplier of a neighbour, gate_d = its derivative dG/dh, gate_adj += dE/dh
  const Real* gate_m = nullptr;
  const Real* gate_d = nullptr;
  Real* gate_adj = nullptr;
  int v_accumulate = 0;  // forces: add the per-atom virial to v_s instead of storing it
  // deterministic summation (null otherwise): f_s holds 64-bit fixed-point sums
  // (fix_encode); *fix_flag is set when a contribution is out of range
  int* fix_flag = nullptr;
  // parameter gradient: interleaved copies of the per-pair LDS accumulators (power of two)
  int pg_copies = 1;
  // launch shape: `threads` per workgroup (64, 128 or 256) and `npl` consecutive lanes per
  // atom (1, 2, 4 or 8) that split its neighbour list (edge k goes to lane k mod npl) and
  // sum their partial results with lane shuffles; a workgroup covers threads / npl atoms
  int threads = 256;
  int npl = 1;
  // matrix-core basic moments (FP32 engine): the shared radial coefficient matrix c[mu][n]
  // ([Mr][rb]) and the pair factors sigma_p ([S * S]) with c_p = sigma_p c
  const Real* mfma_rc = nullptr;
  const Real* pair_sigma = nullptr;
};

// Basic moments as two dense matrix-core products per atom (T^T Y over the edges, then the
// radial coefficients): FP32 operands, or BF16 operands with FP32 accumulation.  For FP32
// engines with one radial scaling per pair (uniform), no gate and l_max <= 3,
// k_max (l_max + 1) <= 16, rb <= 16 (mfma_basics_supported).
bool mfma_basics_supported(int l_max, int k_max, int rb, int S, bool uniform, bool gate);
bool launch_basics_mfma(const EdgeLaunch<float>& a, float* basics, bool bf16, hipStream_t s);

// Feasible edge-kernel launch shapes (threads per workgroup, lanes per atom).
inline bool edge_shape_valid(int threads, int npl)
{
  return (threads == 64 || threads == 128 || threads == 256) &&
         (npl == 1 || npl == 2 || npl == 4 || npl == 8);
}

template <typename Real>
bool launch_basics(const EdgeLaunch<Real>& a, Real* basics, hipStream_t s);

// force/virial: grads[B][N] -> f_s[3][N] (atomic, double) and v_s[9][N].
// pg_rad (optional): accumulates dE/d(radial coeff, type coeff, scal) per pair into
// pg_rad[S*S][Mr*rb + Mr*2] (double, atomic).
// win (optional): tile windows from launch_windows(_batch) with capacity win_cap;
// neighbour forces are then summed per tile in LDS and flushed once per neighbour.
template <typename Real>
bool launch_forces(const EdgeLaunch<Real>& a, const Real* grads, double* f_s, double* v_s,
                   double* pg_rad, const int* win, int win_cap, hipStream_t s);
// The l_max-specific halves of launch_basics / launch_forces (one compiled unit each).
template <typename Real, int L>
bool launch_basics_l(const EdgeLaunch<Real>& a, Real* basics, hipStream_t s);
template <typename Real, int L>
bool launch_forces_l(const EdgeLaunch<Real>& a, const Real* grads, double* f_s, double* v_s,
                     double* pg_rad, const int* win, int win_cap, hipStream_t s);

// --- graph kernel (graph_kernel.hip) ---
// mode 0: rows / back chunks read per thread; mode 1: interleaved per-thread instruction
// streams (gt.fwd_ins / gt.bwd_ins).  pg != nullptr selects the persistent
// parameter-gradient variant (row mode) with at most persistent_blocks workgroups.
template <typename Real>
bool launch_graph(int tile, int mode, int N, const Real* basics, const int* type_s,
                  const GraphTable<Real>& gt, double* e_s, Real* grads, const ParamGradOut* pg,
                  hipStream_t s, size_t max_lds, int persistent_blocks);

template <typename Real>
size_t graph_lds_bytes(int tile, int slots);

// Lane program tables (see lane_program.hpp): one atom per lane, wave-uniform program.
template <typename Real>
struct LaneTable {
  const int4* segs;        // (kind, first record, groups per wave, 0)
  const uint32_t* recs;    // one record per group (encode_lane_records)
  const double* shift;     // [S]
  const double* sigma;     // [S]
  int nseg;
  int basics;
  int ones_slot;
  int v_base;              // first component-value slot (reused for the per-wave energies)
  int slots;               // per-lane slots: LDS of a workgroup = slots * 64 * sizeof(Real)
  int nslot;               // accumulators per wave (instructions per group)
  int waves;               // waves sharing one tile
  int atoms;               // atoms per lane (the tile holds 64 atoms x atoms)
};

// Lane graph kernel (lane_kernel.hip): one 64-atom tile per workgroup of lt.waves waves;
// returns false when the tile's LDS would exceed max_lds or the program shape is unknown.
// Whether a lane kernel is compiled for `waves` waves per tile x `nslot` accumulators x
// `atoms` atoms per lane.
bool lane_shape_supported(int waves, int nslot, int atoms);
template <typename Real>
bool launch_lane_graph(int N, const Real* basics, const int* type_s, const LaneTable<Real>& lt,
                       double* e_s, Real* grads, hipStream_t s, size_t max_lds);

// Block graph kernel (block_kernel.hip, program block_program.hpp): one 64-atom tile per
// workgroup of `waves` (4, 8 or 16) waves.
struct BlockTable {
  const int* seg;          // [batches * waves * 4] word ranges of the FWD / TENSOR epochs
  const uint32_t* words;   // run-encoded records
  const double* shift;     // [S]
  const double* sigma;     // [S]
  int batches;
  int basics;
  int v_base;              // first tensor-value slot (reused for the per-wave energies)
  int slots;               // per-lane slots of the program
  int waves;
};
// LDS bytes of one workgroup: the basic adjoints (basics x 64 doubles) and the program's Real
// slots (at least v_base + waves)
size_t block_graph_lds(const BlockTable& bt, int real_bytes);
// false when the tile's LDS would exceed max_lds or `waves` has no kernel
template <typename Real>
bool launch_block_graph(int N, const Real* basics, const int* type_s, const BlockTable& bt,
                        double* e_s, Real* grads, hipStream_t s, size_t max_lds);

// --- two-layer gate (gate_kernels.hip) ---
// Gate values h_{a,mu} = sum_q weff[mu][q] s_q(a) over the gate scalars s_q of the gate
// program (stored slots), G = 1 + A tanh(a_t h).
template <typename Real>
struct GateTable {
  const Real* weff;        // [Mr][Qg]
  const int* slot;         // [Qg] gate-program slot of scalar q
  const double* additive;  // [S] a_t
  double amplitude;        // A
  int Mr;
  int Qg;
};
// mode 0: gate moments gbasics [B][N] -> mult, dmult [N][Mr]; mode 1: per-atom scalar seeds
// sum_mu adj[a][mu] weff[mu][q] -> gate basic adjoints ggrad [B][N] (no species factor).
template <typename Real>
bool launch_gate_graph(int mode, int tile, int N, const Real* gbasics, const int* type_s,
                       const GraphTable<Real>& gt, const GateTable<Real>& ga, Real* mult,
                       Real* dmult, const Real* adj, Real* ggrad, hipStream_t s, size_t max_lds);
size_t gate_graph_lds_bytes(int tile, int slots, int real_bytes);

// --- velocity-Verlet NVE driver (md.hip) ---
void launch_half_kick(int N, double half_dt, const int* type, const double* mass, const double* f,
                      double* v, hipStream_t s);
void launch_drift(int N, double dt, const double* v, double* x, hipStream_t s);
// The reductions below sum in a fixed order (bit-identical between runs) through `partial`,
// a device buffer of md_partials_size(N) doubles.
size_t md_partials_size(int N);
// step == nullptr: out[0..1] = (potential, kinetic); otherwise out[2 * *step ..] =, and
// launch_step_advance(step) increments the device step counter
void launch_step_advance(int* step, hipStream_t s);
void launch_energies(int N, const int* type, const double* mass, const double* pot,
                     const double* v, double* out, const int* step, double* partial,
                     hipStream_t s);
// NPT: out[0..7) = (sum m v_x^2, m v_y^2, m v_z^2 in eV, sum potential, sum virial xx, yy, zz)
void launch_npt_reduce(int N, const int* type, const double* mass, const double* pot,
                       const double* v, const double* vir, double* out, double* partial,
                       hipStream_t s);
void launch_scale3(int N, double s0, double s1, double s2, double* x, hipStream_t s);

// --- scatter of per-sorted-atom results to the caller's arrays ---
// fixed != 0: f_s holds 64-bit fixed-point sums (deterministic summation)
void launch_scatter(int N, const int* rank, const double* e_s, const double* f_s,
                    const double* v_s, double* potential, double* force, double* virial,
                    int fixed, hipStream_t s);

}  // namespace sus2sh
