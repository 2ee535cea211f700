/*
 * sus2sh.h -- C ABI of the MI355X SUS2-SH potential engine.
 *
 * Drop-in boundary for the SUS2-SH backend of GPUMD-SUS2-SH.  Every entry point names
 * the reference interface it replaces.  All device pointers are plain HIP device
 * pointers; `stream` is a hipStream_t (NULL = the default stream).  No entry point
 * exits the process: failures return a non-zero status and sus2sh_last_error() holds
 * the message (the reference prints and calls exit(1), src/force/sus2_sh.cu:269-273).
 */
#ifndef SUS2SH_H
#define SUS2SH_H

#ifdef __cplusplus
extern "C" {
#endif

#define SUS2SH_OK 0
#define SUS2SH_ERR_INPUT 1       /* model file or option error */
#define SUS2SH_ERR_ARGUMENT 2    /* invalid argument to an entry point */
#define SUS2SH_ERR_DEVICE 3      /* HIP runtime or kernel launch failure */
#define SUS2SH_ERR_UNSUPPORTED 4 /* valid model that this engine does not evaluate */
#define SUS2SH_ERR_CAPACITY 5    /* deferred-check evaluations ran out of neighbour capacity */

typedef struct sus2sh_handle sus2sh_handle;

/* GPUMD Box: h[0..8] is the row-major 3x3 matrix whose COLUMNS are the cell vectors
 * a, b, c (x = h0 sa + h1 sb + h2 sc); h[9..17] is its inverse (row-major).
 * pbc[d] = 1 for a periodic direction.
 * Reference: GPUMD Box::cpu_h / pbc_x,y,z, used at src/force/sus2_sh.cu:7701-7703 and
 * src/force/force.cu:439-465. */
typedef struct sus2sh_box {
  double h[18];
  int pbc[3];
} sus2sh_box;

typedef struct sus2sh_model_info {
  int species_count;
  int l_max;
  int k_max;
  int body_order;
  int radial_basis_size;
  int radial_funcs_count;
  int moments_count;      /* alpha_moments_count */
  int basic_count;        /* alpha_index_basic_count */
  int scalar_count;       /* alpha_scalar_moments */
  int product_count;      /* sh_product_count */
  int layer_count;
  int row_count;          /* explicit rows (one per defined target) */
  int back_row_count;
  int back_term_count;
  int standard_graph;     /* 1 when the file graph equals the regenerated standard graph */
  int use_float;          /* 1: FP32 moments/adjoint, 0: FP64 throughout */
  int param_count;        /* length of the dE/dtheta vector */
  int stored_slots;       /* per-atom on-chip graph state (values and adjoints each) */
  int graph_tile;         /* atoms per workgroup in the graph kernel */
  int graph_variant;      /* graph kernel variant: 0 rows, 1 balanced streams, 2 round-robin,
                             3 lane program (one atom per lane) */
  int neighbor_capacity;  /* current per-atom neighbour capacity */
  double max_dist;
  int lane_program;       /* 1 when the lane program (graph depth <= 2) is available */
  int lane_slots;         /* per-lane on-chip slots of the lane program */
  int lane_instructions;  /* lane-program instructions per atom (including padding) */
  int lane_waves;         /* lane-program waves per 64-atom tile */
  int lane_nslot;         /* lane-program accumulators per wave */
  int lane_atoms;         /* lane-program atoms per lane */
} sus2sh_model_info;

/* Model identity: first token "MTP" and the substring "potential_tag = SUS2-SH".
 * Replaces is_sus2_sh_potential_file (src/force/sus2_sh.cu:7771-7780) together with the
 * "MTP" dispatch in Force::parse_potential (src/force/force.cu:82-87,112-114).
 * Returns 1 for a SUS2-SH model file, 0 otherwise. */
int sus2sh_probe(const char* path);

/* Constructor.  Replaces SUS2_SH::SUS2_SH(file_potential, num_atoms,
 * num_potential_options, potential_options) (src/force/sus2_sh.cuh:12-16, called from
 * src/force/force.cu:114).  `opts` are the run.in tokens after the model file: the
 * species_count element symbols, then key=value options (src/force/sus2_sh.cu:527-534).
 * Honoured options (like the reference, only key=value tokens are options; bare tokens are
 * ignored): sus2_float / sus2_sh_float (precision, env SUS2_GPUMD_FLOAT),
 * sus2_sh_graph_tile=<n>, sus2_sh_graph_variant=<0|1|2|3>, sus2_sh_lane_waves=<1|2|4|8|16>,
 * sus2_sh_lane_shape=<waves>x<accumulators>x<min slots>[x<atoms per lane>] (e.g. 8x2x208 or
 * 8x2x160x2: fixes the lane kernel's full shape), sus2_sh_autotune=<0|1>, sus2_sh_profile=<0|1>,
 * sus2_sh_deferred_check=<0|1> (steady calls make no host synchronisation and can be
 * captured in a HIP graph; see sus2sh_check), sus2_sh_md_graph=<0|1> (sus2sh_md_nve replays
 * captured steps, default 1), sus2_sh_deterministic=<0|1> (energies, forces and virials
 * bit-identical between calls, handles and MD runs: force contributions summed in 64-bit
 * fixed point, quantum 2^-34 eV/A, with pair terms below 131072 eV/A and at most 1023
 * neighbours per atom, else SUS2SH_ERR_ARGUMENT; no autotuner; not for dE/dtheta or
 * two-layer gate models; no counterpart in the reference, whose float force atomics vary
 * between runs, docs/gpumd-sus2-sh-interface-plan.md:1099-1107).  Validated like the reference: sus2_sh_factor_pruning=
 * <legacy|q-total> (src/force/sus2_sh.cu:558-587; the file's graph is always evaluated as an
 * explicit graph) and sus2_radial_direct=<1> (0 is rejected, src/force/sus2_sh.cu:7914-7916).
 * Other sus2_* tokens are accepted and ignored.  Uses the current HIP device. */
int sus2sh_create(const char* mtp_path, int num_atoms, int nopts, const char** opts,
                  sus2sh_handle** out);

/* Destructor.  Replaces SUS2_SH::~SUS2_SH (src/force/sus2_sh.cuh:17). */
void sus2sh_destroy(sus2sh_handle* handle);

/* One evaluation.  Replaces SUS2_SH::compute(Box&, type, position, potential, force,
 * virial) (src/force/sus2_sh.cuh:19-25, src/force/sus2_sh.cu:9620-10675), called from
 * Force::compute (src/force/force.cu:538-539).
 *   d_type      int[N]       species index of each atom (order of symbols in run.in)
 *   d_position  double[3N]   x[N], y[N], z[N]
 *   d_potential double[N]    += per-atom energy
 *   d_force     double[3N]   += fx[N], fy[N], fz[N]
 *   d_virial    double[9N]   += per-atom virial rows xx yy zz xy xz yz yx zx zy
 * Outputs are accumulated; the caller zeroes them (src/force/force.cu:513-519).
 * The neighbour list is rebuilt on every call (src/force/sus2_sh.cu:9648). */
int sus2sh_compute(sus2sh_handle* handle, const sus2sh_box* box, int num_atoms,
                   const int* d_type, const double* d_position, double* d_potential,
                   double* d_force, double* d_virial, void* stream);

/* Allocation-free evaluation.  No counterpart in the reference: for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc).  Same inputs,
 * outputs (+=) and results as sus2sh_compute, but every work buffer (cell sort, neighbour
 * list, basic moments and adjoints, per-atom energy / force / virial before the scatter) is
 * carved from one fixed device array of 128 MiB per GPU, held from the library's first
 * sus2sh_create on that GPU until the process ends; the call makes no device allocation or
 * free and no allocating library call.  Limit: the per-atom buffers plus a neighbour list
 * large enough for the densest atom must fit 128 MiB (about 65,000 atoms of an l3k3, rb 8,
 * rc 5 A model in FP32, 30,000 in FP64); larger inputs return SUS2SH_ERR_ARGUMENT with the
 * limit in the message.  The call synchronises `stream` once (neighbour count check), never
 * runs the autotuner (it uses the graph kernel the handle holds) and cannot be captured.
 * All calls on one GPU share its array and leave their contents in it: calls that overlap
 * in time on one GPU (other streams, threads or handles) race and must be serialised by the
 * caller.  The handle must have been created on the current device. */
int sus2sh_compute_noalloc(sus2sh_handle* handle, const sus2sh_box* box, int num_atoms,
                           const int* d_type, const double* d_position, double* d_potential,
                           double* d_force, double* d_virial, void* stream);

/* Deferred checks.  With sus2_sh_deferred_check=1, a sus2sh_compute call that follows a
 * synchronised call with the same num_atoms enqueues its kernels without any host
 * synchronisation or allocation (so it can be captured in a HIP graph on a non-default
 * stream), and latches on the device whether some atom had more neighbours than the list
 * capacity or a type out of range.  sus2sh_check synchronises `stream` and reports what
 * was latched since the previous check: SUS2SH_ERR_CAPACITY when neighbours were dropped
 * (the outputs of those calls are incomplete; the capacity is raised and the next
 * uncaptured call sizes the list again), SUS2SH_ERR_ARGUMENT for a bad type, else
 * SUS2SH_OK.  The reference synchronises on every call instead (thrust max_element over the
 * neighbour counts, src/force/sus2_sh.cu:9596-9598); no reference counterpart. */
int sus2sh_check(sus2sh_handle* handle, void* stream);

/* Model-parameter gradient of the total energy, dE/dtheta, accumulated (+=) into
 * d_param_grad (double[param_count]).  Parameter order:
 *   [shift_coeffs S][species_coeffs S][moment_coeffs Q][radial_coeffs M_r*rb]
 *   [radial type coefficients S][scal_coeffs 2 S^2 M_r]
 * No counterpart in the reference backend (training lives in the separate SUS2-SH
 * trainers).  Energy/forces/virial are accumulated exactly as in sus2sh_compute. */
int sus2sh_compute_param_grad(sus2sh_handle* handle, const sus2sh_box* box, int num_atoms,
                              const int* d_type, const double* d_position,
                              double* d_potential, double* d_force, double* d_virial,
                              double* d_param_grad, void* stream);

/* Batched evaluation of independent periodic structures in one pass of every kernel (one
 * launch per stage for the whole batch).  Structure s owns atoms [atom_offsets[s],
 * atom_offsets[s+1]) of the concatenated SoA arrays (N = atom_offsets[num_structures]) and
 * has box boxes[s]; boxes and atom_offsets are host arrays.  Outputs are per atom and
 * accumulated as in sus2sh_compute; d_param_grad (may be NULL) += dE/dtheta summed over all
 * structures.  No counterpart in the reference (GPUMD evaluates one box per call). */
int sus2sh_compute_batch(sus2sh_handle* handle, int num_structures, const sus2sh_box* boxes,
                         const int* atom_offsets, const int* d_type, const double* d_position,
                         double* d_potential, double* d_force, double* d_virial,
                         double* d_param_grad, void* stream);

/* Velocity-Verlet NVE driver around sus2sh_compute: nsteps steps of dt_fs femtoseconds.
 * masses: host double[species_count] in amu; d_position / d_velocity (A, A/fs, SoA [3N]) are
 * advanced in place; d_potential / d_force / d_virial hold the last evaluation (overwritten,
 * not accumulated); energy_trace (host, may be NULL) receives double[nsteps+1][2] = (sum of
 * potential, kinetic) in eV for the initial state and after each step.  No counterpart in
 * the reference overlay (GPUMD's integrators are upstream code); used for the NVE
 * energy-drift configurations. */
int sus2sh_md_nve(sus2sh_handle* handle, const sus2sh_box* box, int num_atoms,
                  const int* d_type, const double* masses, double* d_position,
                  double* d_velocity, double dt_fs, int nsteps, double* d_potential,
                  double* d_force, double* d_virial, double* energy_trace, void* stream);

/* NPT molecular dynamics, the integrator of the reference's benchmark protocol (GPUMD
 * `ensemble npt_mttk temp T1 T2 aniso p1 p2 tperiod 50 pperiod 500`,
 * docs/gpumd-sus2-sh-interface-plan.md:264-271; GPUMD's integrator is not in the reference
 * overlay): the Martyna-Tobias-Klein barostat with Nose-Hoover chains (length 3) on the
 * particles and on the barostat, in the reversible splitting of Shinoda et al. (2004).
 * couple = 0 (iso: the mean pressure drives three identical axes) or 1 (aniso: each axis its
 * own diagonal pressure); temperatures in K and pressures in GPa ramp linearly from start to
 * stop over the run; tperiod / pperiod are the thermostat / barostat periods in time steps.
 * The box must be periodic in x, y, z; its matrix (and inverse) is updated in place, and the
 * positions are dilated with it.  d_potential / d_force / d_virial hold the last evaluation;
 * trace (host, may be NULL) receives double[nsteps + 1][8] = (potential, kinetic, conserved
 * extended energy in eV, volume A^3, temperature K, pressure GPa, cube root of the volume A,
 * target temperature K).  No counterpart in the reference overlay. */
int sus2sh_md_npt_mttk(sus2sh_handle* handle, sus2sh_box* box, int num_atoms, const int* d_type,
                       const double* masses, double* d_position, double* d_velocity,
                       double dt_fs, int nsteps, double t_start, double t_stop, double p_start,
                       double p_stop, double tperiod, double pperiod, int couple,
                       double* d_potential, double* d_force, double* d_virial, double* trace,
                       void* stream);

/* Read / replace the model coefficients in the dE/dtheta order (host double[param_count]).
 * set_params re-uploads the device tables; the graph topology is unchanged.  For training
 * and gradient checks; no reference counterpart (the reference reads coefficients once). */
int sus2sh_set_params(sus2sh_handle* handle, const double* theta, int n);
int sus2sh_get_params(const sus2sh_handle* handle, double* theta, int n);

/* Last error message of a handle, or of the last failed sus2sh_create on this thread
 * when handle is NULL.  Never NULL. */
const char* sus2sh_last_error(const sus2sh_handle* handle);

/* Neighbour cutoff (reference: Potential::rc = max_dist, src/force/sus2_sh.cu:7821). */
double sus2sh_cutoff(const sus2sh_handle* handle);

int sus2sh_get_model_info(const sus2sh_handle* handle, sus2sh_model_info* out);

/* Per-stage GPU times of the last compute when profiling is on (sus2_sh_profile=1):
 * ms[0..n) = neighbour, basic, graph, force, scatter.  Returns the number written.
 * Replaces the reference's host-clock stage profile (src/force/sus2_sh.cu:9631-9645). */
int sus2sh_get_stage_times(const sus2sh_handle* handle, float* ms, int n);

/* Autotuner measurements of the graph kernel (ms per launch, -1 when not tried):
 * ms[v * 5 + log2(tile)] for variant v = 0..2 and tile 1..16, and ms[15 + c] for lane-program
 * (variant 3) shape c = 0..6: (waves per 64-atom tile, accumulators per wave, minimum
 * per-lane slots) = (4,4,160), (8,2,160), (4,2,160), (8,2,192), (8,2,256), (2,4,160),
 * (8,2,208), then with two atoms per lane (8,2,160), (4,2,160), (8,2,208) at c = 7..9, then
 * sixteen waves per tile (16,2,192), (16,2,320) at c = 10, 11.  When
 * the lane shape is fixed (sus2_sh_lane_waves / sus2_sh_lane_shape) only
 * that shape is timed and its time is in ms[15].  ms[27 + c] for the block program
 * (variant 4) with 8, 16, 4 waves per 64-atom tile at c = 0, 1, 2 (ms[27] only when
 * sus2_sh_block_waves fixes the shape).  Returns the count written. */
int sus2sh_get_tuning(const sus2sh_handle* handle, float* ms, int n);

/* Launch shapes of the edge kernels (no reference counterpart: the reference launches one
 * fixed shape).  shape[0..5) = basic-moment kernel threads per workgroup and lanes per atom,
 * force kernel threads and lanes per atom, and the basic-moment kernel kind (0 vector,
 * 1 FP32 matrix cores, 2 BF16 matrix cores; sus2_sh_basic_mfma).  ms[c] = the autotuner's
 * vector basic-kernel time of candidate c and ms[12 + c] its force-kernel time, candidates
 * c = 0..11 being threads 256, 128, 64 (c / 4) x lanes per atom 1, 2, 4, 8 (c mod 4);
 * ms[24] = the FP32 matrix-core basic kernel (-1: not tried).  Returns the number of ms
 * entries written (0 before the autotuner ran or without it). */
int sus2sh_get_edge_tuning(const sus2sh_handle* handle, int* shape, float* ms, int n);

/* Mean number of neighbours per atom found by the last compute. */
double sus2sh_mean_neighbors(const sus2sh_handle* handle);

/* ---------------- host-only model utilities (no GPU needed) ---------------- */

/* Clebsch-Gordan coefficient <j1 m1 j2 m2 | j m> (src/force/sus2_sh.cu:1107-1142). */
double sus2sh_clebsch_gordan(int j1, int m1, int j2, int m2, int j, int m);

/* Real-basis coupling coefficient (src/force/sus2_sh.cu:1186-1218). */
double sus2sh_real_cg(int l1, int m1, int l2, int m2, int L, int M);

typedef struct sus2sh_graph sus2sh_graph;

/* Standard real-CG graph (SHStandardGraphBuilderHost::BuildGraph,
 * src/force/sus2_sh.cu:1231-1249).  body_l_max holds the sh_body_l_max entries for body
 * orders 2.. (4 values for body order <= 5, 5 for body order 6), or NULL for l_max. */
int sus2sh_standard_graph(int l_max, int k_max, int body_order, const int* body_l_max,
                          int n_body_l_max, sus2sh_graph** out);
/* dims[0..5) = node_count, basic_count, product_count, scalar_count, max_layer */
int sus2sh_graph_dims(const sus2sh_graph* graph, int* dims);
/* basic_klm[3B] = (k,l,m); product_idx[3T] = (left,right,target); product_coeff[T];
 * scalars[Q]; scalar_body[Q] (body order of each scalar). */
int sus2sh_graph_export(const sus2sh_graph* graph, int* basic_klm, int* product_idx,
                        double* product_coeff, int* scalars, int* scalar_body);
void sus2sh_graph_free(sus2sh_graph* graph);

/* Host only: the ZBL pair tables of a model with zbl_enabled -- atomic number per species
 * (file metadata zbl_atomic_numbers, else the element symbols), and [S*S] inner / outer
 * cutoffs (type-wise when zbl_typewise_cutoff_enabled or a factor is given).  Replaces the
 * table set-up of src/force/sus2_sh.cu:7824-7860 with src/force/sus2_zbl_common.cuh:77-168;
 * errors carry the reference's messages (sus2sh_last_error(NULL)). */
int sus2sh_zbl_tables(const char* mtp_path, int nsym, const char** symbols, int* atomic_numbers,
                      double* inner, double* outer);

/* Parse and validate a model file on the host only (the reference loader,
 * src/force/sus2_sh.cu:2105-2482).  dims[0..16) = species_count, l_max, k_max,
 * body_order, radial_basis_size, radial_funcs_count, moments_count, basic_count,
 * scalar_count, product_count, layer_count, row_count, row_term_count,
 * back_row_count, back_term_count, standard_graph. */
int sus2sh_model_check(const char* mtp_path, int* dims);

/* Host-only check of the lane program (variant 3 of the graph kernel; no reference
 * counterpart): compile it for the model with at most max_slots per-lane slots, `waves`
 * (1, 2, 4, 8 or 16) waves per 64-atom tile and `nslot` (1, 2, 4 or 8; 0: 2 with 8 or 16
 * waves, else 4)
 * accumulators per wave, and run it on the host for n basic vectors
 * basics[n][B] -> e_prime[n] = sum_s w_s M[map[s]] and grad[n][B] = de_prime/dbasic (without
 * the species factor).  real_bytes = 0 runs the compiled program; 4 or 8 runs its device
 * encoding for FP32 / FP64 kernels (the records the kernel reads, coefficients rounded to
 * that precision) laid out for `atoms` (1 or 2) atoms per lane.  info[0..6) = ok, slots, instructions per atom, graph components,
 * batches, largest component.  When the model has no lane program (info[0] = 0) nothing is
 * evaluated.  Returns a status code. */
int sus2sh_lane_check(const char* mtp_path, int max_slots, int waves, int nslot, int real_bytes,
                      int atoms, int n, const double* basics, double* e_prime, double* grad,
                      int* info);

/* Host check of the block graph kernel's program (sus2_sh_graph_variant=4; no reference
 * counterpart): compile it for the model with `waves` (1, 2, 4, 8 or 16) waves per 64-atom
 * tile, coefficients of real_bytes (4 or 8) and at most max_slots per-lane slots, halving the
 * large degree classes when that brings it to target_slots (<= 0: never), and run its device
 * words on the host for n basic vectors basics[n][B] -> e_prime[n] and grad[n][B] as
 * sus2sh_lane_check does.  info[0..9) = ok, slots, batches, quadratic-form tensor records,
 * layer-1 blocks, forward terms, adjoint terms, reverse terms (per atom), units (classes or
 * class halves).  When the model has no block program (info[0] = 0) nothing is evaluated.
 * Returns a status code. */
int sus2sh_block_check(const char* mtp_path, int max_slots, int target_slots, int waves,
                       int real_bytes, int n, const double* basics, double* e_prime,
                       double* grad, int* info);

#ifdef __cplusplus
}
#endif

#endif /* SUS2SH_H */
