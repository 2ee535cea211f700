#include "engine.hpp"

#include "scratch.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>

namespace sus2sh {

namespace {

void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

bool starts_with(const std::string& s, const std::string& p)
{
  return s.size() >= p.size() && s.compare(0, p.size(), p) == 0;
}

bool parse_bool(const std::string& v, const std::string& name)
{
  if (v == "1" || v == "true" || v == "TRUE" || v == "yes" || v == "on") return true;
  if (v == "0" || v == "false" || v == "FALSE" || v == "no" || v == "off") return false;
  throw InputError("Invalid boolean value for " + name + ": " + v);
}

// Options after the species symbols (src/force/sus2_sh.cu:520-535): sus2_float /
// sus2_sh_float select the precision (env SUS2_GPUMD_FLOAT); the reference's other
// sus2_* switches select CUDA micro-variants and are accepted and ignored here, except the
// two whose values the reference validates: the factor-pruning mode (legacy | q-total,
// src/force/sus2_sh.cu:558-587; this engine always evaluates the file's explicit graph, which
// is what q-total selects) and direct radial evaluation, which must stay on
// (src/force/sus2_sh.cu:538-556, 7914-7916).  Like the reference, only `key=value` tokens are
// options: a bare token (no '=') is ignored.
EngineOptions parse_options(const Model& m, int nopts, const char** opts)
{
  EngineOptions o;
  if (const char* env = std::getenv("SUS2_GPUMD_FLOAT")) o.use_float = parse_bool(env, "SUS2_GPUMD_FLOAT");
  if (const char* env = std::getenv("SUS2_GPUMD_PROFILE")) o.profile = parse_bool(env, "SUS2_GPUMD_PROFILE");
  bool radial_direct = true;
  if (const char* env = std::getenv("SUS2_GPUMD_RADIAL_DIRECT"))
    radial_direct = parse_bool(env, "SUS2_GPUMD_RADIAL_DIRECT");
  std::string pruning = "legacy";
  if (const char* env = std::getenv("SUS2_SH_GPUMD_FACTOR_PRUNING"))
    pruning = env;
  else if (const char* env2 = std::getenv("SUS2_GPUMD_FACTOR_PRUNING"))
    pruning = env2;
  const int begin = std::min(nopts, m.species_count);
  for (int i = begin; i < nopts; ++i) {
    const std::string opt = opts[i] ? opts[i] : "";
    const size_t eq = opt.find('=');
    if (eq == std::string::npos) continue;  // not an option token
    const std::string key = opt.substr(0, eq);
    const std::string val = opt.substr(eq + 1);
    if (key == "sus2_radial_direct" || key == "sus2_sh_radial_direct" || key == "radial_direct") {
      radial_direct = parse_bool(val, key);
    } else if (key == "sus2_sh_factor_pruning" || key == "sus2_factor_pruning" ||
               key == "sh_factor_pruning") {
      pruning = val;
    } else if (key == "sus2_sh_lane_shape") {
      // waves x accumulators x minimum slot budget [x atoms per lane], e.g. 8x2x208 or
      // 8x2x160x2
      int w = 0, a = 0, s = 0, na = 1;
      char x1 = 0, x2 = 0, x3 = 0;
      std::istringstream in(val);
      in >> w >> x1 >> a >> x2 >> s;
      if (in && !in.eof()) in >> x3 >> na;
      if (!in || x1 != 'x' || x2 != 'x' || (x3 != 0 && x3 != 'x') || !in.eof())
        throw InputError("sus2_sh_lane_shape must be <waves>x<accumulators>x<slots>[x<atoms>]: " +
                         val);
      o.lane_waves = w;
      o.lane_nslot = a;
      o.lane_min_slots = s;
      o.lane_atoms = na;
    } else if (key == "sus2_float" || key == "sus2_sh_float") {
      o.use_float = parse_bool(val, key);
    } else if (key == "sus2_sh_profile") {
      o.profile = parse_bool(val, key);
    } else if (key == "sus2_sh_autotune") {
      o.autotune = parse_bool(val, key);
    } else if (key == "sus2_sh_graph_variant") {
      const int v = std::atoi(val.c_str());
      if (v < 0 || v > 4 || val.empty() || val.find_first_not_of("0123456789") != std::string::npos)
        throw InputError("sus2_sh_graph_variant must be 0, 1, 2, 3 or 4: " + val);
      o.graph_variant = v;
    } else if (key == "sus2_sh_block_waves") {
      const int w = std::atoi(val.c_str());
      if ((w != 4 && w != 8 && w != 16) || val.find_first_not_of("0123456789") != std::string::npos)
        throw InputError("sus2_sh_block_waves must be 4, 8 or 16: " + val);
      o.block_waves = w;
    } else if (key == "sus2_sh_lane_waves") {
      const int w = std::atoi(val.c_str());
      if (w != 1 && w != 2 && w != 4 && w != 8 && w != 16)
        throw InputError("sus2_sh_lane_waves must be 1, 2, 4, 8 or 16: " + val);
      // fixes the lane program shape: 2 accumulators per wave with 8 or 16 waves, else 4;
      // the smallest slot budget that holds the program, from 160 up
      o.lane_waves = w;
      o.lane_nslot = w >= 8 ? 2 : 4;
      o.lane_min_slots = 160;
      o.lane_atoms = 1;
    } else if (key == "sus2_sh_deferred_check") {
      o.deferred_check = parse_bool(val, key);
    } else if (key == "sus2_sh_md_graph") {
      o.md_graph = parse_bool(val, key);
    } else if (key == "sus2_sh_deterministic") {
      o.deterministic = parse_bool(val, key);
    } else if (key == "sus2_sh_basic_shape" || key == "sus2_sh_force_shape") {
      // threads per workgroup x lanes per atom of an edge kernel, e.g. 256x1 or 128x4
      int t = 0, n = 0;
      char x = 0;
      std::istringstream in(val);
      in >> t >> x >> n;
      if (!in || x != 'x' || !in.eof() || !edge_shape_valid(t, n))
        throw InputError(key + " must be <threads>x<lanes per atom> with threads 64, 128 or 256 "
                         "and 1, 2, 4 or 8 lanes per atom: " + val);
      (key == "sus2_sh_basic_shape" ? o.basic_threads : o.force_threads) = t;
      (key == "sus2_sh_basic_shape" ? o.basic_npl : o.force_npl) = n;
    } else if (key == "sus2_sh_basic_mfma") {
      if (val == "0" || val == "off") o.basic_mfma = 0;
      else if (val == "1" || val == "fp32") o.basic_mfma = 1;
      else if (val == "2" || val == "bf16") o.basic_mfma = 2;
      else if (val == "auto") o.basic_mfma = -1;
      else
        throw InputError("sus2_sh_basic_mfma must be auto, 0 (off), 1 (fp32) or 2 (bf16): " + val);
    } else if (key == "sus2_sh_graph_tile") {
      const int t = std::atoi(val.c_str());
      if (t != 1 && t != 2 && t != 4 && t != 8 && t != 16)
        throw InputError("sus2_sh_graph_tile must be 1, 2, 4, 8 or 16: " + val);
      o.graph_tile = t;
    }
  }
  if (!(pruning == "legacy" || pruning == "Legacy" || pruning == "LEGACY" || pruning == "q-total" ||
        pruning == "q_total" || pruning == "total" || pruning == "QTotal" || pruning == "QTOTAL"))
    throw InputError("Invalid SUS2_SH factor pruning mode: " + pruning +
                     " (expected legacy or q-total).");
  if (!radial_direct)
    throw InputError("SUS2_SH first backend currently requires direct radial evaluation.");
  if (o.lane_nslot > 0) {
    if (!lane_shape_supported(o.lane_waves, o.lane_nslot, o.lane_atoms))
      throw InputError("sus2_sh_lane_shape: no lane kernel for this waves x accumulators pair");
    if (o.lane_min_slots < 1 || o.lane_min_slots > kLaneMaxSlot + 1)
      throw InputError("sus2_sh_lane_shape: slot budget out of range");
  }
  return o;
}

}  // namespace

DeviceBuffer::~DeviceBuffer()
{
  if (ptr_) (void)hipFree(ptr_);
}

void DeviceBuffer::resize(size_t bytes)
{
  if (bytes <= bytes_ && ptr_) return;
  if (ptr_) hip_check(hipFree(ptr_), "hipFree");
  ptr_ = nullptr;
  bytes_ = 0;
  if (bytes == 0) return;
  hip_check(hipMalloc(&ptr_, bytes), "hipMalloc");
  bytes_ = bytes;
}

void DeviceBuffer::upload(const void* host, size_t bytes)
{
  resize(std::max<size_t>(bytes, 16));
  if (bytes) hip_check(hipMemcpy(ptr_, host, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

Engine::Engine(const std::string& path, int num_atoms, int nopts, const char** opts)
{
  model_ = load_model(path);
  if (model_.l_max > kMaxL)
    throw UnsupportedError("SUS2_SH: this engine evaluates sh_l_max <= 6 only.");
  // k chunks of at most kMaxK (sizes differ by at most one)
  kchunks_ = (model_.k_max + kMaxK - 1) / kMaxK;
  if (kchunks_ > kMaxKChunks)
    throw UnsupportedError("SUS2_SH: this engine supports sh_k_max <= 48.");
  for (int c = 0, k0 = 0; c < kchunks_; ++c) {
    kchunk_k0_[c] = k0;
    kchunk_K_[c] = model_.k_max / kchunks_ + (c < model_.k_max % kchunks_ ? 1 : 0);
    k0 += kchunk_K_[c];
  }
  if (model_.gate_enabled && kchunks_ > 1)
    throw UnsupportedError("SUS2_SH: two-layer gate models are evaluated for sh_k_max <= 6 only.");
  if (model_.species_count > kMaxSpecies)
    throw UnsupportedError("SUS2_SH: this engine supports at most 16 species.");
  opt_ = parse_options(model_, nopts, opts);
  if (opt_.deterministic && model_.gate_enabled)
    throw UnsupportedError("SUS2_SH: sus2_sh_deterministic=1 is not available for two-layer gate "
                           "models.");
  prog_ = compile_graph(model_, graph_chunk_terms());
  if (model_.zbl_enabled) {
    // atomic numbers and pair cutoffs (src/force/sus2_sh.cu:7824-7860); the neighbour
    // cutoff of the reference is max(max_dist, largest ZBL outer cutoff)
    zbl_ = build_zbl_tables(model_, nopts, opts);
    if (zbl_.outer_max > model_.max_dist)
      throw UnsupportedError("SUS2_SH: ZBL outer cutoff beyond max_dist is not supported by "
                             "this engine.");
  }

  hip_check(hipGetDevice(&device_), "hipGetDevice");
  // the allocation-free evaluation's scratch array: its address on this device, looked up
  // once per device (compute_noalloc then makes no runtime call to find it)
  hip_check(scratch_init_current_device(), "hipGetSymbolAddress(scratch)");
  int lds = 0;
  hip_check(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device_),
            "hipDeviceGetAttribute");
  max_lds_ = static_cast<size_t>(lds);
  hip_check(hipDeviceGetAttribute(&cu_count_, hipDeviceAttributeMultiprocessorCount, device_),
            "hipDeviceGetAttribute");

  const size_t rsz = opt_.use_float ? sizeof(float) : sizeof(double);
  // LDS radial tables of the edge kernels: per pair, coefficients and (a, s), each padded by
  // at most 31 words (1 mod 32, or 4 mod 32 for the 16-byte rows; edge_kernels.hip)
  const auto padded = [](size_t n) { return n + 31; };
  // (rows of the uniform kernels are zero-padded to 8, 12 or 16 coefficients)
  const size_t row = model_.rb_size <= 8 ? 8 : (model_.rb_size <= 12 ? 12 : 16);
  const size_t mr_chunk = static_cast<size_t>(kchunk_K_[0]) * (model_.l_max + 1);  // largest
  const size_t table = static_cast<size_t>(model_.species_count) * model_.species_count *
                       (padded(mr_chunk * row) + padded(mr_chunk * 2)) * rsz;
  if (table > 65536)
    throw UnsupportedError("SUS2_SH: radial tables exceed 64 KiB of LDS (too many species).");

  // graph tile: largest power of two whose LDS footprint leaves room for 2 workgroups/CU
  const auto lds_of = [&](int t) {
    return opt_.use_float ? graph_lds_bytes<float>(t, prog_.slot_count)
                          : graph_lds_bytes<double>(t, prog_.slot_count);
  };
  if (opt_.graph_variant >= 0) variant_ = opt_.graph_variant;
  if (opt_.graph_tile > 0) {
    tuned_ = true;  // a fixed tile disables the autotuner
    tile_ = opt_.graph_tile;
    if (lds_of(tile_) > max_lds_)
      throw InputError("sus2_sh_graph_tile too large for this model's on-chip graph state.");
  } else {
    const size_t budget = std::min<size_t>(max_lds_, 64 * 1024);
    tile_ = 1;
    for (int t : {16, 8, 4, 2, 1})
      if (lds_of(t) <= budget) {
        tile_ = t;
        break;
      }
    if (lds_of(tile_) > max_lds_)
      throw UnsupportedError("SUS2_SH: graph state of one atom exceeds the LDS of a workgroup.");
  }

  compile_lane();
  block_ = build_block(opt_.block_waves > 0 ? opt_.block_waves : block_waves_list()[0]);
  if (opt_.graph_variant == 4) {
    if (!block_.ok)
      throw InputError("sus2_sh_graph_variant=4: no block program for this model (" + block_.reason +
                       ").");
    if (opt_.deterministic)
      throw InputError("sus2_sh_graph_variant=4 sums basic adjoints with atomics; it is not "
                       "available with sus2_sh_deterministic=1.");
    tuned_ = true;  // a fixed variant disables the autotuner
  }
  if (opt_.graph_variant == 3) {
    if (!lane_.ok)
      throw InputError("sus2_sh_graph_variant=3: no lane program for this model (" + lane_.reason +
                       ").");
    tuned_ = true;  // a fixed variant disables the autotuner
  }
  if (opt_.deterministic) {
    // a timed choice of kernel could differ between runs (and so could its rounding): the
    // lane program when the model has one, else the balanced-stream row kernel
    tuned_ = true;
    if (opt_.graph_variant < 0) variant_ = lane_.ok ? 3 : 1;
  } else if (opt_.graph_variant < 0) {
    // until the autotuner has timed the candidates (or when it does not run): the block
    // program, else the lane program, else the balanced-stream row kernel
    variant_ = block_.ok ? 4 : (lane_.ok ? 3 : 1);
  }
  // edge-kernel shapes: fixed ones stay; deterministic mode keeps one lane per atom (one
  // summation order for every atom count), else the autotuner or the size rule decides
  if (opt_.basic_threads > 0) basic_shape_ = EdgeShape{opt_.basic_threads, opt_.basic_npl};
  if (opt_.force_threads > 0) force_shape_ = EdgeShape{opt_.force_threads, opt_.force_npl};
  if (opt_.deterministic || (opt_.basic_threads > 0 && opt_.force_threads > 0)) edge_tuned_ = true;
  if (opt_.use_float)
    upload_tables<float>();
  else
    upload_tables<double>();
  if (opt_.basic_mfma > 0 && !mfma_ok_)
    throw InputError("sus2_sh_basic_mfma: the matrix-core basic moments need the FP32 engine, "
                     "one radial scaling (a, s) per species pair, no gate, l_max <= 3, "
                     "k_max (l_max + 1) <= 16 and radial_basis_size <= 16.");
  if (opt_.basic_mfma >= 0) basic_shape_.mfma = opt_.basic_mfma;
  if (opt_.basic_mfma > 0 && opt_.force_threads > 0) edge_tuned_ = true;
  if (model_.zbl_enabled) {
    zbl_z_.upload(zbl_.atomic_number.data(), zbl_.atomic_number.size() * sizeof(int));
    zbl_inner_.upload(zbl_.inner.data(), zbl_.inner.size() * sizeof(double));
    zbl_outer_.upload(zbl_.outer.data(), zbl_.outer.size() * sizeof(double));
  }
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&host_max_), 8 * sizeof(int), 0),
            "hipHostMalloc");
  latch_.resize(8 * sizeof(int));
  launch_fill_int(latch_.as<int>(), 0, 8, nullptr);
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  (void)num_atoms;
}

Engine::~Engine()
{
  for (hipEvent_t e : events_) (void)hipEventDestroy(e);
  if (md_event_) (void)hipEventDestroy(md_event_);
  if (md_stream_) (void)hipStreamDestroy(md_stream_);
  if (host_max_) (void)hipHostFree(host_max_);
  if (host_red_) (void)hipHostFree(host_red_);
}

int Engine::param_count() const
{
  const int S = model_.species_count;
  return 2 * S + model_.scalar_count + model_.radial_funcs_count * model_.rb_size + S +
         2 * S * S * model_.radial_funcs_count;
}

// Lane program of one shape: the smallest per-lane slot budget >= min_slots that holds every
// graph component (160 slots are 40 KiB of LDS per 64-atom FP32 tile).
const std::vector<int>& Engine::block_waves_list()
{
  static const std::vector<int> waves = {8, 16, 4};
  return waves;
}

BlockProgram Engine::build_block(int waves) const
{
  const size_t rsz = opt_.use_float ? sizeof(float) : sizeof(double);
  // Real slots left beside the basic adjoints (B x 64 doubles) in a workgroup's LDS, and in
  // half of it (two workgroups per CU: large degree classes are then halved if that fits)
  const size_t g_bytes = static_cast<size_t>(model_.basic_count) * 64 * sizeof(double);
  const auto slots_in = [&](size_t lds) {
    return lds > g_bytes ? static_cast<int>((lds - g_bytes) / (64 * rsz)) : 0;
  };
  // (halves are compiled for the FP32 kernel only)
  BlockProgram p = compile_block_program(model_, waves, static_cast<int>(rsz), slots_in(max_lds_),
                                         opt_.use_float ? slots_in(max_lds_ / 2) : 0);
  if (p.ok) {
    BlockTable bt{};
    bt.basics = model_.basic_count;
    bt.slots = p.slots;
    bt.v_base = p.v_base;
    bt.waves = waves;
    if (block_graph_lds(bt, static_cast<int>(rsz)) > max_lds_) {
      p.ok = false;
      p.reason = "block program exceeds the LDS of a workgroup";
    }
  }
  return p;
}

void Engine::upload_block()
{
  block_seg_.upload(block_.seg.data(), block_.seg.size() * sizeof(int));
  block_words_.upload(block_.words.data(), std::max<size_t>(block_.words.size(), 1) * sizeof(uint32_t));
}

LaneProgram Engine::build_lane(const LaneShape& shape) const
{
  const size_t rsz = opt_.use_float ? sizeof(float) : sizeof(double);
  LaneProgram best;
  best.reason = "no slot budget fits the LDS of a workgroup";
  // 208 FP32 slots: three 64-atom tiles (52 KiB each) still share one CU's 160 KiB
  for (int max_slots : {160, 192, 208, 256, 320, 384, 512, 640, 1024}) {
    if (max_slots < shape.min_slots) continue;
    if (static_cast<size_t>(max_slots) * 64 * shape.atoms * rsz > max_lds_) break;
    LaneProgram p = compile_lane_program(model_, max_slots, shape.nslot, shape.waves);
    p.atoms = shape.atoms;
    if (p.ok) return p;
    best.reason = p.reason;
    if (p.reason.find("slot budget") == std::string::npos) break;  // not a budget limit
  }
  return best;
}

// Edge-kernel shapes the autotuner tries: threads per workgroup x lanes per atom.
const std::vector<EdgeShape>& Engine::edge_candidates()
{
  static const std::vector<EdgeShape> c = {{256, 1}, {256, 2}, {256, 4}, {256, 8},
                                           {128, 1}, {128, 2}, {128, 4}, {128, 8},
                                           {64, 1},  {64, 2},  {64, 4},  {64, 8}};
  return c;
}

// Shape of the edge kernels before (or without) the autotuner: as many lanes per atom as
// keep the lane count N * npl at or below 2^17 (two waves per SIMD of a 256-CU device), so
// that small structures still fill the device; one lane per atom from 131,072 atoms on.
EdgeShape Engine::default_edge_shape(int N)
{
  EdgeShape e;
  while (e.npl < 8 && static_cast<long long>(N) * e.npl * 2 <= (1LL << 17)) e.npl *= 2;
  return e;
}

// Shapes the autotuner tries (the first feasible one is the default): waves per tile x
// accumulators per wave x minimum slot budget x atoms per lane.  Two atoms per lane share
// each LDS address and each program record (half the record traffic per atom) at twice the
// LDS per tile.  Sixteen waves per tile (1024-thread workgroups) keep four waves per SIMD
// when one tile fills the CU's LDS (three-layer graphs need 384 FP32 slots, 96 KiB).
const std::vector<Engine::LaneShape>& Engine::lane_shapes()
{
  static const std::vector<LaneShape> shapes = {
    {4, 4, 160, 1}, {8, 2, 160, 1}, {4, 2, 160, 1}, {8, 2, 192, 1}, {8, 2, 256, 1},
    {2, 4, 160, 1}, {8, 2, 208, 1}, {8, 2, 160, 2}, {4, 2, 160, 2}, {8, 2, 208, 2},
    {16, 2, 192, 1}, {16, 2, 320, 1}};
  return shapes;
}

void Engine::compile_lane()
{
  if (opt_.lane_nslot > 0) {
    lane_ = build_lane(
      LaneShape{opt_.lane_waves, opt_.lane_nslot, opt_.lane_min_slots, opt_.lane_atoms});
    return;
  }
  lane_ = LaneProgram{};
  for (const LaneShape& sh : lane_shapes()) {
    lane_ = build_lane(sh);
    if (lane_.ok) return;
  }
}

void Engine::upload_lane()
{
  const LaneRecords rec =
    encode_lane_records(lane_, opt_.use_float ? static_cast<int>(sizeof(float)) : static_cast<int>(sizeof(double)),
                        64 * lane_.atoms);
  const size_t nseg = lane_.seg_kind.size();
  std::vector<int4> segs(nseg);
  for (size_t k = 0; k < nseg; ++k)
    segs[k] = make_int4(lane_.seg_kind[k], rec.seg_first[k], lane_.seg_groups[k],
                        lane_.seg_sync[k]);
  lane_segs_.upload(segs.data(), segs.size() * sizeof(int4));
  lane_ins_.upload(rec.words.data(), rec.words.size() * sizeof(uint32_t));
}

template <typename Real>
void Engine::upload_tables()
{
  const Model& m = model_;
  const GraphProgram& g = prog_;
  std::vector<double> coeff, scal;
  build_radial_tables(m, coeff, scal);
  const int L = m.l_max, K = m.k_max, Mr = m.radial_funcs_count, rb = m.rb_size;
  const int pairs = m.species_count * m.species_count;
  for (int c = 0; c < kchunks_; ++c) {
    // the chunk's radial rows mu in [k0 (L+1), (k0 + Kc)(L+1)) of every pair
    const int mu0 = kchunk_k0_[c] * (L + 1), mrc = kchunk_K_[c] * (L + 1);
    std::vector<Real> cR(static_cast<size_t>(pairs) * mrc * rb), sR(static_cast<size_t>(pairs) * mrc * 2);
    for (int p = 0; p < pairs; ++p)
      for (int mu = 0; mu < mrc; ++mu) {
        const size_t src = static_cast<size_t>(p) * Mr + mu0 + mu, dst = static_cast<size_t>(p) * mrc + mu;
        for (int n = 0; n < rb; ++n) cR[dst * rb + n] = static_cast<Real>(coeff[src * rb + n]);
        sR[dst * 2] = static_cast<Real>(scal[src * 2]);
        sR[dst * 2 + 1] = static_cast<Real>(scal[src * 2 + 1]);
      }
    radial_coeff_[c].upload(cR.data(), cR.size() * sizeof(Real));
    radial_scal_[c].upload(sR.data(), sR.size() * sizeof(Real));
  }
  // uniform kernels (one map for every mu, Chebyshev rows) are for the Chebyshev bases
  radial_uniform_ = m.radial_kind != RadialKind::LaguerreLog1p;
  for (int p = 0; p < pairs && radial_uniform_; ++p)
    for (int mu = 1; mu < m.radial_funcs_count; ++mu)
      if (scal[(p * m.radial_funcs_count + mu) * 2] != scal[(p * m.radial_funcs_count) * 2] ||
          scal[(p * m.radial_funcs_count + mu) * 2 + 1] != scal[(p * m.radial_funcs_count) * 2 + 1]) {
        radial_uniform_ = false;
        break;
      }
  // every pair's coefficient rows are one shared matrix times the pair factor sigma_p
  // (build_radial_tables): c_p[mu][n] = sigma_p c[mu][n]
  {
    std::vector<Real> sig(pairs);
    for (int ti = 0; ti < m.species_count; ++ti)
      for (int tj = 0; tj < m.species_count; ++tj)
        sig[ti * m.species_count + tj] = static_cast<Real>(
          m.scaling * m.radial_type_coeffs[ti] * m.radial_type_coeffs[tj]);
    pair_sigma_.upload(sig.data(), sig.size() * sizeof(Real));
    // the uniform edge kernels' rows, zero-padded to their row length (8, 12 or 16)
    shared_nb_ = rb <= 8 ? 8 : (rb <= 12 ? 12 : 16);
    for (int c = 0; c < kchunks_; ++c) {
      const int mu0 = kchunk_k0_[c] * (L + 1), mrc = kchunk_K_[c] * (L + 1);
      std::vector<Real> sh(static_cast<size_t>(mrc) * shared_nb_, Real(0));
      for (int mu = 0; mu < mrc; ++mu)
        for (int n = 0; n < rb; ++n)
          sh[static_cast<size_t>(mu) * shared_nb_ + n] =
            static_cast<Real>(m.radial_coeffs[static_cast<size_t>(mu0 + mu) * rb + n]);
      shared_rc_[c].upload(sh.data(), sh.size() * sizeof(Real));
    }
  }
  // matrix-core basic moments: the shared matrix unpadded ([Mr][rb])
  mfma_ok_ = sizeof(Real) == sizeof(float) && kchunks_ == 1 &&
             mfma_basics_supported(L, K, rb, m.species_count, radial_uniform_, m.gate_enabled);
  if (mfma_ok_) {
    std::vector<Real> rcm(static_cast<size_t>(Mr) * rb);
    for (size_t i = 0; i < rcm.size(); ++i) rcm[i] = static_cast<Real>(m.radial_coeffs[i]);
    mfma_rc_.upload(rcm.data(), rcm.size() * sizeof(Real));
  }

  // each chunk's full (k,l,m) grid -> basic index (the kernels' full_index<L, Kc> with the
  // chunk-local k); a basic belongs to the chunk that holds its k
  (void)K;
  for (int c = 0; c < kchunks_; ++c) {
    const int k0 = kchunk_k0_[c], Kc = kchunk_K_[c];
    std::vector<int> fmap(Kc * (L + 1) * (L + 1), -1);
    for (int b = 0; b < m.basic_count; ++b) {
      const int k = m.basic_k[b] - k0;
      if (k < 0 || k >= Kc) continue;
      int off = 0;
      for (int lp = L; lp > m.basic_l[b]; --lp) off += Kc * (2 * lp + 1);
      fmap[off + (Kc - 1 - k) * (2 * m.basic_l[b] + 1) + m.basic_m[b] + m.basic_l[b]] = b;
    }
    fmap_[c].upload(fmap.data(), fmap.size() * sizeof(int));
  }

  const int rows = static_cast<int>(g.fwd_row_target.size());
  std::vector<int4> frows(rows);
  std::vector<Real> fw(rows);
  for (int r = 0; r < rows; ++r) {
    frows[r] = make_int4(g.fwd_row_target[r], g.fwd_row_begin[r], g.fwd_row_count[r], 0);
    fw[r] = static_cast<Real>(g.fwd_row_weight[r]);
  }
  fwd_rows_.upload(frows.data(), frows.size() * sizeof(int4));
  fwd_w_.upload(fw.data(), fw.size() * sizeof(Real));
  std::vector<Term<Real>> ft(g.fwd_term_idx.size());
  for (size_t t = 0; t < ft.size(); ++t) {
    ft[t] = Term<Real>{};
    ft[t].idx = g.fwd_term_idx[t];
    ft[t].c = static_cast<Real>(g.fwd_term_coeff[t]);
  }
  fwd_terms_.upload(ft.data(), ft.size() * sizeof(Term<Real>));
  fwd_layer_.upload(g.fwd_layer_begin.data(), g.fwd_layer_begin.size() * sizeof(int));

  const int chunks = static_cast<int>(g.bwd_chunk_src.size());
  std::vector<int4> bch(chunks);
  for (int c = 0; c < chunks; ++c)
    bch[c] = make_int4(g.bwd_chunk_src[c], g.bwd_chunk_begin[c], g.bwd_chunk_count[c],
                       g.bwd_chunk_atomic[c]);
  bwd_chunks_.upload(bch.data(), bch.size() * sizeof(int4));
  std::vector<Term<Real>> bt(g.bwd_term_idx.size());
  for (size_t t = 0; t < bt.size(); ++t) {
    bt[t] = Term<Real>{};
    bt[t].idx = g.bwd_term_idx[t];
    bt[t].c = static_cast<Real>(g.bwd_term_coeff[t]);
  }
  bwd_terms_.upload(bt.data(), bt.size() * sizeof(Term<Real>));
  bwd_layer_.upload(g.bwd_layer_begin.data(), g.bwd_layer_begin.size() * sizeof(int));

  auto upload_stream = [&](const std::vector<uint32_t>& ab, const std::vector<double>& c,
                           DeviceBuffer& dst) {
    std::vector<Term<Real>> v(ab.size());
    for (size_t k = 0; k < v.size(); ++k) {
      v[k] = Term<Real>{};
      v[k].idx = ab[k];
      v[k].c = static_cast<Real>(c[k]);
    }
    dst.upload(v.data(), v.size() * sizeof(Term<Real>));
  };
  for (int v = 0; v < 2; ++v) {
    const GraphProgram::Streams& st = g.streams[v];
    upload_stream(st.fwd_ab, st.fwd_c, fwd_ins_[v]);
    upload_stream(st.bwd_ab, st.bwd_c, bwd_ins_[v]);
    fwd_sbegin_[v].upload(st.fwd_begin.data(), st.fwd_begin.size() * sizeof(int));
    fwd_slen_[v].upload(st.fwd_len.data(), st.fwd_len.size() * sizeof(int));
    bwd_sbegin_[v].upload(st.bwd_begin.data(), st.bwd_begin.size() * sizeof(int));
    bwd_slen_[v].upload(st.bwd_len.data(), st.bwd_len.size() * sizeof(int));
  }
  e_slot_.upload(g.energy_slot.data(), g.energy_slot.size() * sizeof(int));
  std::vector<Real> ew(g.energy_weight.begin(), g.energy_weight.end());
  e_w_.upload(ew.data(), ew.size() * sizeof(Real));
  std::vector<Real> seed(g.seed.begin(), g.seed.end());
  seed_.upload(seed.data(), seed.size() * sizeof(Real));
  shift_.upload(m.shift_coeffs.data(), m.shift_coeffs.size() * sizeof(double));
  sigma_.upload(m.species_coeffs.data(), m.species_coeffs.size() * sizeof(double));
  row_sbegin_.upload(g.row_scalar_begin.data(), g.row_scalar_begin.size() * sizeof(int));
  row_slist_.upload(g.row_scalar_list.data(), g.row_scalar_list.size() * sizeof(int));
  stored_index_.upload(g.stored_scalar_index.data(), g.stored_scalar_index.size() * sizeof(int));
  stored_slot_.upload(g.stored_scalar_slot.data(), g.stored_scalar_slot.size() * sizeof(int));
  if (lane_.ok) upload_lane();
  if (block_.ok) upload_block();
  if (m.gate_enabled) upload_gate_tables<Real>();
}

// Gate layer tables (two-layer mu gate): the gate radial coefficients in the main table's
// layout, the gate program -- the model's graph compiled for the gate scalars, each stored
// in a slot -- and h = weff s with weff[mu][q] the mu-scalar-full weights or, for
// mu-body-linear-combo, c_{mu, body(q)} w_q (h = sum_p c_{mu,p} sum_{q in p} w_q s_q).
template <typename Real>
void Engine::upload_gate_tables()
{
  const Model& m = model_;
  std::vector<double> coeff, scal;
  build_radial_tables(m, coeff, scal, &m.gate_radial_coeffs);
  std::vector<Real> cR(coeff.begin(), coeff.end());
  gate_coeff_.upload(cR.data(), cR.size() * sizeof(Real));
  {
    // shared rows of the gate coefficients (gate models have one k chunk)
    const int Mr0 = m.radial_funcs_count, rb0 = m.rb_size;
    const int nb = rb0 <= 8 ? 8 : (rb0 <= 12 ? 12 : 16);
    std::vector<Real> sh(static_cast<size_t>(Mr0) * nb, Real(0));
    for (int mu = 0; mu < Mr0; ++mu)
      for (int n = 0; n < rb0; ++n)
        sh[static_cast<size_t>(mu) * nb + n] =
          static_cast<Real>(m.gate_radial_coeffs[static_cast<size_t>(mu) * rb0 + n]);
    gate_shared_.upload(sh.data(), sh.size() * sizeof(Real));
  }

  Model gm = m;  // same graph, the gate scalars as its scalar list
  const int Qg = static_cast<int>(m.gate_moment_index.size());
  gm.scalar_count = Qg;
  gm.moment_mapping = m.gate_moment_index;
  gm.moment_coeffs.assign(Qg, 1.0);
  gate_prog_ = compile_graph(gm, 8, 256, true);
  const GraphProgram& g = gate_prog_;
  for (int t : g.fwd_row_target)
    if (t < 0) throw DeviceError("gate program: a row without a slot");
  for (uint32_t idx : g.bwd_term_idx)
    if ((idx >> 16) == kNoTarget) throw DeviceError("gate program: a terminal back term");
  std::vector<int> slot(Qg);
  for (int q = 0; q < Qg; ++q) {
    slot[q] = g.node_slot[m.gate_moment_index[q]];
    if (slot[q] < 0) throw DeviceError("gate program: a gate scalar without a slot");
  }
  const int Mr = m.radial_funcs_count, K = m.gate_body_count;
  std::vector<Real> weff(static_cast<size_t>(Mr) * Qg);
  for (int mu = 0; mu < Mr; ++mu)
    for (int q = 0; q < Qg; ++q)
      weff[static_cast<size_t>(mu) * Qg + q] = static_cast<Real>(
        m.gate_mode == kGateScalarFull
          ? m.gate_weights[static_cast<size_t>(mu) * Qg + q]
          : m.gate_body_mix[static_cast<size_t>(mu) * K + m.gate_body_id[q]] * m.gate_weights[q]);
  gate_weff_.upload(weff.data(), weff.size() * sizeof(Real));
  gate_slot_.upload(slot.data(), slot.size() * sizeof(int));
  gate_add_.upload(m.gate_additive.data(), m.gate_additive.size() * sizeof(double));

  const int rows = static_cast<int>(g.fwd_row_target.size());
  std::vector<int4> frows(rows);
  for (int r = 0; r < rows; ++r)
    frows[r] = make_int4(g.fwd_row_target[r], g.fwd_row_begin[r], g.fwd_row_count[r], 0);
  g_fwd_rows_.upload(frows.data(), frows.size() * sizeof(int4));
  std::vector<Term<Real>> ft(g.fwd_term_idx.size());
  for (size_t t = 0; t < ft.size(); ++t) {
    ft[t] = Term<Real>{};
    ft[t].idx = g.fwd_term_idx[t];
    ft[t].c = static_cast<Real>(g.fwd_term_coeff[t]);
  }
  g_fwd_terms_.upload(ft.data(), ft.size() * sizeof(Term<Real>));
  g_fwd_layer_.upload(g.fwd_layer_begin.data(), g.fwd_layer_begin.size() * sizeof(int));
  const int chunks = static_cast<int>(g.bwd_chunk_src.size());
  std::vector<int4> bch(chunks);
  for (int c = 0; c < chunks; ++c)
    bch[c] = make_int4(g.bwd_chunk_src[c], g.bwd_chunk_begin[c], g.bwd_chunk_count[c],
                       g.bwd_chunk_atomic[c]);
  g_bwd_chunks_.upload(bch.data(), bch.size() * sizeof(int4));
  std::vector<Term<Real>> bt(g.bwd_term_idx.size());
  for (size_t t = 0; t < bt.size(); ++t) {
    bt[t] = Term<Real>{};
    bt[t].idx = g.bwd_term_idx[t];
    bt[t].c = static_cast<Real>(g.bwd_term_coeff[t]);
  }
  g_bwd_terms_.upload(bt.data(), bt.size() * sizeof(Term<Real>));
  g_bwd_layer_.upload(g.bwd_layer_begin.data(), g.bwd_layer_begin.size() * sizeof(int));

  // tile: the largest power of two whose values and adjoints fit 64 KiB of LDS
  const size_t budget = std::min<size_t>(max_lds_, 64 * 1024);
  gate_tile_ = 0;
  for (int t : {16, 8, 4, 2, 1})
    if (gate_graph_lds_bytes(t, g.slot_count, static_cast<int>(sizeof(Real))) <= budget) {
      gate_tile_ = t;
      break;
    }
  if (gate_tile_ == 0) {
    if (gate_graph_lds_bytes(1, g.slot_count, static_cast<int>(sizeof(Real))) > max_lds_)
      throw UnsupportedError("SUS2_SH: the gate layer's graph state exceeds the LDS of a workgroup.");
    gate_tile_ = 1;
  }
}

void Engine::ensure_atoms(int N, size_t edge_bytes)
{
  const size_t n = static_cast<size_t>(N);
  const size_t rsz = opt_.use_float ? sizeof(float) : sizeof(double);
  if (N > atoms_alloc_) {
    key_.resize(n * sizeof(int));
    slot_.resize(n * sizeof(int));
    arrival_.resize(n * sizeof(int));
    skey_.resize(n * sizeof(int));
    satom_.resize(n * sizeof(int));
    rank_.resize(n * sizeof(int));
    wrapped_.resize(3 * n * sizeof(double));
    pos_s_.resize(3 * n * sizeof(double));
    packed_.resize(n * sizeof(PackedAtom));
    type_s_.resize(n * sizeof(int));
    count_.resize(n * sizeof(int));
    basics_.resize(static_cast<size_t>(model_.basic_count) * n * rsz);
    grads_.resize(static_cast<size_t>(model_.basic_count) * n * rsz);
    e_s_.resize(n * sizeof(double));
    f_s_.resize(3 * n * sizeof(double));
    v_s_.resize(9 * n * sizeof(double));
    if (model_.gate_enabled) {
      const size_t nm = n * static_cast<size_t>(model_.radial_funcs_count) * rsz;
      gbasics_.resize(static_cast<size_t>(model_.basic_count) * n * rsz);
      ggrads_.resize(static_cast<size_t>(model_.basic_count) * n * rsz);
      gate_m_.resize(nm);
      gate_d_.resize(nm);
      gate_adj_.resize(nm);
    }
    atoms_alloc_ = N;
  }
  max_count_.resize(64);
  nl_.resize(static_cast<size_t>(cap_) * n * edge_bytes);
}

void Engine::bind_buffers(int N, int ncell, int n_totals)
{
  (void)N;
  (void)ncell;
  (void)n_totals;
  ws_.key = key_.as<int>();
  ws_.slot = slot_.as<int>();
  ws_.arrival = arrival_.as<int>();
  ws_.skey = skey_.as<int>();
  ws_.satom = satom_.as<int>();
  ws_.rank = rank_.as<int>();
  ws_.type_s = type_s_.as<int>();
  ws_.count = count_.as<int>();
  ws_.wrapped = wrapped_.as<double>();
  ws_.pos_s = pos_s_.as<double>();
  ws_.e_s = e_s_.as<double>();
  ws_.f_s = f_s_.as<double>();
  ws_.v_s = v_s_.as<double>();
  ws_.packed = packed_.as<PackedAtom>();
  ws_.nl = nl_.as<void>();
  ws_.basics = basics_.as<void>();
  ws_.grads = grads_.as<void>();
  ws_.cell_count = cell_count_.as<int>();
  ws_.cell_begin = cell_begin_.as<int>();
  ws_.scan_totals = scan_totals_.as<int>();
  ws_.max_count = max_count_.as<int>();
  ws_.win = win_.as<int>();
  ws_.gbasics = gbasics_.as<void>();
  ws_.ggrads = ggrads_.as<void>();
  ws_.gate_m = gate_m_.as<void>();
  ws_.gate_d = gate_d_.as<void>();
  ws_.gate_adj = gate_adj_.as<void>();
  ws_.cap = cap_;
}

// Work buffers of one evaluation as 256-byte aligned pieces of the device's scratch array;
// what is left after the per-atom and per-cell buffers holds the neighbour list, whose
// capacity (neighbours per atom) follows from it.  Nothing here allocates or calls the HIP
// runtime.
void Engine::carve_scratch(int N, int ncell, int n_totals, size_t edge_bytes, int win_words)
{
  unsigned char* base = static_cast<unsigned char*>(scratch_device_ptr(device_));
  if (!base) throw DeviceError("compute_noalloc: no scratch array on this device");
  const size_t n = static_cast<size_t>(N);
  const size_t rsz = opt_.use_float ? sizeof(float) : sizeof(double);
  size_t off = 0;
  auto take = [&](size_t bytes) -> void* {
    void* p = base + off;
    off += (bytes + 255) & ~size_t(255);
    return p;
  };
  Workspace w;
  w.key = static_cast<int*>(take(n * sizeof(int)));
  w.slot = static_cast<int*>(take(n * sizeof(int)));
  w.arrival = static_cast<int*>(take(n * sizeof(int)));
  w.skey = static_cast<int*>(take(n * sizeof(int)));
  w.satom = static_cast<int*>(take(n * sizeof(int)));
  w.rank = static_cast<int*>(take(n * sizeof(int)));
  w.type_s = static_cast<int*>(take(n * sizeof(int)));
  w.count = static_cast<int*>(take(n * sizeof(int)));
  w.wrapped = static_cast<double*>(take(3 * n * sizeof(double)));
  w.pos_s = static_cast<double*>(take(3 * n * sizeof(double)));
  w.e_s = static_cast<double*>(take(n * sizeof(double)));
  w.f_s = static_cast<double*>(take(3 * n * sizeof(double)));
  w.v_s = static_cast<double*>(take(9 * n * sizeof(double)));
  w.packed = static_cast<PackedAtom*>(take(n * sizeof(PackedAtom)));
  w.basics = take(static_cast<size_t>(model_.basic_count) * n * rsz);
  w.grads = take(static_cast<size_t>(model_.basic_count) * n * rsz);
  w.cell_count = static_cast<int*>(take((static_cast<size_t>(ncell) + 1) * sizeof(int)));
  w.cell_begin = static_cast<int*>(take((static_cast<size_t>(ncell) + 1) * sizeof(int)));
  w.scan_totals = static_cast<int*>(take(static_cast<size_t>(n_totals) * sizeof(int)));
  w.max_count = static_cast<int*>(take(16 * sizeof(int)));
  w.win = static_cast<int*>(take(static_cast<size_t>(win_words) * sizeof(int)));
  const size_t left = off < kScratchBytes ? kScratchBytes - off : 0;
  const size_t cap = std::min<size_t>(left / (n * edge_bytes), 4096);
  if (off > kScratchBytes || cap < 1) {
    std::ostringstream msg;
    msg << "SUS2_SH: compute_noalloc: " << N << " atoms do not fit the "
        << (kScratchBytes >> 20) << " MiB scratch array of this device";
    throw ArgumentError(msg.str());
  }
  w.nl = base + off;
  w.cap = static_cast<int>(cap);
  ws_ = w;
}

CellGrid Engine::make_grid(const sus2sh_box& box, int N, int* ncell) const
{
  CellGrid g{};
  for (int k = 0; k < 9; ++k) {
    g.h[k] = box.h[k];
    g.hi[k] = box.h[9 + k];
  }
  const double* h = box.h;
  const double a[3] = {h[0], h[3], h[6]}, b[3] = {h[1], h[4], h[7]}, c[3] = {h[2], h[5], h[8]};
  const double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
  const double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
  const double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double vol = std::fabs(a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2]);
  if (!(vol > 0.0) || !std::isfinite(vol)) throw ArgumentError("box matrix is singular");
  const double t[3] = {vol / std::sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]),
                       vol / std::sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]),
                       vol / std::sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2])};
  const double rc = model_.max_dist;
  long total = 1;
  for (int d = 0; d < 3; ++d) {
    g.pbc[d] = box.pbc[d] ? 1 : 0;
    g.n[d] = std::max(1, static_cast<int>(std::floor(t[d] / rc)));
    g.n[d] = std::min(g.n[d], 1024);
    total *= g.n[d];
  }
  // keep the cell count bounded by the atom count (sparse systems)
  const long limit = std::max<long>(27, 4L * std::max(N, 1));
  while (total > limit) {
    int dmax = 0;
    for (int d = 1; d < 3; ++d)
      if (g.n[d] > g.n[dmax]) dmax = d;
    g.n[dmax] = std::max(1, g.n[dmax] / 2);
    total = static_cast<long>(g.n[0]) * g.n[1] * g.n[2];
  }
  for (int d = 0; d < 3; ++d)
    g.R[d] = std::max(1, static_cast<int>(std::ceil(rc * g.n[d] / t[d] - 1e-12)));
  *ncell = static_cast<int>(total);
  return g;
}

void Engine::compute(const sus2sh_box& box, int N, const int* type, const double* pos,
                     double* pot, double* force, double* virial, hipStream_t stream,
                     double* param_grad)
{
  if (N < 0) throw ArgumentError("num_atoms must be >= 0");
  if (N == 0) return;
  if (!type || !pos || !pot || !force || !virial)
    throw ArgumentError("null device pointer passed to compute");
  const int offsets[2] = {0, N};
  if (opt_.use_float)
    run<float>(1, &box, offsets, N, type, pos, pot, force, virial, stream, param_grad, false);
  else
    run<double>(1, &box, offsets, N, type, pos, pot, force, virial, stream, param_grad, false);
}

void Engine::compute_noalloc(const sus2sh_box& box, int N, const int* type, const double* pos,
                             double* pot, double* force, double* virial, hipStream_t stream)
{
  if (N < 0) throw ArgumentError("num_atoms must be >= 0");
  if (N == 0) return;
  if (!type || !pos || !pot || !force || !virial)
    throw ArgumentError("null device pointer passed to compute_noalloc");
  int dev = -1;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev != device_)
    throw ArgumentError("compute_noalloc: the current device is not the handle's device");
  const int offsets[2] = {0, N};
  if (opt_.use_float)
    run<float>(1, &box, offsets, N, type, pos, pot, force, virial, stream, nullptr, true);
  else
    run<double>(1, &box, offsets, N, type, pos, pot, force, virial, stream, nullptr, true);
}

void Engine::compute_batch(int nstruct, const sus2sh_box* boxes, const int* atom_offsets,
                           const int* type, const double* pos, double* pot, double* force,
                           double* virial, hipStream_t stream, double* param_grad)
{
  if (nstruct <= 0 || !boxes || !atom_offsets) throw ArgumentError("compute_batch: empty batch");
  if (atom_offsets[0] != 0) throw ArgumentError("compute_batch: atom_offsets[0] must be 0");
  for (int s = 0; s < nstruct; ++s)
    if (atom_offsets[s + 1] < atom_offsets[s])
      throw ArgumentError("compute_batch: atom_offsets must be non-decreasing");
  const int N = atom_offsets[nstruct];
  if (N == 0) return;
  if (!type || !pos || !pot || !force || !virial)
    throw ArgumentError("null device pointer passed to compute_batch");
  if (opt_.use_float)
    run<float>(nstruct, boxes, atom_offsets, N, type, pos, pot, force, virial, stream, param_grad,
               false);
  else
    run<double>(nstruct, boxes, atom_offsets, N, type, pos, pot, force, virial, stream, param_grad,
                false);
}

template <typename Real>
bool Engine::launch_graph_variant(int variant, int tile, int N, const GraphTable<Real>& gt,
                                  hipStream_t s)
{
  if (variant == 4) {
    if (!block_.ok) return false;
    BlockTable bt{};
    bt.seg = block_seg_.as<int>();
    bt.words = block_words_.as<uint32_t>();
    bt.shift = shift_.as<double>();
    bt.sigma = sigma_.as<double>();
    bt.batches = block_.batches;
    bt.basics = model_.basic_count;
    bt.v_base = block_.v_base;
    bt.slots = block_.slots;
    bt.waves = block_.waves;
    return launch_block_graph<Real>(N, static_cast<const Real*>(ws_.basics), ws_.type_s, bt,
                                    ws_.e_s, static_cast<Real*>(ws_.grads), s, max_lds_);
  }
  if (variant == 3) {
    if (!lane_.ok) return false;
    LaneTable<Real> lt{};
    lt.segs = lane_segs_.as<int4>();
    lt.recs = lane_ins_.as<uint32_t>();
    lt.shift = shift_.as<double>();
    lt.sigma = sigma_.as<double>();
    lt.nseg = static_cast<int>(lane_.seg_kind.size());
    lt.basics = model_.basic_count;
    lt.ones_slot = lane_.ones_slot;
    lt.v_base = lane_.v_base;
    lt.slots = lane_.slots;
    lt.nslot = lane_.nslot;
    lt.waves = lane_.waves;
    lt.atoms = lane_.atoms;
    return launch_lane_graph<Real>(N, static_cast<const Real*>(ws_.basics), ws_.type_s, lt,
                                   ws_.e_s, static_cast<Real*>(ws_.grads), s, max_lds_);
  }
  GraphTable<Real> t = gt;
  const int k = variant == 2 ? 1 : 0;
  t.fwd_ins = fwd_ins_[k].as<Term<Real>>();
  t.fwd_sbegin = fwd_sbegin_[k].as<int>();
  t.fwd_slen = fwd_slen_[k].as<int>();
  t.bwd_ins = bwd_ins_[k].as<Term<Real>>();
  t.bwd_sbegin = bwd_sbegin_[k].as<int>();
  t.bwd_slen = bwd_slen_[k].as<int>();
  return launch_graph<Real>(tile, variant == 0 ? 0 : 1, N, static_cast<const Real*>(ws_.basics),
                            ws_.type_s, t, ws_.e_s, static_cast<Real*>(ws_.grads), nullptr, s,
                            max_lds_, 0);
}

template <typename Real>
void Engine::run(int nstruct, const sus2sh_box* boxes, const int* atom_offsets, int N,
                 const int* type, const double* pos, double* pot, double* force, double* virial,
                 hipStream_t s, double* param_grad, bool noalloc)
{
  if (N >= kMaxEdgeAtoms)
    throw ArgumentError("SUS2_SH: at most 134217727 atoms per call (neighbour index width)");
  const bool gate = model_.gate_enabled;
  if (gate && param_grad)
    throw UnsupportedError("SUS2_SH: dE/dtheta is not implemented for two-layer gate models.");
  if (gate && noalloc)
    throw UnsupportedError("SUS2_SH: compute_noalloc does not evaluate two-layer gate models "
                           "(use sus2sh_compute).");
  if (param_grad && opt_.deterministic)
    throw UnsupportedError("SUS2_SH: dE/dtheta is not available with sus2_sh_deterministic=1.");
  if (param_grad && model_.radial_kind == RadialKind::LaguerreLog1p)
    throw UnsupportedError("SUS2_SH: dE/dtheta is not implemented for RBLaguerre_log1p models.");
  const bool batch = nstruct > 1;
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  hip_check(hipStreamIsCapturing(s, &capture), "hipStreamIsCapturing");
  const bool capturing = capture != hipStreamCaptureStatusNone;
  if (capturing && batch) throw ArgumentError("SUS2_SH: batched evaluations cannot be captured");
  // cell grids (host; the per-structure tables only for a batch, so that a single structure
  // makes no host allocation either)
  CellGrid grid{};
  int ncell = 0;
  std::vector<CellGrid> grids;
  std::vector<int> cell_off;
  if (!batch) {
    grid = make_grid(boxes[0], N, &ncell);
  } else {
    grids.resize(nstruct);
    cell_off.assign(nstruct + 1, 0);
    for (int st = 0; st < nstruct; ++st) {
      int nc = 0;
      grids[st] = make_grid(boxes[st], atom_offsets[st + 1] - atom_offsets[st], &nc);
      if (static_cast<long>(cell_off[st]) + nc > (1L << 30)) throw ArgumentError("too many cells");
      cell_off[st + 1] = cell_off[st] + nc;
    }
    ncell = cell_off[nstruct];
    grid = grids[0];
  }
  if (batch) {
    batch_grids_.upload(grids.data(), grids.size() * sizeof(CellGrid));
    batch_cell_off_.upload(cell_off.data(), cell_off.size() * sizeof(int));
    batch_atom_off_.upload(atom_offsets, (nstruct + 1) * sizeof(int));
    sid_.resize(static_cast<size_t>(N) * sizeof(int));
    sid_s_.resize(static_cast<size_t>(N) * sizeof(int));
  }
  // deferred: no host synchronisation; the neighbour capacity was validated by an earlier
  // synchronised call with the same atom count, and overflows are latched for check()
  const bool deferred = (opt_.deferred_check || md_capture_) && cap_known_ && N == cap_atoms_ &&
                        !batch && !param_grad && !noalloc;
  if (capturing && !deferred)
    throw ArgumentError(
      "SUS2_SH: a captured evaluation needs sus2_sh_deferred_check=1, one structure, no "
      "parameter gradient, and a preceding uncaptured call with the same atom count");
  const bool profile = opt_.profile && !capturing && !noalloc;
  if (cap_ == 0) cap_ = 64;
  const int n_totals = static_cast<int>(cell_sort_totals(ncell));
  // force-tile neighbour windows (no parameter gradient)
  const bool use_win = !param_grad;
  const int win_words =
    use_win ? static_cast<int>((static_cast<size_t>(N) + kWinTile - 1) / kWinTile) * kWinStride : 0;
  if (noalloc) {
    if (capturing) throw ArgumentError("SUS2_SH: compute_noalloc cannot be captured");
    carve_scratch(N, ncell, n_totals, sizeof(Edge<Real>), win_words);
  } else {
    if (capturing && (N > atoms_alloc_ || nl_.bytes() < static_cast<size_t>(cap_) * N * sizeof(Edge<Real>) ||
                      cell_begin_.bytes() < (static_cast<size_t>(ncell) + 1) * sizeof(int) ||
                      scan_totals_.bytes() < static_cast<size_t>(n_totals) * sizeof(int) ||
                      win_.bytes() < static_cast<size_t>(win_words) * sizeof(int)))
      throw ArgumentError("SUS2_SH: a captured evaluation cannot resize work buffers (the atom "
                          "count or the cell grid grew since the last uncaptured call)");
    ensure_atoms(N, sizeof(Edge<Real>));
    cell_count_.resize((static_cast<size_t>(ncell) + 1) * sizeof(int));
    cell_begin_.resize((static_cast<size_t>(ncell) + 1) * sizeof(int));
    scan_totals_.resize(static_cast<size_t>(n_totals) * sizeof(int));
    if (use_win) win_.resize(static_cast<size_t>(win_words) * sizeof(int));
    bind_buffers(N, ncell, n_totals);
  }

  if (profile && events_.empty()) {
    events_.resize(6);
    for (hipEvent_t& e : events_) hip_check(hipEventCreate(&e), "hipEventCreate");
  }
  if (profile) hip_check(hipEventRecord(events_[0], s), "hipEventRecord");

  launch_fill_int(ws_.max_count, 0, 4, s);
  CellSortArgs cs{};
  cs.N = N;
  cs.S = model_.species_count;
  cs.ncell = ncell;
  cs.pos = pos;
  cs.type = type;
  cs.grid = grid;
  cs.grids = batch ? batch_grids_.as<CellGrid>() : nullptr;
  cs.cell_off = batch ? batch_cell_off_.as<int>() : nullptr;
  cs.atom_off = batch ? batch_atom_off_.as<int>() : nullptr;
  cs.nstruct = nstruct;
  cs.key = ws_.key;
  cs.slot = ws_.slot;
  cs.arrival = ws_.arrival;
  cs.cell_count = ws_.cell_count;
  cs.cell_begin = ws_.cell_begin;
  cs.scan_totals = ws_.scan_totals;
  cs.wrapped = ws_.wrapped;
  cs.sid = batch ? sid_.as<int>() : nullptr;
  cs.satom = ws_.satom;
  cs.rank = ws_.rank;
  cs.skey = ws_.skey;
  cs.pos_s = ws_.pos_s;
  cs.packed = ws_.packed;
  cs.type_s = ws_.type_s;
  cs.bad_type = ws_.max_count + 1;
  cs.sid_s = batch ? sid_s_.as<int>() : nullptr;
  launch_cell_sort(cs, s);
  const int* cell_begin = ws_.cell_begin;
  const int* cell_end = cell_begin + 1;
  NeighborArgs na{};
  na.N = N;
  na.grid = grid;
  na.grids = batch ? batch_grids_.as<CellGrid>() : nullptr;
  na.cell_off = batch ? batch_cell_off_.as<int>() : nullptr;
  na.sid_s = batch ? sid_s_.as<int>() : nullptr;
  na.atom_off = batch ? batch_atom_off_.as<int>() : nullptr;
  na.rc2 = model_.max_dist * model_.max_dist;
  if (deferred) {
    na.cap = ws_.cap;
    launch_neighbors<Real>(na, ws_.pos_s, ws_.packed, ws_.skey, cell_begin, cell_end, ws_.count,
                           static_cast<Edge<Real>*>(ws_.nl), ws_.max_count, s);
    launch_latch(ws_.max_count, latch_.as<int>(), s);
  } else {
    for (int attempt = 0; attempt < 2; ++attempt) {
      na.cap = ws_.cap;
      if (attempt > 0) launch_fill_int(ws_.max_count, 0, 1, s);
      launch_fill_int(ws_.max_count + 2, 0, 2, s);
      launch_neighbors<Real>(na, ws_.pos_s, ws_.packed, ws_.skey, cell_begin, cell_end, ws_.count,
                             static_cast<Edge<Real>*>(ws_.nl), ws_.max_count, s);
      hip_check(hipMemcpyAsync(host_max_, ws_.max_count, 4 * sizeof(int),
                               hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync D2H");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
      const int maxc = host_max_[0];
      const unsigned long long total =
        static_cast<unsigned long long>(static_cast<unsigned>(host_max_[2])) |
        (static_cast<unsigned long long>(static_cast<unsigned>(host_max_[3])) << 32);
      mean_neighbors_ = static_cast<double>(total) / N;
      if (host_max_[1] != 0)
        throw ArgumentError("atom type outside [0, species_count) passed to compute");
      if (opt_.deterministic && maxc > kFixMaxNeighbors) throw ArgumentError(fix_range_message());
      if (maxc <= ws_.cap) break;
      if (noalloc) {
        std::ostringstream msg;
        msg << "SUS2_SH: compute_noalloc: " << maxc << " neighbours of one atom exceed the "
            << ws_.cap << " that the scratch array (" << (kScratchBytes >> 20)
            << " MiB) holds for " << N << " atoms";
        throw ArgumentError(msg.str());
      }
      cap_ = ((maxc + 8 + 7) / 8) * 8;
      ensure_atoms(N, sizeof(Edge<Real>));
      ws_.nl = nl_.as<void>();
      ws_.cap = cap_;
    }
    if (!batch && !noalloc) {
      cap_known_ = true;
      cap_atoms_ = N;
    }
  }
  if (profile) hip_check(hipEventRecord(events_[1], s), "hipEventRecord");

  EdgeLaunch<Real> el{};
  el.N = N;
  el.L = model_.l_max;
  el.K = kchunk_K_[0];
  el.B = model_.basic_count;
  el.count = ws_.count;
  el.nl = static_cast<const Edge<Real>*>(ws_.nl);
  el.type_s = ws_.type_s;
  el.rt.coeff = radial_coeff_[0].as<Real>();
  el.rt.scal = radial_scal_[0].as<Real>();
  el.rt.shared = shared_rc_[0].as<Real>();
  el.rt.sigma = pair_sigma_.as<Real>();
  el.rt.shared_nb = shared_nb_;
  el.rt.S = model_.species_count;
  el.rt.rb = model_.rb_size;
  el.rt.rc = static_cast<Real>(model_.max_dist);
  el.rt.kind = static_cast<int>(model_.radial_kind);  // 0 tanh, 1 rational, 2 Laguerre
  el.rt.uniform = radial_uniform_ ? 1 : 0;
  el.fmap = fmap_[0].as<int>();
  // the edge launch of k chunk c (chunk 0 is el itself)
  const auto chunk_launch = [&](int c) {
    EdgeLaunch<Real> ec = el;
    ec.K = kchunk_K_[c];
    ec.rt.coeff = radial_coeff_[c].as<Real>();
    ec.rt.scal = radial_scal_[c].as<Real>();
    ec.rt.shared = shared_rc_[c].as<Real>();
    ec.fmap = fmap_[c].as<int>();
    return ec;
  };
  el.fix_flag = opt_.deterministic ? latch_.as<int>() + kLatchFix : nullptr;
  // two-layer gate (src/force/sus2_sh.cu:9821-9893): gate moments with the gate radial
  // coefficients, the gate layer's graph, G = 1 + A tanh(a h) per atom and channel; the main
  // moments then weight every edge by the neighbour's G
  GraphTable<Real> gg{};
  GateTable<Real> ga{};
  if (gate) {
    EdgeLaunch<Real> eg = el;
    eg.rt.coeff = gate_coeff_.as<Real>();
    eg.rt.shared = gate_shared_.as<Real>();
    if (!launch_basics<Real>(eg, static_cast<Real*>(ws_.gbasics), s))
      throw UnsupportedError("SUS2_SH: no basic-moment kernel for this (l_max, k_max).");
    gg.fwd_rows = g_fwd_rows_.as<int4>();
    gg.fwd_terms = g_fwd_terms_.as<Term<Real>>();
    gg.fwd_layer = g_fwd_layer_.as<int>();
    gg.bwd_chunks = g_bwd_chunks_.as<int4>();
    gg.bwd_terms = g_bwd_terms_.as<Term<Real>>();
    gg.bwd_layer = g_bwd_layer_.as<int>();
    gg.slots = gate_prog_.slot_count;
    gg.basics = model_.basic_count;
    gg.layers = gate_prog_.layer_count;
    ga.weff = gate_weff_.as<Real>();
    ga.slot = gate_slot_.as<int>();
    ga.additive = gate_add_.as<double>();
    ga.amplitude = model_.gate_amplitude;
    ga.Mr = model_.radial_funcs_count;
    ga.Qg = static_cast<int>(model_.gate_moment_index.size());
    if (!launch_gate_graph<Real>(0, gate_tile_, N, static_cast<const Real*>(ws_.gbasics),
                                 ws_.type_s, gg, ga, static_cast<Real*>(ws_.gate_m),
                                 static_cast<Real*>(ws_.gate_d), nullptr, nullptr, s, max_lds_))
      throw UnsupportedError("SUS2_SH: gate graph tile does not fit in LDS.");
    el.gate_m = static_cast<const Real*>(ws_.gate_m);
  }
  // edge-kernel shapes: the autotuner times every candidate once per handle on the first
  // call that can be timed (one k chunk, no gate, no parameter gradient, not captured, from
  // 512 atoms: below that the launches are too short to rank); until then the size rule
  const bool edge_tune = opt_.autotune && !edge_tuned_ && !param_grad && !capturing && !noalloc &&
                         !gate && kchunks_ == 1 && N >= 512;
  if (!edge_tuned_ && !edge_tune) {
    if (opt_.basic_threads <= 0) {
      const int mfma = basic_shape_.mfma;  // the kernel kind is not the rule's to change
      basic_shape_ = default_edge_shape(N);
      basic_shape_.mfma = mfma;
    }
    if (opt_.force_threads <= 0) force_shape_ = default_edge_shape(N);
  }
  const std::vector<EdgeShape>& edge_cand = edge_candidates();
  hipEvent_t et0 = nullptr, et1 = nullptr;
  if (edge_tune) {
    hip_check(hipEventCreate(&et0), "hipEventCreate");
    hip_check(hipEventCreate(&et1), "hipEventCreate");
    // [0, C) vector basic kernel, [C, 2C) forces, [2C] the FP32 matrix-core basic kernel
    edge_tune_ms_.assign(2 * edge_cand.size() + 1, -1.0f);
  }
  el.mfma_rc = mfma_ok_ ? mfma_rc_.as<Real>() : nullptr;
  el.pair_sigma = mfma_ok_ ? pair_sigma_.as<Real>() : nullptr;
  if (basic_shape_.mfma > 0 && !mfma_ok_)
    throw UnsupportedError("SUS2_SH: the model no longer fits the matrix-core basic moments "
                           "(sus2_sh_basic_mfma): one radial scaling per species pair is needed.");
  // the basic-moment launch of the current shape (matrix cores or the vector kernel)
  const auto basics_launch = [&](const EdgeLaunch<Real>& e, const EdgeShape& sh) {
    if constexpr (sizeof(Real) == sizeof(float)) {
      if (sh.mfma > 0)
        return launch_basics_mfma(e, static_cast<float*>(ws_.basics), sh.mfma == 2, s);
    }
    return launch_basics<Real>(e, static_cast<Real*>(ws_.basics), s);
  };
  // mean time of two launches after one untimed one (-1: the shape does not launch)
  const auto time_edge = [&](const auto& launch) {
    if (!launch()) return -1.0f;
    float ms = 0.0f;
    hip_check(hipEventRecord(et0, s), "hipEventRecord");
    for (int rep = 0; rep < 2; ++rep) launch();
    hip_check(hipEventRecord(et1, s), "hipEventRecord");
    hip_check(hipEventSynchronize(et1), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(&ms, et0, et1), "hipEventElapsedTime");
    return ms / 2;
  };
  if (edge_tune && opt_.basic_mfma <= 0) {
    float best = 1e30f;
    if (opt_.basic_threads <= 0) {
      for (size_t c = 0; c < edge_cand.size(); ++c) {
        EdgeLaunch<Real> et = el;
        et.threads = edge_cand[c].threads;
        et.npl = edge_cand[c].npl;
        const float ms = time_edge([&] { return basics_launch(et, edge_cand[c]); });
        edge_tune_ms_[c] = ms;
        if (ms >= 0.0f && ms < best) {
          best = ms;
          basic_shape_ = edge_cand[c];
        }
      }
    }
    if (opt_.basic_mfma < 0 && mfma_ok_) {
      // the FP32 matrix-core kernel (same precision class as the vector kernel)
      const EdgeShape mf{256, 1, 1};
      const float ms = time_edge([&] { return basics_launch(el, mf); });
      edge_tune_ms_[2 * edge_cand.size()] = ms;
      if (ms >= 0.0f && ms < best) basic_shape_ = mf;
    }
  }
  el.threads = basic_shape_.threads;
  el.npl = basic_shape_.npl;
  bool basics_ok = basics_launch(el, basic_shape_);
  el.gate_m = nullptr;
  for (int c = 1; c < kchunks_ && basics_ok; ++c)
    basics_ok = launch_basics<Real>(chunk_launch(c), static_cast<Real*>(ws_.basics), s);
  if (!basics_ok)
    throw UnsupportedError("SUS2_SH: no basic-moment kernel for this (l_max, k_max).");
  if (profile) hip_check(hipEventRecord(events_[2], s), "hipEventRecord");

  GraphTable<Real> gt{};
  gt.fwd_rows = fwd_rows_.as<int4>();
  gt.fwd_w = fwd_w_.as<Real>();
  gt.fwd_terms = fwd_terms_.as<Term<Real>>();
  gt.fwd_layer = fwd_layer_.as<int>();
  gt.bwd_chunks = bwd_chunks_.as<int4>();
  gt.bwd_terms = bwd_terms_.as<Term<Real>>();
  gt.bwd_layer = bwd_layer_.as<int>();
  gt.e_slot = e_slot_.as<int>();
  gt.e_w = e_w_.as<Real>();
  gt.seed = seed_.as<Real>();
  gt.shift = shift_.as<double>();
  gt.sigma = sigma_.as<double>();
  gt.ones_slot = prog_.ones_slot;
  gt.n_e = static_cast<int>(prog_.energy_slot.size());
  gt.slots = prog_.slot_count;
  gt.basics = model_.basic_count;
  gt.layers = prog_.layer_count;
  ParamGradOut pgo{};
  const int S = model_.species_count, Q = model_.scalar_count, Mr = model_.radial_funcs_count;
  const size_t n_rad = static_cast<size_t>(S) * S * (Mr * model_.rb_size + 2 * Mr);
  if (param_grad) {
    pg_misc_.resize((2 * S + Q) * sizeof(double));
    pg_rad_.resize(n_rad * sizeof(double));
    launch_fill_double(pg_misc_.as<double>(), 0.0, 2 * S + Q, s);
    launch_fill_double(pg_rad_.as<double>(), 0.0, n_rad, s);
    pgo.d_shift = pg_misc_.as<double>();
    pgo.d_species = pg_misc_.as<double>() + S;
    pgo.d_moment = pg_misc_.as<double>() + 2 * S;
    pgo.row_sbegin = row_sbegin_.as<int>();
    pgo.row_slist = row_slist_.as<int>();
    pgo.stored_index = stored_index_.as<int>();
    pgo.stored_slot = stored_slot_.as<int>();
    pgo.n_stored = static_cast<int>(prog_.stored_scalar_index.size());
    pgo.Q = Q;
    pgo.S = S;
  }
  if (opt_.autotune && !tuned_ && !param_grad && !capturing && !noalloc && N >= 4096) {
    // Autotuner: time every feasible candidate -- (variant, tile) of the workgroup kernel
    // and the lane program -- and keep the fastest (HIP events on the engine's stream; the
    // graph kernels are idempotent, so the trial launches only rewrite e_s and grads).
    hipEvent_t t0, t1;
    hip_check(hipEventCreate(&t0), "hipEventCreate");
    hip_check(hipEventCreate(&t1), "hipEventCreate");
    float best = 1e30f;
    int best_tile = tile_, best_variant = variant_;
    const std::vector<LaneShape>& shapes = lane_shapes();
    const int n_lane = opt_.lane_nslot > 0 ? 1 : static_cast<int>(shapes.size());
    const std::vector<int>& bwaves = block_waves_list();
    const int n_block = opt_.block_waves > 0 ? 1 : static_cast<int>(bwaves.size());
    tune_ms_.assign(15 + shapes.size() + bwaves.size(), -1.0f);
    LaneProgram best_lane = lane_;
    BlockProgram best_block = block_;
    auto time_it = [&](int v, int t) {
      float ms = 0.0f;
      if (!launch_graph_variant<Real>(v, t, N, gt, s)) return -1.0f;
      hip_check(hipEventRecord(t0, s), "hipEventRecord");
      for (int rep = 0; rep < 2; ++rep) launch_graph_variant<Real>(v, t, N, gt, s);
      hip_check(hipEventRecord(t1, s), "hipEventRecord");
      hip_check(hipEventSynchronize(t1), "hipEventSynchronize");
      hip_check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
      return ms / 2;
    };
    for (int v = 0; v < 5; ++v) {
      if (opt_.graph_variant >= 0 && v != opt_.graph_variant) continue;  // fixed variant
      for (int li = 0; li < (v == 4 ? n_block : v == 3 ? n_lane : 5); ++li) {
        const int t = 1 << li;
        if (v < 3 && graph_lds_bytes<Real>(t, prog_.slot_count) > max_lds_) continue;
        if (v == 4) {
          if (opt_.block_waves <= 0) {
            block_ = build_block(bwaves[li]);
            if (!block_.ok) continue;
            upload_block();
          } else if (!block_.ok) {
            continue;
          }
        }
        if (v == 3) {
          if (opt_.lane_nslot <= 0) {
            lane_ = build_lane(shapes[li]);
            if (!lane_.ok) continue;
            upload_lane();
          } else if (!lane_.ok) {
            continue;
          }
        }
        const float ms = time_it(v, t);
        if (ms < 0.0f) continue;
        tune_ms_[v == 4 ? 15 + static_cast<int>(shapes.size()) + li : v == 3 ? 15 + li : v * 5 + li] = ms;
        if (ms < best) {
          best = ms;
          best_tile = t;
          best_variant = v;
          if (v == 3) best_lane = lane_;
          if (v == 4) best_block = block_;
        }
      }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (best_variant < 3) tile_ = best_tile;
    variant_ = best_variant;
    if (best_block.ok && (best_block.waves != block_.waves || !block_.ok)) {
      block_ = std::move(best_block);
      upload_block();
    }
    if (best_lane.ok && (best_lane.waves != lane_.waves || best_lane.nslot != lane_.nslot ||
                         best_lane.slots != lane_.slots || best_lane.atoms != lane_.atoms)) {
      lane_ = std::move(best_lane);
      upload_lane();
    }
    tuned_ = true;
  }
  if (param_grad) {
    // the parameter gradient needs per-scalar sums (dE/dw_s = sum_i sigma_i s_i): the
    // persistent row-mode kernel; with a block or lane program that kernel writes the
    // energies and adjoints and the row kernel only runs the forward pass for the sums
    const int ea = block_.ok ? 4 : (lane_.ok ? 3 : -1);
    pgo.fwd_only = ea >= 0 ? 1 : 0;
    if (ea >= 0 && !launch_graph_variant<Real>(ea, tile_, N, gt, s))
      throw UnsupportedError("SUS2_SH: graph kernel variant does not fit in LDS.");
    if (!launch_graph<Real>(tile_, 0, N, static_cast<const Real*>(ws_.basics), ws_.type_s, gt,
                            ws_.e_s, static_cast<Real*>(ws_.grads), &pgo, s, max_lds_,
                            4 * cu_count_))
      throw UnsupportedError("SUS2_SH: graph tile does not fit in LDS.");
  } else if (!launch_graph_variant<Real>(variant_, tile_, N, gt, s)) {
    throw UnsupportedError("SUS2_SH: graph kernel variant does not fit in LDS.");
  }
  if (profile) hip_check(hipEventRecord(events_[3], s), "hipEventRecord");

  // neighbour windows of the force tiles: LDS accumulation of the
  // neighbour forces, budget 48 KiB of 8-byte accumulators (double, or 64-bit fixed point in
  // deterministic mode) per 128-atom tile
  const int* win = nullptr;
  int win_cap = 0;
  if (use_win) {
    win_cap = static_cast<int>(48 * 1024 / (3 * sizeof(double)));
    if (batch)
      launch_windows_batch(N, sid_s_.as<int>(), batch_atom_off_.as<int>(), ws_.win, win_cap, s);
    else
      launch_windows(N, grid, ws_.skey, ws_.win, win_cap, s);
    win = ws_.win;
  }
  if (edge_tune && opt_.force_threads <= 0) {
    // trial launches accumulate into f_s and overwrite v_s: both are reset below
    float best = 1e30f;
    for (size_t c = 0; c < edge_cand.size(); ++c) {
      EdgeLaunch<Real> et = el;
      et.threads = edge_cand[c].threads;
      et.npl = edge_cand[c].npl;
      const float ms = time_edge([&] {
        return launch_forces<Real>(et, static_cast<const Real*>(ws_.grads), ws_.f_s, ws_.v_s,
                                   nullptr, win, win_cap, s);
      });
      edge_tune_ms_[edge_cand.size() + c] = ms;
      if (ms >= 0.0f && ms < best) {
        best = ms;
        force_shape_ = edge_cand[c];
      }
    }
  }
  if (edge_tune) {
    (void)hipEventDestroy(et0);
    (void)hipEventDestroy(et1);
    edge_tuned_ = true;
  }
  el.threads = force_shape_.threads;
  el.npl = force_shape_.npl;
  if (param_grad && opt_.force_threads <= 0) {
    // the parameter-gradient kernel zeroes, fills and flushes per-pair accumulators in every
    // workgroup: it keeps 256-thread workgroups (the timed shapes do not cover it)
    el.threads = 256;
    el.npl = default_edge_shape(N).npl;
  }
  launch_fill_double(ws_.f_s, 0.0, 3 * static_cast<size_t>(N), s);
  if (gate) {
    // gated main forces and dE/dh scattered to the neighbours, the reverse through the gate
    // layer's graph, then the gate moments' own forces (src/force/sus2_sh.cu:10052-10200)
    const size_t nm = static_cast<size_t>(N) * model_.radial_funcs_count;
    launch_fill_int(static_cast<int*>(ws_.gate_adj), 0, nm * sizeof(Real) / sizeof(int), s);
    EdgeLaunch<Real> ef = el;
    ef.gate_m = static_cast<const Real*>(ws_.gate_m);
    ef.gate_d = static_cast<const Real*>(ws_.gate_d);
    ef.gate_adj = static_cast<Real*>(ws_.gate_adj);
    if (!launch_forces<Real>(ef, static_cast<const Real*>(ws_.grads), ws_.f_s, ws_.v_s, nullptr,
                             nullptr, 0, s))
      throw UnsupportedError("SUS2_SH: no force kernel for this (l_max, k_max).");
    if (!launch_gate_graph<Real>(1, gate_tile_, N, static_cast<const Real*>(ws_.gbasics),
                                 ws_.type_s, gg, ga, nullptr, nullptr,
                                 static_cast<const Real*>(ws_.gate_adj),
                                 static_cast<Real*>(ws_.ggrads), s, max_lds_))
      throw UnsupportedError("SUS2_SH: gate graph tile does not fit in LDS.");
    EdgeLaunch<Real> e2 = el;
    e2.rt.coeff = gate_coeff_.as<Real>();
    e2.rt.shared = gate_shared_.as<Real>();
    e2.v_accumulate = 1;
    if (!launch_forces<Real>(e2, static_cast<const Real*>(ws_.ggrads), ws_.f_s, ws_.v_s, nullptr,
                             win, win_cap, s))
      throw UnsupportedError("SUS2_SH: no force kernel for this (l_max, k_max).");
  } else {
    // k chunks after the first add their virials to the first's
    for (int c = 0; c < kchunks_; ++c) {
      EdgeLaunch<Real> ec = chunk_launch(c);
      ec.v_accumulate = c > 0 ? 1 : 0;
      if (param_grad) {
        // copies of the per-pair dE/dtheta accumulators in LDS (lanes spread over them):
        // as many as fit 72 KiB, at most 16
        const size_t mrc = static_cast<size_t>(kchunk_K_[c]) * (model_.l_max + 1);
        const size_t per_copy = sizeof(double) * model_.species_count * model_.species_count *
                                (mrc * model_.rb_size + 2 * mrc);
        ec.pg_copies = 1;
        while (ec.pg_copies < 16 && 2 * ec.pg_copies * per_copy <= 72 * 1024) ec.pg_copies *= 2;
      }
      if (!launch_forces<Real>(ec, static_cast<const Real*>(ws_.grads), ws_.f_s, ws_.v_s,
                               param_grad ? pg_rad_.as<double>() + pg_chunk_offset(c) : nullptr,
                               win, win_cap, s))
        throw UnsupportedError("SUS2_SH: no force kernel for this (l_max, k_max).");
    }
  }
  if (model_.zbl_enabled) {
    const ZblTable zt{zbl_z_.as<int>(), zbl_inner_.as<double>(), zbl_outer_.as<double>(),
                      model_.species_count};
    launch_zbl<Real>(N, ws_.count, static_cast<const Edge<Real>*>(ws_.nl), ws_.type_s, zt,
                     ws_.e_s, ws_.f_s, ws_.v_s, el.fix_flag, s);
  }
  if (profile) hip_check(hipEventRecord(events_[4], s), "hipEventRecord");

  launch_scatter(N, ws_.rank, ws_.e_s, ws_.f_s, ws_.v_s, pot, force, virial,
                 opt_.deterministic ? 1 : 0, s);
  hip_check(hipGetLastError(), "kernel launch");
  if (opt_.deterministic && !deferred) {
    // out-of-range contributions of the fixed-point sums (a deferred call reports them at
    // check())
    hip_check(hipMemcpyAsync(host_max_ + kLatchFix, latch_.as<int>() + kLatchFix, sizeof(int),
                             hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (host_max_[kLatchFix] != 0) {
      launch_fill_int(latch_.as<int>() + kLatchFix, 0, 1, s);
      throw ArgumentError(fix_range_message());
    }
  }
  if (param_grad) finish_param_grad(param_grad, s);
  if (profile) {
    hip_check(hipEventRecord(events_[5], s), "hipEventRecord");
    hip_check(hipEventSynchronize(events_[5]), "hipEventSynchronize");
    stage_ms_.assign(5, 0.0f);
    for (int k = 0; k < 5; ++k)
      hip_check(hipEventElapsedTime(&stage_ms_[k], events_[k], events_[k + 1]),
                "hipEventElapsedTime");
  }
}

void Engine::get_params(double* theta, int n) const
{
  if (!theta || n != param_count()) throw ArgumentError("get_params: wrong parameter count");
  const Model& m = model_;
  double* o = theta;
  o = std::copy(m.shift_coeffs.begin(), m.shift_coeffs.end(), o);
  o = std::copy(m.species_coeffs.begin(), m.species_coeffs.end(), o);
  o = std::copy(m.moment_coeffs.begin(), m.moment_coeffs.end(), o);
  o = std::copy(m.radial_coeffs.begin(), m.radial_coeffs.end(), o);
  o = std::copy(m.radial_type_coeffs.begin(), m.radial_type_coeffs.end(), o);
  std::copy(m.scal_coeffs.begin(), m.scal_coeffs.end(), o);
}

void Engine::set_params(const double* theta, int n)
{
  if (!theta || n != param_count()) throw ArgumentError("set_params: wrong parameter count");
  Model& m = model_;
  const double* p = theta;
  auto take = [&](std::vector<double>& v) {
    std::copy(p, p + v.size(), v.begin());
    p += v.size();
  };
  take(m.shift_coeffs);
  take(m.species_coeffs);
  take(m.moment_coeffs);
  take(m.radial_coeffs);
  take(m.radial_type_coeffs);
  take(m.scal_coeffs);
  prog_ = compile_graph(m, graph_chunk_terms());
  compile_lane();
  // the block program folds the energy weights into its records: rebuilt with its current shape
  block_ = build_block(block_.ok ? block_.waves
                                 : (opt_.block_waves > 0 ? opt_.block_waves : block_waves_list()[0]));
  if (opt_.use_float)
    upload_tables<float>();
  else
    upload_tables<double>();
}

size_t Engine::pg_chunk_offset(int c) const
{
  const size_t S = model_.species_count;
  size_t off = 0;
  for (int k = 0; k < c; ++k) {
    const size_t mrc = static_cast<size_t>(kchunk_K_[k]) * (model_.l_max + 1);
    off += S * S * (mrc * model_.rb_size + 2 * mrc);
  }
  return off;
}

std::string Engine::fix_range_message()
{
  std::ostringstream msg;
  msg << "SUS2_SH: sus2_sh_deterministic=1: a force contribution is outside the fixed-point "
         "range (pair terms below "
      << kFixPairMax << " eV/A, per-atom sums below " << kFixSumMax << " eV/A, at most "
      << kFixMaxNeighbors << " neighbours per atom)";
  return msg.str();
}

void Engine::reset_latch(hipStream_t s) { launch_fill_int(latch_.as<int>(), 0, 8, s); }

void Engine::read_latch(hipStream_t s, int* out)
{
  hip_check(hipMemcpyAsync(host_max_, latch_.as<int>(), 8 * sizeof(int), hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync D2H");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  std::memcpy(out, host_max_, 8 * sizeof(int));
}

void Engine::check(hipStream_t s)
{
  int l[8];
  read_latch(s, l);
  reset_latch(s);
  if (l[4] == 0) return;
  if (opt_.deterministic && (l[kLatchFix] != 0 || l[0] > kFixMaxNeighbors))
    throw ArgumentError(fix_range_message());
  const unsigned long long total = static_cast<unsigned long long>(static_cast<unsigned>(l[2])) |
                                   (static_cast<unsigned long long>(static_cast<unsigned>(l[3])) << 32);
  if (cap_atoms_ > 0) mean_neighbors_ = static_cast<double>(total) / cap_atoms_;
  if (l[1] != 0) throw ArgumentError("atom type outside [0, species_count) passed to compute");
  if (l[0] > cap_) {
    std::ostringstream msg;
    msg << "SUS2_SH: neighbour capacity exceeded in a deferred-check evaluation (" << l[0]
        << " neighbours, capacity " << cap_ << "): the outputs of the " << l[4]
        << " evaluation(s) since the previous check miss neighbours; the capacity is raised "
           "for the next (synchronised) call";
    cap_ = ((l[0] + 8 + 7) / 8) * 8;
    cap_known_ = false;
    throw CapacityError(msg.str());
  }
}

void Engine::md_nve(const sus2sh_box& box, int N, const int* type, const double* masses,
                    double* pos, double* vel, double dt_fs, int nsteps, double* pot, double* force,
                    double* virial, double* energy_trace, hipStream_t s)
{
  if (N <= 0 || nsteps < 0) throw ArgumentError("md_nve: N must be > 0 and nsteps >= 0");
  if (!masses || !pos || !vel) throw ArgumentError("md_nve: null pointer");
  const int S = model_.species_count;
  for (int t = 0; t < S; ++t)
    if (!(masses[t] > 0.0)) throw ArgumentError("md_nve: masses must be positive");
  // graph mode runs on the engine's own stream (the legacy default stream cannot be
  // captured), ordered after the caller's stream and back
  const hipStream_t caller = s;
  if (opt_.md_graph) {
    if (!md_stream_) hip_check(hipStreamCreateWithFlags(&md_stream_, hipStreamNonBlocking),
                               "hipStreamCreateWithFlags");
    if (!md_event_) hip_check(hipEventCreateWithFlags(&md_event_, hipEventDisableTiming),
                              "hipEventCreate");
    hip_check(hipEventRecord(md_event_, caller), "hipEventRecord");
    hip_check(hipStreamWaitEvent(md_stream_, md_event_, 0), "hipStreamWaitEvent");
    s = md_stream_;
  }
  md_mass_.upload(masses, S * sizeof(double));
  md_energy_.resize(2 * static_cast<size_t>(nsteps + 1) * sizeof(double));
  md_step_.resize(sizeof(int));
  md_partial_.resize(md_partials_size(N) * sizeof(double));
  double* en = md_energy_.as<double>();
  const double* mass = md_mass_.as<double>();
  int* step = md_step_.as<int>();
  launch_fill_double(en, 0.0, 2 * static_cast<size_t>(nsteps + 1), s);
  const size_t n = static_cast<size_t>(N);
  auto evaluate = [&]() {
    launch_fill_double(pot, 0.0, n, s);
    launch_fill_double(force, 0.0, 3 * n, s);
    launch_fill_double(virial, 0.0, 9 * n, s);
    compute(box, N, type, pos, pot, force, virial, s, nullptr);
  };
  // one velocity-Verlet step; the energies go to en[2 * step] and the counter advances
  auto md_step = [&]() {
    launch_half_kick(N, 0.5 * dt_fs, type, mass, force, vel, s);
    launch_drift(N, dt_fs, vel, pos, s);
    evaluate();
    launch_half_kick(N, 0.5 * dt_fs, type, mass, force, vel, s);
    launch_energies(N, type, mass, pot, vel, en, step, md_partial_.as<double>(), s);
    launch_step_advance(step, s);
  };
  evaluate();  // synchronised: sizes the neighbour capacity
  launch_fill_int(step, 0, 1, s);
  launch_energies(N, type, mass, pot, vel, en, step, md_partial_.as<double>(), s);
  launch_step_advance(step, s);
  if (!opt_.md_graph) {
    for (int k = 1; k <= nsteps; ++k) md_step();
  } else {
    // Steady steps replay one captured step (no host synchronisation inside).  Chunks of
    // kChunk steps start from a snapshot of positions and velocities; a chunk whose
    // neighbour counts overflowed the capacity (latched on the device) is rolled back and
    // run again with a larger capacity.
    constexpr int kChunk = 50;
    md_x0_.resize(3 * n * sizeof(double));
    md_v0_.resize(3 * n * sizeof(double));
    hipGraphExec_t exec = nullptr;
    auto drop = [&]() {
      if (exec) (void)hipGraphExecDestroy(exec);
      exec = nullptr;
    };
    try {
      int done = 0, retries = 0;
      while (done < nsteps) {
        const int chunk = std::min(kChunk, nsteps - done);
        hip_check(hipMemcpyAsync(md_x0_.as<double>(), pos, 3 * n * sizeof(double),
                                 hipMemcpyDeviceToDevice, s), "hipMemcpyAsync D2D");
        hip_check(hipMemcpyAsync(md_v0_.as<double>(), vel, 3 * n * sizeof(double),
                                 hipMemcpyDeviceToDevice, s), "hipMemcpyAsync D2D");
        reset_latch(s);
        if (!exec) {
          hipGraph_t graph = nullptr;
          hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal),
                    "hipStreamBeginCapture");
          md_capture_ = true;
          try {
            md_step();
          } catch (...) {
            md_capture_ = false;
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
          }
          md_capture_ = false;
          hip_check(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
          const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          hip_check(e, "hipGraphInstantiate");
        }
        for (int k = 0; k < chunk; ++k) hip_check(hipGraphLaunch(exec, s), "hipGraphLaunch");
        int l[8];
        read_latch(s, l);
        if (l[1] != 0) throw ArgumentError("atom type outside [0, species_count) passed to compute");
        if (opt_.deterministic && (l[kLatchFix] != 0 || l[0] > kFixMaxNeighbors))
          throw ArgumentError(fix_range_message());
        if (l[0] > cap_) {
          if (++retries > 8) throw DeviceError("md_nve: neighbour capacity did not settle");
          // roll back to the chunk start and size the capacity with a synchronised call
          cap_ = ((l[0] + 8 + 7) / 8) * 8;
          cap_known_ = false;
          drop();
          hip_check(hipMemcpyAsync(pos, md_x0_.as<double>(), 3 * n * sizeof(double),
                                   hipMemcpyDeviceToDevice, s), "hipMemcpyAsync D2D");
          hip_check(hipMemcpyAsync(vel, md_v0_.as<double>(), 3 * n * sizeof(double),
                                   hipMemcpyDeviceToDevice, s), "hipMemcpyAsync D2D");
          launch_fill_int(step, done + 1, 1, s);
          launch_fill_double(en + 2 * static_cast<size_t>(done + 1), 0.0, 2 * static_cast<size_t>(chunk), s);
          evaluate();
          continue;
        }
        const unsigned long long total =
          static_cast<unsigned long long>(static_cast<unsigned>(l[2])) |
          (static_cast<unsigned long long>(static_cast<unsigned>(l[3])) << 32);
        mean_neighbors_ = static_cast<double>(total) / N;
        done += chunk;
      }
    } catch (...) {
      md_capture_ = false;
      drop();
      throw;
    }
    drop();
  }
  hip_check(hipGetLastError(), "md kernel launch");
  if (energy_trace) {
    hip_check(hipMemcpyAsync(energy_trace, en, 2 * static_cast<size_t>(nsteps + 1) * sizeof(double),
                             hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  }
  if (s != caller) {
    hip_check(hipEventRecord(md_event_, s), "hipEventRecord");
    hip_check(hipStreamWaitEvent(caller, md_event_, 0), "hipStreamWaitEvent");
  }
}

// dE/dtheta from the per-type / per-pair adjoint sums (host, double):
//   shift, species, moment_coeffs directly; radial[mu][n] and type coefficients through
//   c^p_{mu n} = scaling w_ti w_tj radial[mu][n]; scal_coeffs per (mu, pair).
void Engine::finish_param_grad(double* d_param_grad, hipStream_t s)
{
  const Model& m = model_;
  const int S = m.species_count, Q = m.scalar_count, Mr = m.radial_funcs_count, rb = m.rb_size;
  std::vector<double> misc(2 * S + Q), rad(static_cast<size_t>(S) * S * (Mr * rb + 2 * Mr));
  const int P = param_count();
  std::vector<double> grad(P);
  hip_check(hipMemcpyAsync(misc.data(), pg_misc_.as<double>(), misc.size() * sizeof(double),
                           hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(rad.data(), pg_rad_.as<double>(), rad.size() * sizeof(double),
                           hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(grad.data(), d_param_grad, P * sizeof(double), hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  const int o_shift = 0, o_species = S, o_w = 2 * S, o_rad = 2 * S + Q, o_type = o_rad + Mr * rb,
            o_scal = o_type + S;
  for (int t = 0; t < S; ++t) {
    grad[o_shift + t] += misc[t];
    grad[o_species + t] += misc[S + t];
  }
  for (int q = 0; q < Q; ++q) grad[o_w + q] += misc[2 * S + q];
  // per k chunk: [S*S][mrc*rb + 2*mrc] for its radial rows mu0 .. mu0 + mrc - 1
  for (int c = 0; c < kchunks_; ++c) {
    const int mu0 = kchunk_k0_[c] * (m.l_max + 1), mrc = kchunk_K_[c] * (m.l_max + 1);
    const int stride = mrc * rb + 2 * mrc;
    for (int ti = 0; ti < S; ++ti)
      for (int tj = 0; tj < S; ++tj) {
        const double* D = rad.data() + pg_chunk_offset(c) + static_cast<size_t>(ti * S + tj) * stride;
        const double wi = m.radial_type_coeffs[ti], wj = m.radial_type_coeffs[tj];
        for (int ml = 0; ml < mrc; ++ml) {
          const int mu = mu0 + ml;
          for (int n = 0; n < rb; ++n) {
            const double d = D[ml * rb + n];
            grad[o_rad + mu * rb + n] += d * m.scaling * wi * wj;
            const double dv = d * m.scaling * m.radial_coeffs[mu * rb + n];
            grad[o_type + ti] += dv * wj;
            grad[o_type + tj] += dv * wi;
          }
          const size_t so = static_cast<size_t>(2) * mu * S * S + static_cast<size_t>(S) * ti + tj;
          grad[o_scal + so] += D[mrc * rb + ml];
          grad[o_scal + so + static_cast<size_t>(S) * S] += D[mrc * rb + mrc + ml];
        }
      }
  }
  hip_check(hipMemcpyAsync(d_param_grad, grad.data(), P * sizeof(double), hipMemcpyHostToDevice, s),
            "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

template void Engine::upload_tables<float>();
template void Engine::upload_tables<double>();
template void Engine::upload_gate_tables<float>();
template void Engine::upload_gate_tables<double>();

}  // namespace sus2sh
