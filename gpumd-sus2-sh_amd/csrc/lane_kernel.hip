// Lane graph kernel: the depth <= 2 product graph of one atom per lane, driven by a
// wave-uniform program (lane_program.hpp) -- forward values, site energy and the reverse
// pass down to the basic-moment adjoints (semantics: src/force/sus2_sh.cu:4799-4846).
//
// A workgroup owns a tile of 64 atoms.  Each lane owns a column of the tile's LDS: slot s of
// lane i sits at word s * 64 + i, so a uniform slot index reads 64 consecutive words (no
// bank conflicts).  The WPT waves of the workgroup serve the same 64 atoms: the program
// splits every segment's independent rows between them, so a wave runs one WPT-th of the
// instructions and the CU holds WPT times more waves for the same LDS.  A barrier precedes
// every segment that reads what earlier segments wrote (the segment table flags the others:
// a back segment follows its energy segment without one).  A group of NS instructions is one
// pre-decoded record (LDS word offsets, row-end mask, coefficients) that is
// uniform across the wave and fetched with scalar loads, so a group costs no scalar decoding;
// its NS accumulators are independent, so the group issues its NS (or 2 NS) LDS reads before
// the first FMA that needs them, and the reads of the next groups are in flight meanwhile.
// Segment kinds: 0 binary rows stored, 1 unary rows stored, 2 binary rows summed into the
// site energy, 3 binary rows added to their targets.
#include <algorithm>

#include "kernels.hpp"
#include "lane_program.hpp"

namespace sus2sh {

namespace {

constexpr int kLaneWave = 64;

// Read-only program data is read through the constant address space: such loads stay scalar
// across the segment barriers (a barrier is a possible clobber of global memory).
using cu32 = const __attribute__((address_space(4))) uint32_t;

// One group record in scalar registers.
template <typename Real, int NS, bool BINARY>
struct LaneRec {
  uint32_t xo[NS], yo[NS], to[NS], mask;
  Real c[NS];
};

template <typename Real, int NS, bool BINARY>
__device__ __forceinline__ void fetch_rec(cu32* r, LaneRec<Real, NS, BINARY>& q)
{
  constexpr int HW = lane_header_words(NS, static_cast<int>(sizeof(Real)));
  using creal = const __attribute__((address_space(4))) Real;
  creal* f = reinterpret_cast<creal*>(r + HW);
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    q.xo[j] = r[j];
    if (BINARY) q.yo[j] = r[NS + j];
    q.to[j] = r[2 * NS + j];
    q.c[j] = f[j];
  }
  q.mask = r[3 * NS];
}

// Make every field of a fetched record live in scalar registers at this point (an empty asm
// that takes them as SGPR operands), so that a batch of records costs one wait at its start
// instead of loads the compiler would sink into the row-end branches, each followed by a
// wait that also drains the LDS reads in flight.
template <typename Real, int NS, bool BINARY>
__device__ __forceinline__ void pin_rec(const LaneRec<Real, NS, BINARY>& q)
{
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    asm volatile("" : : "s"(q.xo[j]), "s"(q.to[j]), "s"(q.c[j]));
    if (BINARY) asm volatile("" : : "s"(q.yo[j]));
  }
  asm volatile("" : : "s"(q.mask));
}

// Issue the LDS reads of a group: for each of the lane's NA atoms (columns 64 words apart,
// so the second atom's read is the first one's address plus an immediate offset).
template <typename Real, int NS, bool BINARY, int NA>
__device__ __forceinline__ void read_rec(const LaneRec<Real, NS, BINARY>& q, const Real* X,
                                         Real (&xv)[NA][NS], Real (&yv)[NA][NS])
{
#pragma unroll
  for (int j = 0; j < NS; ++j) {
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      xv[a][j] = X[q.xo[j] + a * kLaneWave];
      if (BINARY) yv[a][j] = X[q.yo[j] + a * kLaneWave];
    }
  }
}

// Accumulate a group and end its rows: a row ends on a uniform bit of the record's mask
// (scalar test and branch; most groups end no row), where the accumulator is stored (kinds
// 0, 1) or added (kind 3) to the row's target and restarted.  Mask bit NS ends a row split
// between accumulators 0 and 1: accumulator 1 is added to accumulator 0 first.  Energy rows
// (kind 2) are summed once at the end of the segment.
template <typename Real, int NS, int KIND, int NA>
__device__ __forceinline__ void finish_rec(const LaneRec<Real, NS, KIND != 1>& q,
                                           const Real (&xv)[NA][NS], const Real (&yv)[NA][NS],
                                           Real (&acc)[NA][NS], Real* X)
{
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int j = 0; j < NS; ++j)
      acc[a][j] += KIND != 1 ? (q.c[j] * xv[a][j]) * yv[a][j] : q.c[j] * xv[a][j];
  if (KIND != 2 && q.mask != 0u) {
    if constexpr (NS >= 2) {
      if (q.mask & (1u << NS)) {  // a split row ends: its odd half joins accumulator 0
#pragma unroll
        for (int a = 0; a < NA; ++a) {
          acc[a][0] += acc[a][1];
          acc[a][1] = Real(0);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      if (q.mask & (1u << j)) {
#pragma unroll
        for (int a = 0; a < NA; ++a) {
          if (KIND == 3)
            X[q.to[j] + a * kLaneWave] += acc[a][j];
          else
            X[q.to[j] + a * kLaneWave] = acc[a][j];
          acc[a][j] = Real(0);
        }
      }
    }
  }
}

// Records per batch.  A batch is one set of scalar loads at immediate offsets from one
// pointer and one wait (the program does not fit the scalar cache: each batch is a round
// trip to L2), so the scalar unit spends a pointer increment and a loop test per batch.
template <int NS, typename Real>
constexpr int lane_batch()
{
  // as many records as fit about 64 (FP32) / 48 (FP64) scalar registers, at most 6 / 4
  constexpr int rw = lane_record_words(NS, static_cast<int>(sizeof(Real)));
  constexpr int budget = sizeof(Real) == 4 ? 64 : 48, cap = sizeof(Real) == 4 ? 6 : 4;
  return budget / rw < 1 ? 1 : (budget / rw > cap ? cap : budget / rw);
}

// Wait until at most `ahead` LDS operations are outstanding (the reads of the groups after
// the current one), so that a group's operands cost one wait instead of one per consumer.
// LDS operations of a wave complete in order, so a group's reads, followed by `ahead` later
// ones, are then complete.  The compiler still adds any wait it needs; this one only makes
// its per-use waits redundant.
// s_waitcnt immediate: vmcnt and expcnt at their maxima (no wait), lgkmcnt = ahead.
template <int ahead>
__device__ __forceinline__ void wait_lds_but()
{
  if constexpr (ahead <= 15) {
    __builtin_amdgcn_sched_barrier(0);  // the wait stays between the reads and their uses
    __builtin_amdgcn_s_waitcnt(0xC07F | (ahead << 8));
    __builtin_amdgcn_sched_barrier(0);
  }
}

// CNT records fetched together (one batch of scalar loads, one wait), then run with the
// operands of DEPTH + 1 groups in flight: group k + DEPTH is read before group k is finished
// (a segment never reads what it writes, so reads may pass the row-end stores).
template <typename Real, int NS, int KIND, int CNT, int NA>
__device__ __forceinline__ void run_batch(cu32* R, Real* X, Real (&acc)[NA][NS])
{
  constexpr bool BIN = KIND != 1;
  constexpr int RW = lane_record_words(NS, static_cast<int>(sizeof(Real)));
  constexpr int kReads = (BIN ? 2 * NS : NS) * NA;  // LDS reads of one group
  // up to three groups ahead, as far as the 4-bit LDS counter can express the wait
  constexpr int DEPTH = (CNT > 3 && 3 * kReads <= 15)   ? 3
                        : (CNT > 2 && 2 * kReads <= 15) ? 2
                        : (CNT > 1 ? 1 : 0);
  Real xv[DEPTH + 1][NA][NS], yv[DEPTH + 1][NA][NS];
  LaneRec<Real, NS, BIN> r[CNT];
#pragma unroll
  for (int k = 0; k < CNT; ++k) fetch_rec<Real, NS, BIN>(R + k * RW, r[k]);
#pragma unroll
  for (int k = 0; k < CNT; ++k) pin_rec<Real, NS, BIN>(r[k]);
#pragma unroll
  for (int k = 0; k < DEPTH; ++k) read_rec<Real, NS, BIN, NA>(r[k], X, xv[k], yv[k]);
#pragma unroll
  for (int k = 0; k < CNT; ++k) {
    if (k + DEPTH < CNT) {
      constexpr int d = DEPTH;
      read_rec<Real, NS, BIN, NA>(r[k + d < CNT ? k + d : 0], X, xv[(k + d) % (d + 1)],
                                  yv[(k + d) % (d + 1)]);
    }
    // groups still in flight after group k: min(DEPTH, CNT - 1 - k)
    const int inflight = CNT - 1 - k < DEPTH ? CNT - 1 - k : DEPTH;
    if (inflight == 3) wait_lds_but<3 * kReads>();
    else if (inflight == 2) wait_lds_but<2 * kReads>();
    else if (inflight == 1) wait_lds_but<kReads>();
    else wait_lds_but<0>();
    finish_rec<Real, NS, KIND, NA>(r[k], xv[k % (DEPTH + 1)], yv[k % (DEPTH + 1)], acc, X);
  }
}

// The last `rest` (< C + 1) records of a segment as one batch.
template <typename Real, int NS, int KIND, int C, int NA>
__device__ __forceinline__ void run_tail(cu32* R, int rest, Real* X, Real (&acc)[NA][NS])
{
  if constexpr (C > 0) {
    if (rest == C)
      run_batch<Real, NS, KIND, C, NA>(R, X, acc);
    else
      run_tail<Real, NS, KIND, C - 1, NA>(R, rest, X, acc);
  }
}

// One segment of one wave: `groups` records from R, SG at a time; the remainder is one
// shorter batch, so a segment costs ceil(groups / SG) scalar-load round trips.
template <typename Real, int NS, int KIND, int NA>
__device__ __forceinline__ void run_segment(cu32* R, int groups, Real* X, Real (&e)[NA])
{
  constexpr int RW = lane_record_words(NS, static_cast<int>(sizeof(Real)));
  constexpr int SG = lane_batch<NS, Real>();
  Real acc[NA][NS];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int j = 0; j < NS; ++j) acc[a][j] = Real(0);
  int g = 0;
  for (; g + SG <= groups; g += SG) {
    run_batch<Real, NS, KIND, SG, NA>(R, X, acc);
    R += SG * RW;
  }
  run_tail<Real, NS, KIND, SG - 1, NA>(R, groups - g, X, acc);
  if (KIND == 2) {
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int j = 0; j < NS; ++j) e[a] += acc[a][j];
  }
}

// NA atoms per lane: a workgroup owns 64 NA atoms; slot s of atom a of lane i sits at word
// s * 64 NA + 64 a + i, so one address serves the lane's NA atoms (immediate offsets of 256 a
// bytes) and one record fetch serves 64 NA atoms.
template <typename Real, int NS, int WPT, int NA>
__global__ __launch_bounds__(kLaneWave * WPT) void k_lane_graph(int N, const Real* __restrict__ basics,
                                                               const int* __restrict__ type_s,
                                                               LaneTable<Real> lt,
                                                               double* __restrict__ e_out,
                                                               Real* __restrict__ grad_out)
{
  constexpr int RW = lane_record_words(NS, static_cast<int>(sizeof(Real)));
  constexpr int COL = kLaneWave * NA;  // words per slot row
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const unsigned lane = threadIdx.x & (kLaneWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kLaneWave));
  Real* X = reinterpret_cast<Real*>(smem_raw) + lane;  // slot s, atom a: X[s * COL + 64 a]
  const unsigned atom0 = blockIdx.x * static_cast<unsigned>(COL) + lane;
  const int B = lt.basics;

  for (int q = wave; q < B; q += WPT) {
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      const unsigned atom = atom0 + a * kLaneWave;
      X[q * COL + a * kLaneWave] =
        atom < static_cast<unsigned>(N) ? basics[static_cast<size_t>(q) * N + atom] : Real(0);
      X[(B + q) * COL + a * kLaneWave] = Real(0);
    }
  }
  if (wave == 0) {
#pragma unroll
    for (int a = 0; a < NA; ++a) X[lt.ones_slot * COL + a * kLaneWave] = Real(1);
  }

  Real e[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) e[a] = Real(0);
  for (int s = 0; s < lt.nseg; ++s) {
    cu32* sg = (cu32*)lt.segs + 4 * s;  // (kind, first record, groups, barrier before)
    const uint32_t kind = sg[0], first = sg[1], groups = sg[2], sync = sg[3];
    if (sync) __syncthreads();  // every wave runs every segment: uniform barrier
    cu32* R = (cu32*)lt.recs + (static_cast<size_t>(first) + static_cast<size_t>(wave) * groups) * RW;
    switch (kind) {
      case 0: run_segment<Real, NS, 0, NA>(R, static_cast<int>(groups), X, e); break;
      case 1: run_segment<Real, NS, 1, NA>(R, static_cast<int>(groups), X, e); break;
      case 2: run_segment<Real, NS, 2, NA>(R, static_cast<int>(groups), X, e); break;
      default: run_segment<Real, NS, 3, NA>(R, static_cast<int>(groups), X, e); break;
    }
  }
  __syncthreads();

  // per-wave energies -> the (now unused) component-value region, summed by wave 0
  if (WPT > 1) {
#pragma unroll
    for (int a = 0; a < NA; ++a) X[(lt.v_base + wave) * COL + a * kLaneWave] = e[a];
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const unsigned atom = atom0 + a * kLaneWave;
    if (atom >= static_cast<unsigned>(N)) continue;
    const int t = type_s[atom];
    if (wave == 0) {
      Real sum = e[a];
#pragma unroll
      for (int w = 1; w < WPT; ++w) sum += X[(lt.v_base + w) * COL + a * kLaneWave];
      const double sg = lt.sigma[t];
      e_out[atom] = lt.shift[t] + sg + sg * static_cast<double>(sum);
    }
    const Real sgr = static_cast<Real>(lt.sigma[t]);
    for (int q = wave; q < B; q += WPT)
      grad_out[static_cast<size_t>(q) * N + atom] = sgr * X[(B + q) * COL + a * kLaneWave];
  }
}

template <typename Real, int NS, int WPT, int NA>
void launch_lane_shape(int N, const Real* basics, const int* type_s, const LaneTable<Real>& lt,
                       double* e_s, Real* grads, hipStream_t s, size_t lds)
{
  auto kern = &k_lane_graph<Real, NS, WPT, NA>;
  if (lds > 65536) allow_large_lds<&k_lane_graph<Real, NS, WPT, NA>>();
  const unsigned per_block = static_cast<unsigned>(kLaneWave * NA);
  const unsigned blocks = (static_cast<unsigned>(N) + per_block - 1) / per_block;
  kern<<<blocks, kLaneWave * WPT, lds, s>>>(N, basics, type_s, lt, e_s, grads);
}

}  // namespace

bool lane_shape_supported(int waves, int nslot, int atoms)
{
  if (atoms == 2) return nslot == 2 && (waves == 4 || waves == 8);
  if (atoms != 1) return false;
  return (nslot == 2 && (waves == 4 || waves == 8 || waves == 16)) ||
         (nslot == 4 && (waves == 1 || waves == 2 || waves == 4)) || (nslot == 8 && waves == 1);
}

template <typename Real>
bool launch_lane_graph(int N, const Real* basics, const int* type_s, const LaneTable<Real>& lt,
                       double* e_s, Real* grads, hipStream_t s, size_t max_lds)
{
  if (N <= 0) return true;
  const size_t lds = static_cast<size_t>(lt.slots) * kLaneWave * lt.atoms * sizeof(Real);
  if (lds > max_lds || lt.slots < lt.v_base + lt.waves) return false;
#define SUS2SH_LANE_CASE(NS, WPT, NA)                                                      \
  if (lt.nslot == NS && lt.waves == WPT && lt.atoms == NA) {                               \
    launch_lane_shape<Real, NS, WPT, NA>(N, basics, type_s, lt, e_s, grads, s, lds);       \
    return true;                                                                           \
  }
  SUS2SH_LANE_CASE(2, 4, 1)
  SUS2SH_LANE_CASE(2, 8, 1)
  SUS2SH_LANE_CASE(2, 16, 1)
  SUS2SH_LANE_CASE(4, 1, 1)
  SUS2SH_LANE_CASE(4, 2, 1)
  SUS2SH_LANE_CASE(4, 4, 1)
  SUS2SH_LANE_CASE(8, 1, 1)
  SUS2SH_LANE_CASE(2, 4, 2)
  SUS2SH_LANE_CASE(2, 8, 2)
#undef SUS2SH_LANE_CASE
  return false;
}

template bool launch_lane_graph<float>(int, const float*, const int*, const LaneTable<float>&,
                                       double*, float*, hipStream_t, size_t);
template bool launch_lane_graph<double>(int, const double*, const int*,
                                        const LaneTable<double>&, double*, double*, hipStream_t,
                                        size_t);

}  // namespace sus2sh
