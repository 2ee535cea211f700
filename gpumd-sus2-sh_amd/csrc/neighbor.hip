// Cell-list neighbour build over all periodic images within rc, and the final scatter of
// per-atom results.  Replaces SUS2_SH::build_neighbor_list (src/force/sus2_sh.cu:9547-9618):
// the reference uses GPUMD's cell list when every periodic thickness is >= 2 rc and an
// O(N^2) multi-image search otherwise (sus2_sh.cu:7668-7767).  One cell search with image
// shifts covers both cases: with n_d = floor(t_d / rc) cells along a direction of thickness
// t_d, every image within rc lies within R_d = ceil(rc n_d / t_d) cells of the home cell,
// and a small cell (n_d < 2 R_d + 1) is simply visited once per image shift.
//
// Atoms are sorted by cell, so every later kernel runs in cell order: neighbours of a
// wavefront's atoms sit in the same few cache lines.  The sort is a counting sort over cell
// keys (count, scan, place) whose order inside a cell is by atom index -- the permutation a
// stable sort of (key, index) gives, independent of atomic ordering -- and it uses kernels
// only (no memset or library call), so an evaluation can be captured in a HIP graph.
#include "kernels.hpp"

namespace sus2sh {

namespace {

// Structure of atom i in a batch (atom_off: [nstruct + 1] prefix offsets).
__device__ __forceinline__ int structure_of(int i, const int* __restrict__ atom_off, int nstruct)
{
  int lo = 0, hi = nstruct - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (atom_off[mid] <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Wrap into the box (fractional coordinates), cell key.  BATCH: every structure has its own
// box and cell grid (grids[s]) and its cells occupy keys [cell_off[s], cell_off[s+1]).
template <bool BATCH>
__global__ void k_bin(int N, const double* __restrict__ pos, CellGrid g1,
                      const CellGrid* __restrict__ grids, const int* __restrict__ cell_off,
                      const int* __restrict__ atom_off, int nstruct, int* __restrict__ key,
                      int* __restrict__ slot, int* __restrict__ cell_count,
                      double* __restrict__ wrapped, int* __restrict__ sid)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int st = 0;
  if (BATCH) {
    st = structure_of(i, atom_off, nstruct);
    sid[i] = st;
  }
  const CellGrid& g = BATCH ? grids[st] : g1;
  const double x = pos[i], y = pos[N + i], z = pos[2 * N + i];
  double s[3] = {g.hi[0] * x + g.hi[1] * y + g.hi[2] * z, g.hi[3] * x + g.hi[4] * y + g.hi[5] * z,
                 g.hi[6] * x + g.hi[7] * y + g.hi[8] * z};
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (g.pbc[d]) {
      s[d] -= floor(s[d]);
      if (s[d] >= 1.0) s[d] = 0.0;
    }
    int cd = static_cast<int>(floor(s[d] * g.n[d]));
    cd = cd < 0 ? 0 : (cd >= g.n[d] ? g.n[d] - 1 : cd);
    c[d] = cd;
  }
  wrapped[i] = g.h[0] * s[0] + g.h[1] * s[1] + g.h[2] * s[2];
  wrapped[N + i] = g.h[3] * s[0] + g.h[4] * s[1] + g.h[5] * s[2];
  wrapped[2 * N + i] = g.h[6] * s[0] + g.h[7] * s[1] + g.h[8] * s[2];
  const int k = (BATCH ? cell_off[st] : 0) + (c[2] * g.n[1] + c[1]) * g.n[0] + c[0];
  key[i] = k;
  slot[i] = atomicAdd(cell_count + k, 1);  // arrival order; made deterministic by k_rank
}

// Exclusive scan of cell counts: per 2048-entry tile (256 threads x 8), tile totals scanned
// by one workgroup, totals added back.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ int wave_inclusive_scan(int x)
{
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(x, off, 64);
    if (lane >= off) x += t;
  }
  return x;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(int n, const int* __restrict__ in,
                                                           int* __restrict__ out,
                                                           int* __restrict__ totals)
{
  __shared__ int v[kScanTile];
  __shared__ int wsum[kScanBlock / 64];
  const int base = blockIdx.x * kScanTile;
  for (int k = threadIdx.x; k < kScanTile; k += kScanBlock)
    v[k] = base + k < n ? in[base + k] : 0;
  __syncthreads();
  int loc[kScanItems];
  int sum = 0;
#pragma unroll
  for (int u = 0; u < kScanItems; ++u) {
    loc[u] = sum;
    sum += v[threadIdx.x * kScanItems + u];
  }
  const int incl = wave_inclusive_scan(sum);
  if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = incl;
  __syncthreads();
  int pre = incl - sum;
  for (int w = 0; w < static_cast<int>(threadIdx.x >> 6); ++w) pre += wsum[w];
#pragma unroll
  for (int u = 0; u < kScanItems; ++u) v[threadIdx.x * kScanItems + u] = pre + loc[u];
  __syncthreads();
  for (int k = threadIdx.x; k < kScanTile; k += kScanBlock)
    if (base + k < n) out[base + k] = v[k];
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) t += wsum[w];
    totals[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(1024) void k_scan_totals(int n, int* __restrict__ totals)
{
  __shared__ int wsum[16];
  int carry = 0;
  for (int base = 0; base < n; base += 1024) {
    const int k = base + static_cast<int>(threadIdx.x);
    const int x = k < n ? totals[k] : 0;
    const int incl = wave_inclusive_scan(x);
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = incl;
    __syncthreads();
    int pre = incl - x, all = 0;
    for (int w = 0; w < 16; ++w) {
      if (w < static_cast<int>(threadIdx.x >> 6)) pre += wsum[w];
      all += wsum[w];
    }
    if (k < n) totals[k] = carry + pre;
    carry += all;
    __syncthreads();
  }
}

__global__ void k_scan_add(int n, int* __restrict__ out, const int* __restrict__ totals)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] += totals[k / kScanTile];
}

__global__ void k_place(int N, const int* __restrict__ key, const int* __restrict__ slot,
                        const int* __restrict__ cbeg, int* __restrict__ arrival)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  arrival[cbeg[key[i]] + slot[i]] = i;
}

__global__ void k_fill_int(int* __restrict__ p, int v, size_t n)
{
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_double(double* __restrict__ p, double v, size_t n)
{
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Final position of each atom: its cell's start plus the number of atoms of the same cell
// with a smaller index; then the gather of wrapped positions and types into that order.
__global__ void k_rank(int N, int S, const int* __restrict__ key, const int* __restrict__ arrival,
                       const int* __restrict__ cbeg, const double* __restrict__ wrapped,
                       const int* __restrict__ type, int* __restrict__ satom,
                       int* __restrict__ skey, double* __restrict__ pos_s,
                       PackedAtom* __restrict__ packed, int* __restrict__ type_s,
                       int* __restrict__ bad_type, const int* __restrict__ sid,
                       int* __restrict__ sid_s, int* __restrict__ rank)
{
  const int q = xcd_block() * blockDim.x + threadIdx.x;
  if (q >= N) return;
  const int i = arrival[q];
  const int k = key[i];
  const int b = cbeg[k], e = cbeg[k + 1];
  int r = 0;
  for (int u = b; u < e; ++u) r += arrival[u] < i;
  const int p = b + r;
  satom[p] = i;
  rank[i] = p;
  skey[p] = k;
  pos_s[p] = wrapped[i];
  pos_s[N + p] = wrapped[N + i];
  pos_s[2 * N + p] = wrapped[2 * N + i];
  int t = type[i];
  if (t < 0 || t >= S) {  // reported to the host; clamped so that no table is overrun
    atomicOr(bad_type, 1);
    t = 0;
  }
  type_s[p] = t;
  packed[p] = PackedAtom{wrapped[i], wrapped[N + i], wrapped[2 * N + i], static_cast<double>(t)};
  if (sid_s) sid_s[p] = sid[i];
}

__device__ __forceinline__ int wrap_cell(int c, int n, int& shift)
{
  // floor division for possibly negative c
  shift = (c >= 0) ? c / n : -((-c + n - 1) / n);
  return c - shift * n;
}

// A candidate is one 32-byte (x, y, z, type) record of the packed copy the cell sort writes:
// two 16-byte loads from one cache line instead of three gathers from three arrays and a
// fourth for the species of each neighbour (the kernel is bound by these gathers' latency).
template <typename Real>
__device__ __forceinline__ void test_candidate(const PackedAtom& pq, int q, int p, bool self_image,
                                               double xp, double yp, double zp, double shx,
                                               double shy, double shz, double rc2, bool valid,
                                               int cap, int N, int& c, Edge<Real>* __restrict__ nl)
{
  if (self_image && q == p) return;
  const double dx = pq.x - xp + shx;
  const double dy = pq.y - yp + shy;
  const double dz = pq.z - zp + shz;
  if (dx * dx + dy * dy + dz * dz < rc2) {
    if (valid && c < cap) {
      Edge<Real> e;
      e.dx = static_cast<Real>(dx);
      e.dy = static_cast<Real>(dy);
      e.dz = static_cast<Real>(dz);
      e.j = q | (static_cast<int>(pq.t) << kEdgeTypeShift);
      nl[static_cast<size_t>(c) * N + p] = e;
    }
    ++c;
  }
}

// Batch: a workgroup's candidates are the atoms of the structures its atoms belong to, one
// contiguous sorted range; up to kBatchStage of them (structures of <= 256 atoms around a
// 256-atom workgroup: at most 766) are copied into LDS once, so that the candidate loads of
// the search -- a serial chain per lane, hundreds of them for a small triclinic cell's
// images -- are LDS reads instead of L2 round trips.
constexpr int kBatchStage = 768;  // 24 KiB
// The cell ranges of those structures, the search's other serial loads, are staged too when
// they number at most kBatchCells.
constexpr int kBatchCells = 256;  // 2 KiB

// The candidates [qb, qe) of one cell, candidate q at src[q - qoff]: four records' loads in
// flight together, then their tests.
template <typename Real>
__device__ __forceinline__ void scan_cell(const PackedAtom* __restrict__ src, int qoff, int qb,
                                          int qe, int p, bool self_image, double xp, double yp,
                                          double zp, double shx, double shy, double shz,
                                          double rc2, bool valid, int cap, int N, int& c,
                                          Edge<Real>* __restrict__ nl)
{
  int q = qb;
  for (; q + 4 <= qe; q += 4) {
    PackedAtom c4[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c4[u] = src[q + u - qoff];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      test_candidate<Real>(c4[u], q + u, p, self_image, xp, yp, zp, shx, shy, shz, rc2, valid,
                           cap, N, c, nl);
  }
  for (; q < qe; ++q)
    test_candidate<Real>(src[q - qoff], q, p, self_image, xp, yp, zp, shx, shy, shz, rc2, valid,
                         cap, N, c, nl);
}

template <typename Real, bool BATCH>
__global__ __launch_bounds__(256) void k_neighbors(NeighborArgs a, const double* __restrict__ pos,
                                                   const PackedAtom* __restrict__ packed,
                                                   const int* __restrict__ skey,
                                                   const int* __restrict__ cbeg,
                                                   const int* __restrict__ cend,
                                                   int* __restrict__ count,
                                                   Edge<Real>* __restrict__ nl,
                                                   int* __restrict__ max_count)
{
  const int N = a.N;
  const int p0 = xcd_block() * blockDim.x + threadIdx.x;
  const bool valid = p0 < N;
  const int p = valid ? p0 : N - 1;  // inactive lanes repeat the last atom, store nothing
  const int st = BATCH ? a.sid_s[p] : 0;
  const CellGrid& g = BATCH ? a.grids[st] : a.grid;
  const int base = BATCH ? a.cell_off[st] : 0;
  const int key = skey[p] - base;
  const int cx = key % g.n[0];
  const int cy = (key / g.n[0]) % g.n[1];
  const int cz = key / (g.n[0] * g.n[1]);
  const double xp = pos[p], yp = pos[N + p], zp = pos[2 * N + p];
  __shared__ PackedAtom s_atoms[BATCH ? kBatchStage : 1];
  __shared__ int s_cb[BATCH ? kBatchCells : 1], s_ce[BATCH ? kBatchCells : 1];
  bool staged = false, cells_staged = false;
  int qoff = 0, coff = 0;
  if constexpr (BATCH) {
    const int pf = xcd_block() * static_cast<int>(blockDim.x);
    const int pl = min(pf + static_cast<int>(blockDim.x), N) - 1;
    const int s0 = a.sid_s[pf], s1 = a.sid_s[pl];
    const int lo = a.atom_off[s0], hi = a.atom_off[s1 + 1];
    const int clo = a.cell_off[s0], nc = a.cell_off[s1 + 1] - clo;
    staged = hi - lo <= kBatchStage;  // workgroup-uniform
    cells_staged = nc <= kBatchCells;
    if (staged) {
      for (int t = threadIdx.x; t < hi - lo; t += blockDim.x) s_atoms[t] = packed[lo + t];
      qoff = lo;
    }
    if (cells_staged) {
      for (int t = threadIdx.x; t < nc; t += blockDim.x) {
        s_cb[t] = cbeg[clo + t];
        s_ce[t] = cend[clo + t];
      }
      coff = clo;
    }
    if (staged || cells_staged) __syncthreads();
  }
  int c = 0;
  for (int oz = -g.R[2]; oz <= g.R[2]; ++oz) {
    int sz = 0, z = cz + oz;
    if (g.pbc[2]) {
      z = wrap_cell(z, g.n[2], sz);
    } else if (z < 0 || z >= g.n[2]) {
      continue;
    }
    for (int oy = -g.R[1]; oy <= g.R[1]; ++oy) {
      int sy = 0, y = cy + oy;
      if (g.pbc[1]) {
        y = wrap_cell(y, g.n[1], sy);
      } else if (y < 0 || y >= g.n[1]) {
        continue;
      }
      for (int ox = -g.R[0]; ox <= g.R[0]; ++ox) {
        int sx = 0, x = cx + ox;
        if (g.pbc[0]) {
          x = wrap_cell(x, g.n[0], sx);
        } else if (x < 0 || x >= g.n[0]) {
          continue;
        }
        const int cc = base + (z * g.n[1] + y) * g.n[0] + x;
        const double shx = sx * g.h[0] + sy * g.h[1] + sz * g.h[2];
        const double shy = sx * g.h[3] + sy * g.h[4] + sz * g.h[5];
        const double shz = sx * g.h[6] + sy * g.h[7] + sz * g.h[8];
        const bool self_image = (sx | sy | sz) == 0;
        int qb, qe;
        if (BATCH && cells_staged) {
          qb = s_cb[cc - coff];
          qe = s_ce[cc - coff];
        } else {
          qb = cbeg[cc];
          qe = cend[cc];
        }
        if (BATCH && staged)
          scan_cell<Real>(s_atoms, qoff, qb, qe, p, self_image, xp, yp, zp, shx, shy, shz, a.rc2,
                          valid, a.cap, N, c, nl);
        else
          scan_cell<Real>(packed, 0, qb, qe, p, self_image, xp, yp, zp, shx, shy, shz, a.rc2,
                          valid, a.cap, N, c, nl);
      }
    }
  }
  if (valid) count[p] = c < a.cap ? c : a.cap;
  // wavefront reduction, then one atomic per wave: max count and total edge count
  int cmax = valid ? c : 0;
  unsigned long long csum = valid ? static_cast<unsigned long long>(c) : 0ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    cmax = max(cmax, __shfl_xor(cmax, off, 64));
    csum += __shfl_xor(csum, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(max_count, cmax);
    atomicAdd(reinterpret_cast<unsigned long long*>(max_count + 2), csum);
  }
}

// The same search for one structure whose periodic cells are all at least rc wide (each
// direction has >= 3 cells and a +-1 cell reach), with the candidates staged in LDS.  A
// wave (one 64-atom workgroup) spans one or two cell rows (y, z); for each of the 9 row
// offsets (oz, oy) -- in the order of k_neighbors, so every list comes out identical -- it
// copies the atoms of the cells its lanes can reach in that band (x from the first lane's
// cell - 1 to the last lane's + 1 per row) from the cell-sorted packed copy into LDS with
// coalesced loads, then every lane scans its three cells there.  A band with more than
// kStageCap atoms, and a wave over more than two rows, read the candidates from global
// memory as k_neighbors does.
// a band of the benchmark's fcc alloy holds ~80-160 atoms (one or two rows of ~10 cells): 256
// staged atoms (8 KiB) cover it and keep the workgroup's LDS small enough for ~16 waves per CU
// (512: 0.87 ms, 256: 0.82 ms, 128: 0.83 ms at 1,024,000 atoms, FP32)
constexpr int kStageCap = 256;  // staged candidate atoms per band (8 KiB)
constexpr int kStageTab = 128;  // staged cells per band


template <typename Real>
__global__ __launch_bounds__(64) void k_neighbors_staged(NeighborArgs a,
                                                         const double* __restrict__ pos,
                                                         const PackedAtom* __restrict__ packed,
                                                         const int* __restrict__ skey,
                                                         const int* __restrict__ cbeg,
                                                         const int* __restrict__ cend,
                                                         int* __restrict__ count,
                                                         Edge<Real>* __restrict__ nl,
                                                         int* __restrict__ max_count)
{
  __shared__ PackedAtom s_cand[kStageCap];
  __shared__ int s_off[kStageTab + 1];
  __shared__ int s_src[kStageTab];
  const CellGrid& g = a.grid;
  const int N = a.N;
  const int lane = threadIdx.x;
  const int p0 = xcd_block() * 64;
  const int p = p0 + lane < N ? p0 + lane : N - 1;
  const bool valid = p0 + lane < N;
  const int nx = g.n[0], ny = g.n[1];
  const int key = skey[p];
  const int kf = __builtin_amdgcn_readfirstlane(skey[p0]);
  const int kl = __builtin_amdgcn_readfirstlane(skey[min(p0 + 63, N - 1)]);
  const int rf = kf / nx, rl = kl / nx;
  const bool wave_direct = rl - rf > 1;
  // x extent of the wave's cells per row, and each row's offset in the band's cell table
  const int xlo0 = kf % nx, xhi0 = rl > rf ? nx - 1 : kl % nx;
  const int xlo1 = 0, xhi1 = kl % nx;
  const int w0 = xhi0 - xlo0 + 3, w1 = rl > rf ? xhi1 - xlo1 + 3 : 0;
  const int ntab = w0 + w1;
  const bool tab_ok = !wave_direct && ntab <= kStageTab;
  const int cx = key % nx;
  const int row = key / nx;
  const int cy = row % ny, cz = row / ny;
  const int ri = row - rf;  // the lane's row: 0 or 1
  const int tbase = ri == 0 ? 0 : w0;
  const int txlo = ri == 0 ? xlo0 : xlo1;
  const double xp = pos[p], yp = pos[N + p], zp = pos[2 * N + p];
  int c = 0;
  for (int oz = -1; oz <= 1; ++oz) {
    for (int oy = -1; oy <= 1; ++oy) {
      // --- the band's cell table and staged atoms (wave-uniform decisions) ---
      bool staged = false;
      if (tab_ok) {
        for (int e2 = 2 * lane; e2 < 2 * lane + 2; ++e2) {
          int src = 0, cnt = 0;
          if (e2 < ntab) {
            const int r = e2 < w0 ? 0 : 1;
            const int rr = rf + r;
            int zb = rr / ny + oz, yb = rr % ny + oy, xb = (r == 0 ? xlo0 : xlo1) - 1 + (e2 - (r == 0 ? 0 : w0));
            int sh;
            bool ok = true;
            if (g.pbc[2]) zb = wrap_cell(zb, g.n[2], sh); else ok = ok && zb >= 0 && zb < g.n[2];
            if (g.pbc[1]) yb = wrap_cell(yb, ny, sh); else ok = ok && yb >= 0 && yb < ny;
            if (g.pbc[0]) xb = wrap_cell(xb, nx, sh); else ok = ok && xb >= 0 && xb < nx;
            if (ok) {
              const int cc = (zb * ny + yb) * nx + xb;
              src = cbeg[cc];
              cnt = cend[cc] - src;
            }
          }
          if (e2 < kStageTab) s_src[e2] = src;
          // counts go through s_off for the scan below
          if (e2 < kStageTab) s_off[e2] = cnt;
        }
        __syncthreads();
        // exclusive prefix over the (at most 128) cell counts: two entries per lane
        const int c0 = s_off[2 * lane], c1 = s_off[2 * lane + 1];
        int incl = c0 + c1;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const int v = __shfl_up(incl, off, 64);
          if (lane >= off) incl += v;
        }
        const int excl = incl - (c0 + c1);
        const int total = __builtin_amdgcn_readlane(incl, 63);
        __syncthreads();
        s_off[2 * lane] = excl;
        s_off[2 * lane + 1] = excl + c0;
        if (lane == 0) s_off[kStageTab] = total;
        __syncthreads();
        staged = total <= kStageCap;
        if (staged) {
          for (int t = lane; t < total; t += 64) {
            int lo = 0, hi = kStageTab - 1;  // last entry with s_off[e] <= t
            while (lo < hi) {
              const int mid = (lo + hi + 1) >> 1;
              if (s_off[mid] <= t)
                lo = mid;
              else
                hi = mid - 1;
            }
            s_cand[t] = packed[s_src[lo] + (t - s_off[lo])];
          }
          __syncthreads();
        }
      }
      // --- this lane's three cells of the band ---
      int sz = 0, z = cz + oz;
      bool band_ok = true;
      if (g.pbc[2]) z = wrap_cell(z, g.n[2], sz); else band_ok = z >= 0 && z < g.n[2];
      int sy = 0, y = cy + oy;
      if (g.pbc[1]) y = wrap_cell(y, ny, sy); else band_ok = band_ok && y >= 0 && y < ny;
      if (band_ok) {
        for (int ox = -1; ox <= 1; ++ox) {
          int sx = 0, x = cx + ox;
          if (g.pbc[0]) {
            x = wrap_cell(x, nx, sx);
          } else if (x < 0 || x >= nx) {
            continue;
          }
          const double shx = sx * g.h[0] + sy * g.h[1] + sz * g.h[2];
          const double shy = sx * g.h[3] + sy * g.h[4] + sz * g.h[5];
          const double shz = sx * g.h[6] + sy * g.h[7] + sz * g.h[8];
          const bool self_image = (sx | sy | sz) == 0;
          if (staged) {
            const int e = tbase + (cx + ox) - (txlo - 1);
            const int o = s_off[e], n = s_off[e + 1] - o, q0 = s_src[e];
            for (int k = 0; k < n; ++k)
              test_candidate<Real>(s_cand[o + k], q0 + k, p, self_image, xp, yp, zp, shx, shy, shz,
                                   a.rc2, valid, a.cap, N, c, nl);
          } else {
            const int cc = (z * ny + y) * nx + x;
            const int qe = cend[cc];
            for (int q = cbeg[cc]; q < qe; ++q)
              test_candidate<Real>(packed[q], q, p, self_image, xp, yp, zp, shx, shy, shz, a.rc2,
                                   valid, a.cap, N, c, nl);
          }
        }
      }
      if (tab_ok) __syncthreads();  // the next band rewrites the staged atoms
    }
  }
  if (valid) count[p] = c < a.cap ? c : a.cap;
  int cmax = valid ? c : 0;
  unsigned long long csum = valid ? static_cast<unsigned long long>(c) : 0ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    cmax = max(cmax, __shfl_xor(cmax, off, 64));
    csum += __shfl_xor(csum, off, 64);
  }
  if (lane == 0) {
    atomicMax(max_count, cmax);
    atomicAdd(reinterpret_cast<unsigned long long*>(max_count + 2), csum);
  }
}

// Caller order: lane i adds the results of its sorted position p = rank[i] -- the caller's
// arrays are read-modify-written with coalesced accesses, the sorted arrays are gathered.
__global__ void k_scatter(int N, const int* __restrict__ rank, const double* __restrict__ e_s,
                          const double* __restrict__ f_s, const double* __restrict__ v_s,
                          double* __restrict__ pot, double* __restrict__ force,
                          double* __restrict__ vir, int fixed)
{
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int p = rank[i];
  pot[i] += e_s[p];
  const long long* f_q = reinterpret_cast<const long long*>(f_s);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t k = static_cast<size_t>(c) * N + p;
    force[static_cast<size_t>(c) * N + i] +=
      fixed ? static_cast<double>(f_q[k]) * kFixInvScale : f_s[k];
  }
#pragma unroll
  for (int c = 0; c < 9; ++c) vir[static_cast<size_t>(c) * N + i] += v_s[static_cast<size_t>(c) * N + p];
}

// Deferred checks: latch[0] = max neighbour count, latch[1] |= bad type, latch[2..3] = the
// last call's total edge count, latch[4] = number of latched calls.
__global__ void k_latch(const int* __restrict__ mc, int* __restrict__ latch)
{
  if (threadIdx.x != 0) return;
  latch[0] = max(latch[0], mc[0]);
  latch[1] |= mc[1];
  latch[2] = mc[2];
  latch[3] = mc[3];
  latch[4] += 1;
}

// First sorted atom whose cell key is >= k.
__device__ __forceinline__ int lower_bound_key(const int* __restrict__ skey, int N, int k)
{
  int lo = 0, hi = N;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (skey[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// One wave per force tile.  Candidate intervals: for every cell row (y, z) the tile touches
// and every row offset (oy, oz) within +-R, the x range of the tile's cells widened by +-R,
// wrapped (two pieces) or clipped like the neighbour search; each cell range becomes an atom
// range by binary search on the sorted keys.  The wave then sorts the ranges by start and
// merges overlapping or touching ones.
constexpr int kWinCand = 128;

__global__ __launch_bounds__(64) void k_windows(int N, CellGrid g, const int* __restrict__ skey,
                                                int* __restrict__ win, int cap)
{
  __shared__ int cs[kWinCand], ce[kWinCand];
  const int tile = blockIdx.x;
  const int p0 = tile * kWinTile;
  const int p1 = min(p0 + kWinTile, N) - 1;
  int* out = win + static_cast<size_t>(tile) * kWinStride;
  const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
  const int k_lo = skey[p0], k_hi = skey[p1];
  const int r_lo = k_lo / n0, r_hi = k_hi / n0;
  const int ny = 2 * g.R[1] + 1, nz = 2 * g.R[2] + 1;
  const int total = (r_hi - r_lo + 1) * ny * nz * 2;
  if (total > kWinCand) {
    if (threadIdx.x == 0) out[0] = -1;
    return;
  }
  for (int c = threadIdx.x; c < total; c += 64) {
    const int piece = c & 1;
    int t = c >> 1;
    const int oz = t % nz - g.R[2];
    t /= nz;
    const int oy = t % ny - g.R[1];
    const int r = r_lo + t / ny;
    const int y = r % n1, z = r / n1;
    const int xa = r == r_lo ? k_lo % n0 : 0;
    const int xb = r == r_hi ? k_hi % n0 : n0 - 1;
    int yy = y + oy, zz = z + oz;
    bool ok = true;
    if (g.pbc[1]) yy = ((yy % n1) + n1) % n1; else ok = ok && yy >= 0 && yy < n1;
    if (g.pbc[2]) zz = ((zz % n2) + n2) % n2; else ok = ok && zz >= 0 && zz < n2;
    int x1 = xa - g.R[0], x2 = xb + g.R[0];
    int a = 0, b = -1;  // cell x range of this piece (empty when b < a)
    if (!g.pbc[0]) {
      if (piece == 0) {
        a = max(x1, 0);
        b = min(x2, n0 - 1);
      }
    } else if (x2 - x1 + 1 >= n0) {
      if (piece == 0) {
        a = 0;
        b = n0 - 1;
      }
    } else if (x1 < 0) {
      a = piece == 0 ? x1 + n0 : 0;
      b = piece == 0 ? n0 - 1 : x2;
    } else if (x2 >= n0) {
      a = piece == 0 ? x1 : 0;
      b = piece == 0 ? n0 - 1 : x2 - n0;
    } else if (piece == 0) {
      a = x1;
      b = x2;
    }
    int s0 = 0, s1 = 0;
    if (ok && b >= a) {
      const int row = (zz * n1 + yy) * n0;
      s0 = lower_bound_key(skey, N, row + a);
      s1 = lower_bound_key(skey, N, row + b + 1);
    }
    cs[c] = s0;
    ce[c] = s1;
  }
  __syncthreads();
  // Sort the ranges by start (rank of each among the non-empty ones, ties by index) and merge
  // overlapping or touching ranges -- the whole wave at once: every lane ranks its two ranges
  // against all of them (broadcast LDS reads), then scans the sorted list.  Empty ranges rank
  // last and are dropped.
  __shared__ int ss[kWinCand], se[kWinCand];
  constexpr int kNone = 0x7fffffff;
  const int lane = threadIdx.x;
  int a0 = kNone, b0 = 0, a1 = kNone, b1 = 0;
  if (lane < total && ce[lane] > cs[lane]) {
    a0 = cs[lane];
    b0 = ce[lane];
  }
  if (lane + 64 < total && ce[lane + 64] > cs[lane + 64]) {
    a1 = cs[lane + 64];
    b1 = ce[lane + 64];
  }
  int r0 = 0, r1 = 0;
  for (int j = 0; j < total; ++j) {
    const int sj = ce[j] > cs[j] ? cs[j] : kNone;
    r0 += (sj < a0 || (sj == a0 && j < lane)) ? 1 : 0;
    r1 += (sj < a1 || (sj == a1 && j < lane + 64)) ? 1 : 0;
  }
  if (a0 != kNone) {
    ss[r0] = a0;
    se[r0] = b0;
  }
  if (a1 != kNone) {
    ss[r1] = a1;
    se[r1] = b1;
  }
  int nvalid = (a0 != kNone ? 1 : 0) + (a1 != kNone ? 1 : 0);
  for (int off = 32; off > 0; off >>= 1) nvalid += __shfl_xor(nvalid, off, 64);
  __syncthreads();
  // sorted element k = 2 lane + v (v = 0, 1): it starts a merged window when its start lies
  // beyond the end of every earlier range (the exclusive prefix max of the ends); a window's
  // end is the prefix max of the ends at its last element
  const int k1 = 2 * lane + 1;
  const bool v0 = 2 * lane < nvalid, v1 = k1 < nvalid;
  const int s0 = v0 ? ss[2 * lane] : 0, e0 = v0 ? se[2 * lane] : -1;
  const int s1 = v1 ? ss[k1] : 0, e1 = v1 ? se[k1] : -1;
  int incl = max(e0, e1);
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off, 64);
    if (lane >= off) incl = max(incl, t);
  }
  int excl = __shfl_up(incl, 1, 64);
  if (lane == 0) excl = -1;
  const int end0 = max(excl, e0), end1 = max(end0, e1);
  const bool st0 = v0 && s0 > excl;
  const bool st1 = v1 && s1 > end0;
  // windows started up to and including element k1 (inclusive prefix count)
  int cst = (st0 ? 1 : 0) + (st1 ? 1 : 0);
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(cst, off, 64);
    if (lane >= off) cst += t;
  }
  const int count = __shfl(cst, 63, 64);
  const int win1 = cst - 1;                  // window of element k1
  const int win0 = cst - (st1 ? 1 : 0) - 1;  // window of element k0
  // start of the window of element k1: starts increase along the sorted list, so it is the
  // prefix max of the starts seen so far
  int gs = st1 ? s1 : (st0 ? s0 : -1);
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(gs, off, 64);
    if (lane >= off) gs = max(gs, t);
  }
  int gs_prev = __shfl_up(gs, 1, 64);
  if (lane == 0) gs_prev = -1;
  const int gs0 = st0 ? s0 : gs_prev;  // start of the window of element k0
  const int gs1 = st1 ? s1 : gs0;
  // element k closes its window when element k + 1 starts one or k is the last
  const int next_st = __shfl_down(st0 ? 1 : 0, 1, 64);  // st0 of the next lane
  const bool last0 = v0 && (!v1 || st1);
  const bool last1 = v1 && (k1 + 1 >= nvalid || next_st != 0);
  if (count > kWinMax) {
    if (lane == 0) out[0] = -1;
    return;
  }
  int sz = 0;
  if (last0) {
    out[2 + 2 * win0] = gs0;
    out[3 + 2 * win0] = end0 - gs0;
    sz += end0 - gs0;
  }
  if (last1) {
    out[2 + 2 * win1] = gs1;
    out[3 + 2 * win1] = end1 - gs1;
    sz += end1 - gs1;
  }
  for (int off = 32; off > 0; off >>= 1) sz += __shfl_xor(sz, off, 64);
  if (lane == 0) {
    out[0] = sz <= cap ? count : -1;
    out[1] = sz;
  }
}

inline int blocks_for(size_t n, int bs) { return static_cast<int>((n + bs - 1) / bs); }

}  // namespace

void launch_cell_sort(const CellSortArgs& a, hipStream_t s)
{
  const int N = a.N;
  const int n = a.ncell + 1;  // counts of every cell plus one zero: cbeg[ncell] = N
  const int tiles = blocks_for(n, kScanTile);
  launch_fill_int(a.cell_count, 0, n, s);
  if (a.grids)
    k_bin<true><<<blocks_for(N, 256), 256, 0, s>>>(N, a.pos, a.grid, a.grids, a.cell_off,
                                                   a.atom_off, a.nstruct, a.key, a.slot,
                                                   a.cell_count, a.wrapped, a.sid);
  else
    k_bin<false><<<blocks_for(N, 256), 256, 0, s>>>(N, a.pos, a.grid, nullptr, nullptr, nullptr,
                                                    1, a.key, a.slot, a.cell_count, a.wrapped,
                                                    nullptr);
  k_scan_tiles<<<tiles, kScanBlock, 0, s>>>(n, a.cell_count, a.cell_begin, a.scan_totals);
  if (tiles > 1) {
    k_scan_totals<<<1, 1024, 0, s>>>(tiles, a.scan_totals);
    k_scan_add<<<blocks_for(n, 256), 256, 0, s>>>(n, a.cell_begin, a.scan_totals);
  }
  k_place<<<blocks_for(N, 256), 256, 0, s>>>(N, a.key, a.slot, a.cell_begin, a.arrival);
  k_rank<<<blocks_for(N, 256), 256, 0, s>>>(N, a.S, a.key, a.arrival, a.cell_begin, a.wrapped,
                                             a.type, a.satom, a.skey, a.pos_s, a.packed,
                                             a.type_s, a.bad_type, a.sid, a.sid_s, a.rank);
}

void launch_latch(const int* max_count, int* latch, hipStream_t s)
{
  k_latch<<<1, 64, 0, s>>>(max_count, latch);
}

size_t cell_sort_totals(int ncell) { return static_cast<size_t>(blocks_for(ncell + 1, kScanTile)); }

void launch_fill_int(int* p, int value, size_t n, hipStream_t s)
{
  if (n == 0) return;
  k_fill_int<<<blocks_for(n, 256), 256, 0, s>>>(p, value, n);
}

void launch_fill_double(double* p, double value, size_t n, hipStream_t s)
{
  if (n == 0) return;
  k_fill_double<<<blocks_for(n, 256), 256, 0, s>>>(p, value, n);
}

template <typename Real>
void launch_neighbors(const NeighborArgs& a, const double* pos_s, const PackedAtom* packed,
                      const int* sorted_key, const int* cell_begin, const int* cell_end,
                      int* count, Edge<Real>* nl, int* max_count, hipStream_t s)
{
  // the staged search needs every periodic direction split into >= 3 cells reached by +-1
  bool stage = !a.grids;
  for (int d = 0; d < 3 && stage; ++d)
    stage = a.grid.R[d] == 1 && (a.grid.n[d] >= 3 || !a.grid.pbc[d]);
  if (a.grids)
    k_neighbors<Real, true><<<blocks_for(a.N, 256), 256, 0, s>>>(
      a, pos_s, packed, sorted_key, cell_begin, cell_end, count, nl, max_count);
  else if (stage)
    k_neighbors_staged<Real><<<blocks_for(a.N, 64), 64, 0, s>>>(
      a, pos_s, packed, sorted_key, cell_begin, cell_end, count, nl, max_count);
  else
    k_neighbors<Real, false><<<blocks_for(a.N, 256), 256, 0, s>>>(
      a, pos_s, packed, sorted_key, cell_begin, cell_end, count, nl, max_count);
}

template void launch_neighbors<float>(const NeighborArgs&, const double*, const PackedAtom*,
                                      const int*, const int*, const int*, int*, Edge<float>*,
                                      int*, hipStream_t);
template void launch_neighbors<double>(const NeighborArgs&, const double*, const PackedAtom*,
                                       const int*, const int*, const int*, int*, Edge<double>*,
                                       int*, hipStream_t);

void launch_windows(int N, const CellGrid& g, const int* sorted_key, int* win, int cap,
                    hipStream_t s)
{
  if (N <= 0) return;
  k_windows<<<blocks_for(N, kWinTile), 64, 0, s>>>(N, g, sorted_key, win, cap);
}

namespace {

// Batch windows, one lane per tile: the structures of a batch occupy consecutive ranges of
// the sorted order (atom_off), and every neighbour of an atom belongs to its own structure,
// so a tile's neighbours lie in the one range spanning its first to its last structure.
__global__ void k_windows_batch(int N, const int* __restrict__ sid_s,
                                const int* __restrict__ atom_off, int* __restrict__ win, int cap)
{
  const int tile = blockIdx.x * blockDim.x + threadIdx.x;
  const int p0 = tile * kWinTile;
  if (p0 >= N) return;
  const int p1 = min(p0 + kWinTile, N) - 1;
  const int a = atom_off[sid_s[p0]], b = atom_off[sid_s[p1] + 1];
  int* out = win + static_cast<size_t>(tile) * kWinStride;
  if (b - a <= cap) {
    out[0] = 1;
    out[1] = b - a;
    out[2] = a;
    out[3] = b - a;
  } else {
    out[0] = -1;
  }
}

}  // namespace

void launch_windows_batch(int N, const int* sid_s, const int* atom_off, int* win, int cap,
                          hipStream_t s)
{
  if (N <= 0) return;
  const int tiles = blocks_for(N, kWinTile);
  k_windows_batch<<<blocks_for(tiles, 64), 64, 0, s>>>(N, sid_s, atom_off, win, cap);
}

void launch_scatter(int N, const int* rank, const double* e_s, const double* f_s,
                    const double* v_s, double* potential, double* force, double* virial,
                    int fixed, hipStream_t s)
{
  k_scatter<<<blocks_for(N, 256), 256, 0, s>>>(N, rank, e_s, f_s, v_s, potential, force, virial,
                                               fixed);
}

}  // namespace sus2sh
