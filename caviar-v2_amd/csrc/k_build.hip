// k_build.hip -- periodic wrap, cell sort and Verlet-list build on the device.
//
// A build runs only when the device flag F_REBUILD is set (by the
// displacement check in the integrator kernels, or by the host for a forced
// build); every kernel below returns at once otherwise, so a whole MD step
// can be queued without a host round trip.
//
// Reference behaviour restated here:
//   wrap        Domain::fix_position (domain.cpp:443-493)
//   cells       Neighborlist::update_cell_list / binlist_index
//               (neighborlist.cpp:392-455) -- own grid, see below
//   list        Neighborlist::update_verlet_list (neighborlist.cpp:247-297)
//               with ghost images as exchange_ghost_single_md_domain.cpp:67-215
//               makes them.
//
// Pair-set parity: the predicate is evaluated exactly as the reference does,
// in FP64 with no contraction: the ghost coordinate is formed first,
// g = p_j - v*L (one rounding), then dr = g - p_i and
// r^2 = (dx*dx + dy*dy) + dz*dz < (cutoff*coef)^2.  Ghost images are not
// stored: for each candidate the only image that can lie within the list
// cutoff (the minimum image, boxes being wider than twice the cutoff) is
// generated on the fly, and kept only if the reference would have created
// that ghost (owner within ghost_cutoff of the face, same priority rule).
// The cell grid is only a candidate filter (cells at least as wide as the
// list cutoff), so the pair set does not depend on it.
#include <algorithm>
#include <cmath>

#include "cv2_internal.h"

namespace cv2 {

namespace {

__device__ inline int axis_cell(double p, double glo, double cinv, int nc) {
  int c = (int)floor((p - glo) * cinv);
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__device__ inline int cell_of(const double4 &p, const BoxDev &b) {
  const int cx = axis_cell(p.x, b.glo[0], b.cinv[0], b.nc[0]);
  const int cy = axis_cell(p.y, b.glo[1], b.cinv[1], b.nc[1]);
  const int cz = axis_cell(p.z, b.glo[2], b.cinv[2], b.nc[2]);
  return cx + b.nc[0] * (cy + b.nc[1] * cz);
}

// Wrap, bin, count.  One thread per owned atom.  The atoms are still in the
// previous build's cell order, so neighbouring lanes mostly share a cell:
// each run of equal cells in consecutive lanes takes its slots with one
// atomic add (the in-cell order is fixed afterwards by k_rank).
__global__ void k_wrap_bin(double4 *__restrict__ pos, int n, BoxDev b, int *__restrict__ cell,
                           int *__restrict__ slot, int *__restrict__ cell_count,
                           int *__restrict__ flags) {
  if (!build_pending(flags)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {  // statistics of this build (k_build_stats)
    flags[F_MAXNBR] = 0;
    flags[F_ENTRIES] = 0;
  }
  const int lane = threadIdx.x & 63;
  const bool valid = i < n;
  int ci = -1;
  if (valid) {
    double4 p = pos[i];
    double *c[3] = {&p.x, &p.y, &p.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (b.wrap[d] != 1) continue;
      if (*c[d] < b.lo[d])
        *c[d] += b.L[d];
      else if (*c[d] > b.hi[d])
        *c[d] -= b.L[d];
    }
    pos[i] = p;
    ci = cell_of(p, b);
    cell[i] = ci;
  }
  const int prev = __shfl_up(ci, 1, 64);
  const bool lead = valid && (lane == 0 || prev != ci);
  const unsigned long long lm = __ballot(lead);
  const int nvalid = __popcll(__ballot(valid));  // valid lanes are a prefix
  const unsigned long long upto = lane == 63 ? lm : lm & ((2ull << lane) - 1ull);
  const int leader = upto ? 63 - __clzll(upto) : 0;
  int base = 0;
  if (lead) {
    const unsigned long long above = lane == 63 ? 0ull : lm & ~((2ull << lane) - 1ull);
    const int next = above ? __ffsll((long long)above) - 1 : nvalid;
    base = atomicAdd(&cell_count[ci], next - lane);
  }
  base = __shfl(base, leader, 64);
  if (valid) slot[i] = base + lane - leader;
}

// Exclusive scan of cell counts in two passes over tiles of kScanTile cells:
// (1) each work-group scans its tile in LDS and stores the tile total;
// (2) each work-group adds the sum of the totals before it.
constexpr int kScanTile = 1024;

__device__ inline int block_exclusive_scan(int v, int *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanTile; off <<= 1) {
    const int a = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  return sh[t] - v;
}

__global__ void __launch_bounds__(kScanTile)
    k_scan_tiles(const int *__restrict__ count, int *__restrict__ start,
                 int *__restrict__ tile_sum, int ncell, const int *__restrict__ flags) {
  if (!build_pending(flags)) return;
  __shared__ int sh[kScanTile];
  const int c = blockIdx.x * kScanTile + threadIdx.x;
  const int v = c < ncell ? count[c] : 0;
  const int ex = block_exclusive_scan(v, sh);
  if (c < ncell) start[c] = ex;
  if (threadIdx.x == kScanTile - 1) tile_sum[blockIdx.x] = ex + v;
}

__global__ void __launch_bounds__(kScanTile)
    k_scan_add(int *__restrict__ start, const int *__restrict__ tile_sum, int ncell, int ntiles,
               const int *__restrict__ flags) {
  if (!build_pending(flags)) return;
  __shared__ int sh[kScanTile];
  // offset of this tile: sum of the totals of the tiles before it
  int s = 0;
  for (int k = threadIdx.x; k < (int)blockIdx.x; k += kScanTile) s += tile_sum[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = kScanTile / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const int off = sh[0];
  const int c = blockIdx.x * kScanTile + threadIdx.x;
  if (c < ncell) start[c] += off;
  if ((int)blockIdx.x == ntiles - 1 && threadIdx.x == 0)
    start[ncell] = off + tile_sum[blockIdx.x];
}

// Cell starts of a build in one pass: each work-group scans a tile of
// kScanTile cells into tile-local starts (local) and stores the tile total;
// the last work-group to finish turns the tile totals into tile offsets (in
// place) and stores the grand total in start[ncell].  A start is then
// local[c] + tile_sum[c / kScanTile] (CellStarts); k_scatter writes them out.
__global__ void __launch_bounds__(kScanTile)
    k_scan_local(const int *__restrict__ count, int *__restrict__ local, int *tile_sum,
                 int *__restrict__ start, int ncell, int *__restrict__ flags) {
  if (!build_pending(flags)) return;
  __shared__ int sh[kScanTile];
  __shared__ int last_sh;
  const int c = blockIdx.x * kScanTile + threadIdx.x;
  const int v = c < ncell ? count[c] : 0;
  const int ex = block_exclusive_scan(v, sh);
  if (c < ncell) local[c] = ex;
  if (threadIdx.x == kScanTile - 1) tile_sum[blockIdx.x] = ex + v;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    last_sh = atomicAdd(&flags[F_BINBLK], 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (!last_sh) return;
  __threadfence();
  const int ntiles = gridDim.x;
  int carry = 0;
  for (int t0 = 0; t0 < ntiles; t0 += kScanTile) {
    const int t = t0 + threadIdx.x;
    const int tv = t < ntiles ? tile_sum[t] : 0;
    __syncthreads();  // sh of the previous chunk has been read
    const int tex = block_exclusive_scan(tv, sh);
    if (t < ntiles) tile_sum[t] = carry + tex;
    carry += sh[kScanTile - 1];
  }
  if (threadIdx.x == 0) {
    start[ncell] = carry;
    flags[F_BINBLK] = 0;
  }
}

// Cell starts of a build from the tile-local starts and tile offsets of
// k_scan_local (start[ncell], the total, already written by it).
struct CellStarts {
  const int *local, *tile_off, *start;
  int ncell;
  __device__ int operator()(int c) const {
    return c < ncell ? local[c] + tile_off[c / kScanTile] : start[ncell];
  }
};

constexpr int kRowWaves = 4;  // x-rows partitioned per work-group (one wavefront each)
constexpr int kMaxRowCells = 1024;

__device__ void task_row_partition(int row, int lane, int *P, int *H,
                                   const int *__restrict__ count, const CellStarts &cs,
                                   const TaskGeom &g, int4 *__restrict__ tpiece,
                                   int4 *__restrict__ tinfo, int *__restrict__ ctask,
                                   int *__restrict__ rowtasks, int *__restrict__ flags,
                                   const int *__restrict__ gstart, int n);

// Scatter atom indices to their cell slots (work-groups [0, nbs)); the
// next nfix work-groups store the cell starts, and the rest (ntb) cut the
// x-rows of the cell grid into force tasks, one wavefront per row.
__global__ void k_scatter(const int *__restrict__ cell, const int *__restrict__ slot,
                          CellStarts cs, int *start, int *__restrict__ perm, int n,
                          int *__restrict__ flags, int nbs, int nfix, TaskGeom g,
                          const int *__restrict__ count, int4 *__restrict__ tpiece,
                          int4 *__restrict__ tinfo, int *__restrict__ ctask,
                          int *__restrict__ rowtasks, const int *__restrict__ gstart) {
  if (!build_pending(flags)) return;
  const int b = blockIdx.x;
  if (b >= nbs + nfix) {
    // exclusive prefix sums over x of one row's 9-row column counts (P) and
    // of its own counts (H), per wavefront
    __shared__ int P[kRowWaves][kMaxRowCells + 1], H[kRowWaves][kMaxRowCells + 1];
    const int wv = (int)(threadIdx.x >> 6);
    task_row_partition((b - nbs - nfix) * kRowWaves + wv, (int)(threadIdx.x & 63), P[wv], H[wv],
                       count, cs, g, tpiece, tinfo, ctask, rowtasks, flags, gstart, n);
    return;
  }
  if (b >= nbs) {
    const int c = (b - nbs) * blockDim.x + threadIdx.x;
    if (c < cs.ncell) start[c] = cs(c);
    return;
  }
  const int i = b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  perm[cs(cell[i]) + slot[i]] = i;
}

// Fixed-point copy of positions for the FP32 pair kernel: (x - lo) * 2^32 / L
// reduced modulo 2^32, so that int32 differences are minimum-image
// separations on periodic axes.
__device__ inline int4 qpack(const double4 &p, const BoxDev &b) {
  return pack_q(p.x, p.y, p.z, (int)p.w, b.lo, b.qscale);
}

// The force tasks of a build in compact order: the used task slots of all
// x-rows, in row order, copied to ctinfo / ctpiece (what the force kernel
// walks: one level of indirection less than a map of slot numbers); their
// number goes to flags[F_NTASK].  tm.parts work-groups of k_rank_move
// (256 threads, the first ones of its grid so that they overlap the atom
// moves), each owning a contiguous share of the rows: the tasks of the rows
// before its share give its base (a strided sum), a scan over its share
// gives each row's first compact index, and all threads then copy the
// share's task records.
struct TaskMapArgs {
  const int *rowtasks;
  const int4 *tinfo, *tpiece;  // by slot
  int4 *ctinfo, *ctpiece;      // compact
  int rows, mpr;               // rows = 0: no force tasks
  int parts;                   // work-groups; shares of at most 256 rows
};

// Work-groups of the compaction for rows x-rows: at least 64, and shares
// of at most 256 rows.
inline int compact_parts(int rows) { return rows > 0 ? std::max(64, (rows + 255) / 256) : 0; }

__device__ void compact_tasks(const TaskMapArgs &tm, int part, int *__restrict__ flags) {
  __shared__ int sh[256];
  __shared__ int red[4];
  const int t = threadIdx.x;
  const int r_lo = (int)((long long)tm.rows * part / tm.parts);
  const int r_hi = (int)((long long)tm.rows * (part + 1) / tm.parts);
  // base: tasks of the rows before the share
  int acc = 0;
  for (int r = t; r < r_lo; r += 256) acc += tm.rowtasks[r];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((t & 63) == 0) red[t >> 6] = acc;
  // inclusive scan of the share's row counts (at most 256 rows)
  const int r = r_lo + t;
  const int v = r < r_hi ? tm.rowtasks[r] : 0;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int a = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  const int base = red[0] + red[1] + red[2] + red[3];
  const int nrow = r_hi - r_lo, ntask_share = nrow > 0 ? sh[nrow - 1] : 0;
  // records of the share: task m (compact index base + m) of row r_lo + w,
  // w the row whose inclusive count first exceeds m (binary search in sh)
  for (int it = t; it < ntask_share * (kTaskPieces + 1); it += 256) {
    const int m = it / (kTaskPieces + 1), p = it % (kTaskPieces + 1);
    int lo = 0, hi = nrow - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (sh[mid] > m) hi = mid; else lo = mid + 1;
    }
    const int k = m - (lo > 0 ? sh[lo - 1] : 0);  // task within its row
    const size_t sl = (size_t)(r_lo + lo) * tm.mpr + k, c = (size_t)base + m;
    if (p == kTaskPieces)
      tm.ctinfo[c] = tm.tinfo[sl];
    else
      tm.ctpiece[c * kTaskPieces + p] = tm.tpiece[sl * kTaskPieces + p];
  }
  if (part == tm.parts - 1 && t == 0) flags[F_NTASK] = base + ntask_share;
}

// Order atoms inside each cell by their position in the previous order so
// that the sorted order, and with it every list and sum, is independent of
// atomic arrival order (atoms that stay in a cell keep their relative
// order).  One thread per atom: its rank in its cell is the number of cell
// members with a smaller previous index (perm lists them; indices are
// unique), and it moves its own data to that place in the *_out arrays
// (acc: only when given), its position to pos_ref (the build positions)
// and (q given) its fixed-point coordinates to q.  The cell counts,
// consumed by the scan, are cleared for the next build.
__global__ void k_rank_move(const int *__restrict__ cell, const int *__restrict__ start,
                            const int *__restrict__ perm, int n, const double4 *__restrict__ pos,
                            double4 *__restrict__ pos_out, double4 *__restrict__ pos_ref,
                            const double *__restrict__ vel, double *__restrict__ vel_out,
                            const double *__restrict__ acc, double *__restrict__ acc_out,
                            const int *__restrict__ tag, int *__restrict__ tag_out,
                            const int *__restrict__ id, int *__restrict__ id_out,
                            int *__restrict__ cell_count, int ncell, int4 *__restrict__ q,
                            BoxDev b, int *__restrict__ flags, TaskMapArgs tm) {
  if (!build_pending(flags)) return;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x)
    cell_count[c] = 0;
  const int cp = tm.parts;  // the first cp work-groups compact the tasks
  if ((int)blockIdx.x < cp) {
    compact_tasks(tm, (int)blockIdx.x, flags);
    return;
  }
  const int i = ((int)blockIdx.x - cp) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell[i];
  const int a = start[c], e = start[c + 1];
  int rank = 0;
  // eight members per round: their loads are independent, so they are
  // issued together rather than one round trip each
  int m = a;
  for (; m + 8 <= e; m += 8) {
    int v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = perm[m + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) rank += v[u] < i;
  }
  for (; m < e; ++m) rank += perm[m] < i;
  const int k = a + rank;
  const double4 p = pos[i];
  pos_out[k] = p;
  pos_ref[k] = p;
  if (q) q[k] = qpack(p, b);
  if (isnan(p.x) | isnan(p.y) | isnan(p.z)) raise_fatal(flags, F_NAN);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    vel_out[d * n + k] = vel[d * n + i];
    if (acc) acc_out[d * n + k] = acc[d * n + i];
  }
  tag_out[k] = tag[i];
  id_out[k] = id[i];
}

// The arrays of the new order, copied back to the primary buffers.
struct CopyBack {
  const double4 *pos_alt;
  double4 *pos;
  const double *vel_alt;
  double *vel;
  const double *acc_alt;  // null: accelerations not carried
  double *acc;
  const int *tag_alt;
  int *tag;
  const int *id_alt;
  int *id;
  int n;
  // open wave builds: the sorted ghosts (device ghost r = atom n + r)
  const double4 *gpos_alt;
  const int *gid_alt, *gtag_alt;
  int *gtag, *gslot;
  int ng;
};

__device__ inline void copy_back_ghost(int r, const CopyBack &c) {
  c.pos[c.n + r] = c.gpos_alt[r];
  c.id[c.n + r] = c.gid_alt[r];
  const int h = c.gtag_alt[r];
  c.gtag[r] = h;
  c.gslot[h] = r;
}

__device__ inline void copy_back_atom(int k, const CopyBack &c) {
  const int n = c.n;
  c.pos[k] = c.pos_alt[k];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c.vel[d * n + k] = c.vel_alt[d * n + k];
    if (c.acc_alt) c.acc[d * n + k] = c.acc_alt[d * n + k];
  }
  c.tag[k] = c.tag_alt[k];
  c.id[k] = c.id_alt[k];
}

__global__ void k_copy_back(CopyBack c, const int *__restrict__ flags) {
  if (!build_pending(flags)) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < c.n)
    copy_back_atom(k, c);
  else if (k - c.n < c.ng)
    copy_back_ghost(k - c.n, c);
}

// Host-halo ghosts keep host order; they get their own CSR cell list and
// their build positions.
__global__ void k_ghost_bin(const double4 *__restrict__ pos, double4 *__restrict__ pos_ref,
                            int n, int ng, BoxDev b, int *__restrict__ cell,
                            int *__restrict__ slot, int *__restrict__ cell_count,
                            const int *__restrict__ flags) {
  if (!build_pending(flags)) return;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const double4 p = pos[n + g];
  pos_ref[n + g] = p;
  const int c = cell_of(p, b);
  cell[g] = c;
  slot[g] = atomicAdd(&cell_count[c], 1);
}

__global__ void k_ghost_scatter(const int *__restrict__ cell, const int *__restrict__ slot,
                                const int *__restrict__ start, int *__restrict__ items, int ng,
                                const int *__restrict__ flags) {
  if (!build_pending(flags)) return;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  items[start[cell[g]] + slot[g]] = g;
}

// Open wave builds: device ghost g goes to place r of the cell order, its
// cell's start plus the number of the cell's ghosts with a smaller device
// index (items lists them in arrival order), so the order does not depend
// on the atomics of the binning; its build position and fixed-point
// coordinates are written in the new order, the rest goes to the *_alt
// arrays (k_copy_back).
__global__ void k_ghost_move(const int *__restrict__ items, const int *__restrict__ gcell,
                             const int *__restrict__ gstart, int n, int ng,
                             const double4 *__restrict__ pos, const int *__restrict__ id,
                             const int *__restrict__ gtag, double4 *__restrict__ gpos_alt,
                             int *__restrict__ gid_alt, int *__restrict__ gtag_alt,
                             double4 *__restrict__ pos_ref, int4 *__restrict__ q, BoxDev b,
                             const int *__restrict__ flags) {
  if (!build_pending(flags)) return;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int c = gcell[g], a = gstart[c], e = gstart[c + 1];
  int rank = 0;
  for (int m = a; m < e; ++m) rank += items[m] < g;
  const int r = a + rank;
  const double4 p = pos[n + g];
  gpos_alt[r] = p;
  pos_ref[n + r] = p;
  q[n + r] = pack_q(p.x, p.y, p.z, (int)p.w, b.lo, b.qscale);
  gid_alt[r] = id[n + g];
  gtag_alt[r] = gtag[g];
}

__global__ void k_ghost_sort_cells(int *__restrict__ items, const int *__restrict__ start,
                                   int ncell, const int *__restrict__ flags) {
  if (!build_pending(flags)) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int a = start[c], e = start[c + 1];
  for (int k = a + 1; k < e; ++k) {
    const int v = items[k];
    int m = k - 1;
    while (m >= a && items[m] > v) {
      items[m + 1] = items[m];
      --m;
    }
    items[m + 1] = v;
  }
}

// Candidate cells along one axis: c-2 .. c+2 (cells are at least half the
// list cutoff wide), wrapped on periodic axes and clipped on open ones;
// every cell once, with offset 0 (no pruning), when a periodic axis has
// fewer than 5 cells.
constexpr int kReach = 2;
constexpr int kSpan = 2 * kReach + 1;

__device__ inline int axis_stencil(int c, int nc, int periodic, int out[kSpan], int off[kSpan]) {
  if (periodic && nc < kSpan) {
    for (int k = 0; k < nc; ++k) {
      out[k] = k;
      off[k] = 0;
    }
    return nc;
  }
  int m = 0;
  for (int d = -kReach; d <= kReach; ++d) {
    int v = c + d;
    if (periodic) {
      v = v < 0 ? v + nc : (v >= nc ? v - nc : v);
    } else if (v < 0 || v >= nc) {
      continue;
    }
    off[m] = d;
    out[m++] = v;
  }
  return m;
}

// Image of candidate coordinate pj relative to pi on a periodic axis, and
// whether the reference would have created that ghost
// (exchange_ghost_single_md_domain.cpp:69-100).
__device__ inline bool axis_image(double pi, double pj, double lo, double hi, double L,
                                  double halfL, double g, int periodic, int &v, double &out) {
  v = 0;
  out = pj;
  if (!periodic) return true;
  const double d = pj - pi;
  if (d > halfL)
    v = 1;
  else if (d < -halfL)
    v = -1;
  if (v == 0) return true;
  const int flag = (pj < lo + g) ? -1 : ((pj > hi - g) ? 1 : 0);
  if (flag != v) return false;
  out = __dsub_rn(pj, (double)v * L);
  return true;
}

// Verlet list: one thread per owned atom, full list (both directions).
//
// List layout (shared with the force kernels): entries in groups of four,
// nbr[((k / 4) * stride + i) * 4 + k % 4], so that a lane reads its next four
// entries as one 16-byte load and a wave's loads are contiguous.  Each lane
// collects four entries in registers before storing a group; the last group
// is padded with the atom itself (r = 0, skipped by the force kernels).
//
// Stencil cells whose box lies farther than the list cutoff from the atom
// are skipped (with a relative margin of 1e-9 so that rounding in the cell
// assignment can never drop a neighbour).  Cells are numbered x-fastest, so
// the surviving cells of one (y, z) stencil row that are adjacent in x form
// one contiguous range of sorted atoms, read kBatch candidates at a time.
// Atoms whose stencil does not wrap a periodic face take a fast path without
// image logic: for them every candidate's minimum image is the direct one.
__device__ inline bool list_pair(double px, double py, double pz, double gx, double gy,
                                 double gz, double ce_sq) {
  const double dx = __dsub_rn(gx, px), dy = __dsub_rn(gy, py), dz = __dsub_rn(gz, pz);
  const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
  return r2 < ce_sq;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct ListWriter {
  uint32_t *nbr;
  size_t stride;  // atoms per group row
  int i, cap, cnt;
  u32x4 g;
  __device__ inline void add(uint32_t e) {
    if (cnt < cap) {
      switch (cnt & 3) {
        case 0: g.x = e; break;
        case 1: g.y = e; break;
        case 2: g.z = e; break;
        default:
          g.w = e;
          reinterpret_cast<u32x4 *>(nbr)[(size_t)(cnt >> 2) * stride + i] = g;
      }
    }
    ++cnt;
  }
  __device__ inline void finish() {
    const int kept = cnt < cap ? cnt : cap;
    if (kept & 3) {
      const uint32_t self = (uint32_t)i | (kNoImage << 26);
      if ((kept & 3) < 2) g.y = self;
      if ((kept & 3) < 3) g.z = self;
      g.w = self;
      reinterpret_cast<u32x4 *>(nbr)[(size_t)(kept >> 2) * stride + i] = g;
    }
  }
};

constexpr int kBatch = 8;

template <bool IMAGES>
__device__ inline void scan_range(const double4 *__restrict__ pos, int j0, int j1, int i,
                                  const double4 &pi, const BoxDev &b, ListWriter &out) {
  for (int j = j0; j < j1; j += kBatch) {
    double4 pj[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) pj[u] = pos[min(j + u, j1 - 1)];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int jj = j + u;
      if (jj >= j1 || jj == i) continue;
      if (IMAGES) {
        int vx, vy, vz;
        double gx, gy, gz;
        if (!axis_image(pi.x, pj[u].x, b.lo[0], b.hi[0], b.L[0], b.halfL[0], b.ghost_cut, b.bc[0], vx, gx)) continue;
        if (!axis_image(pi.y, pj[u].y, b.lo[1], b.hi[1], b.L[1], b.halfL[1], b.ghost_cut, b.bc[1], vy, gy)) continue;
        if (!axis_image(pi.z, pj[u].z, b.lo[2], b.hi[2], b.L[2], b.halfL[2], b.ghost_cut, b.bc[2], vz, gz)) continue;
        if (!list_pair(pi.x, pi.y, pi.z, gx, gy, gz, b.ce_sq)) continue;
        out.add((uint32_t)jj | ((uint32_t)(vx + 1) << 26) | ((uint32_t)(vy + 1) << 28) |
                ((uint32_t)(vz + 1) << 30));
      } else {
        if (!list_pair(pi.x, pi.y, pi.z, pj[u].x, pj[u].y, pj[u].z, b.ce_sq)) continue;
        out.add((uint32_t)jj | (kNoImage << 26));
      }
    }
  }
}

__device__ inline void block_stat(int2 *bstat, int cnt, int kept);
__device__ void nlist_atom(const double4 *__restrict__ pos, int n, const BoxDev &b,
                           const int *__restrict__ cstart, const int *__restrict__ gstart,
                           const int *__restrict__ gitems, uint32_t *__restrict__ nbr,
                           int *__restrict__ nnbr, int stride, int cap, int host_halo, int i,
                           int &cnt_out, int &kept_out);

__global__ void __launch_bounds__(256)
    k_nlist(const double4 *__restrict__ pos, int n, BoxDev b, const int *__restrict__ cstart,
            const int *__restrict__ gstart, const int *__restrict__ gitems,
            uint32_t *__restrict__ nbr, int *__restrict__ nnbr, int stride, int cap,
            const int *__restrict__ flags, int host_halo, int2 *__restrict__ wstat) {
  if (!build_pending(flags)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int cnt = 0, kept = 0;
  if (i < n) nlist_atom(pos, n, b, cstart, gstart, gitems, nbr, nnbr, stride, cap, host_halo, i, cnt, kept);
  block_stat(wstat, cnt, kept);
}

// Per-block build statistics (longest list, entries kept), reduced by
// k_build_stats: no same-address atomics from thousands of blocks.
__device__ inline void block_stat(int2 *bstat, int cnt, int kept) {
  __shared__ int2 sh[4];
  for (int o = 32; o > 0; o >>= 1) {
    cnt = max(cnt, __shfl_xor(cnt, o, 64));
    kept += __shfl_xor(kept, o, 64);
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = make_int2(cnt, kept);
  __syncthreads();
  if (threadIdx.x == 0) {
    int2 t = sh[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) t = make_int2(max(t.x, sh[w].x), t.y + sh[w].y);
    bstat[blockIdx.x] = t;
  }
}

__device__ void nlist_atom(const double4 *__restrict__ pos, int n, const BoxDev &b,
                           const int *__restrict__ cstart, const int *__restrict__ gstart,
                           const int *__restrict__ gitems, uint32_t *__restrict__ nbr,
                           int *__restrict__ nnbr, int stride, int cap, int host_halo, int i,
                           int &cnt_out, int &kept_out) {
  const double4 pi = pos[i];
  const double pc[3] = {pi.x, pi.y, pi.z};
  int cc[3], sten[3][kSpan], nst[3];
  double gap[3][kSpan];  // distance from the atom to stencil cell k along axis d
  bool interior = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    cc[d] = axis_cell(pc[d], b.glo[d], b.cinv[d], b.nc[d]);
    int soff[kSpan];
    nst[d] = axis_stencil(cc[d], b.nc[d], b.bc[d], sten[d], soff);
    const double w = 1.0 / b.cinv[d];
    const double off = pc[d] - (b.glo[d] + cc[d] * w);  // position inside the cell
    for (int k = 0; k < nst[d]; ++k) {
      const int o = soff[k];
      double g = o < 0 ? (-o - 1) * w + off : (o > 0 ? (o - 1) * w + (w - off) : 0.0);
      gap[d][k] = g < 0.0 ? 0.0 : g;
    }
    if (b.bc[d] && (cc[d] < kReach || cc[d] > b.nc[d] - 1 - kReach)) interior = false;
  }
  const double prune = b.ce_sq * (1.0 + 1e-9);
  ListWriter out{nbr, (size_t)stride, i, cap, 0, u32x4{0u, 0u, 0u, 0u}};
  for (int a = 0; a < nst[2]; ++a)
    for (int bb = 0; bb < nst[1]; ++bb) {
      const double gyz = gap[2][a] * gap[2][a] + gap[1][bb] * gap[1][bb];
      if (gyz > prune) continue;
      const int row = b.nc[0] * (sten[1][bb] + b.nc[1] * sten[2][a]);
      // merge the surviving x cells of this row into contiguous ranges
      int c = 0;
      while (c < nst[0]) {
        if (gyz + gap[0][c] * gap[0][c] > prune) {
          ++c;
          continue;
        }
        int x0 = sten[0][c], x1 = x0;
        int c2 = c + 1;
        while (c2 < nst[0] && sten[0][c2] == x1 + 1 && gyz + gap[0][c2] * gap[0][c2] <= prune) {
          ++x1;
          ++c2;
        }
        const int j0 = cstart[row + x0], j1 = cstart[row + x1 + 1];
        if (interior)
          scan_range<false>(pos, j0, j1, i, pi, b, out);
        else
          scan_range<true>(pos, j0, j1, i, pi, b, out);
        if (host_halo) {
          for (int x = x0; x <= x1; ++x) {
            const int cell = row + x;
            const int g1 = gstart[cell + 1];
            for (int t = gstart[cell]; t < g1; ++t) {
              const int j = n + gitems[t];
              const double4 pj = pos[j];
              if (list_pair(pi.x, pi.y, pi.z, pj.x, pj.y, pj.z, b.ce_sq)) out.add((uint32_t)j | (kNoImage << 26));
            }
          }
        }
        c = c2;
      }
    }
  out.finish();
  const int kept = out.cnt < cap ? out.cnt : cap;
  nnbr[i] = kept;
  cnt_out = out.cnt;
  kept_out = kept;
}

// Verlet list for fully periodic boxes with at least 5 cells per axis, each
// at least the list cutoff wide (BoxDev::coarse): one wavefront per home
// cell.
//
// The 27 stencil cells form 9 (dy, dz) rows; the x-adjacent cells
// cx-1 .. cx+1 of a row are one contiguous range of sorted atoms (two when
// the row wraps in x).  Slot k of the wave holds up to 64 atoms of range k,
// one per lane, as coordinates relative to the first home atom (from the
// 21-bit coarse positions, so the periodic image is implicit).  The wave
// then takes the home atoms one at a time: every lane tests its candidates
// against the home atom, and the accepted entries are appended to the home
// atom's list in candidate order (ballot + prefix count), coalesced within
// each 4-entry group.  Ranges longer than 64 atoms continue in extra passes.
//
// A candidate is accepted without further work when its coarse squared
// distance is below cthr_in and rejected when it is at or above cthr_out;
// in between, the FP64 predicate of the general build decides (with the
// reference's ghost-existence rule for images when the ghost cutoff is
// below the list cutoff).
// Coordinate relative to a reference atom, from the 32-bit fixed-point
// coordinates (the int32 difference is the minimum image on periodic axes).
__device__ inline float q_rel(int q, int o, float unit) {
  return (float)(int)((unsigned)q - (unsigned)o) * unit;
}

// Coordinates relative to the reference atom po from the FP64 positions, for
// boxes with 3 or 4 cells on some axis: there a candidate can lie more than
// half a box from the reference atom, beyond what an int32 difference of q
// resolves, so the separation of the candidate's image v (position p - v L)
// is formed in FP64 and rounded once.  Home atoms take the same path, so an
// atom's own coordinates are identical as candidate and as home atom.
__device__ inline float3 pos_rel(const double4 *__restrict__ pos, int j, uint32_t img,
                                 const double4 &po, const BoxDev &B) {
  const double *pp = reinterpret_cast<const double *>(pos + j);
  const double x = pp[0], y = pp[1], z = pp[2];  // w is not needed
  const int vx = (int)((img >> 26) & 3u) - 1, vy = (int)((img >> 28) & 3u) - 1,
            vz = (int)(img >> 30) - 1;
  return make_float3((float)((vx ? x - vx * B.L[0] : x) - po.x),
                     (float)((vy ? y - vy * B.L[1] : y) - po.y),
                     (float)((vz ? z - vz * B.L[2] : z) - po.z));
}

// x, y, z of q[j] only (w is not needed, and a loaded but unused w would
// leave a load outstanding for the next reuse of its register to wait for)
__device__ inline int3 load_q3(const int4 *q, int j) {
  const int *p = reinterpret_cast<const int *>(q + j);
  return make_int3(p[0], p[1], p[2]);
}

__device__ inline uint32_t img_bits(int vx, int vy, int vz) {
  return ((uint32_t)(vx + 1) << 26) | ((uint32_t)(vy + 1) << 28) | ((uint32_t)(vz + 1) << 30);
}

constexpr int kWaveRows = 9;
constexpr int kMaxHomeCell = 64;  // home atoms per pass (LDS counts)

// What the main path of k_nlist_cells needs of BoxDev (kernel arguments);
// the rare FP64 path reads the device copy of BoxDev instead, so that its
// constants do not occupy registers for the whole kernel.
struct CellWaveBox {
  int nc[3];
  float cunit[3];
  float cthr_in, cthr_out;
  int wide;  // some axis has 3 or 4 cells: relative coordinates from the FP64 positions
  int open;  // BoxDev::open: rows clipped, part 1 = the ghosts of the part-0 cells
  int n;     // owned atoms (first ghost index)
};

// Range of sorted atoms of stencil row r (dy = r % 3 - 1, dz = r / 3 - 1) of
// home cell (cx, cy, cz).  part 0: the x-cells that do not wrap; part 1: the
// wrapped one (cx == 0: cell nx-1 seen with image +1; cx == nx-1: cell 0
// with image -1), empty otherwise.
// Open grids: part 0 is the owned atoms of x-cells cx-1 .. cx+1 (clipped)
// of the row, part 1 the ghosts of the same cells.
__device__ inline void row_range(const int *__restrict__ cstart, const int *__restrict__ gstart,
                                 const CellWaveBox &b, int cx, int cy, int cz, int r, int part,
                                 int &j0, int &len, uint32_t &img) {
  const int nx = b.nc[0], ny = b.nc[1], nz = b.nc[2];
  int y = cy + r % 3 - 1, z = cz + r / 3 - 1, vy = 0, vz = 0;
  if (b.open) {
    j0 = 0;
    len = 0;
    img = img_bits(0, 0, 0);
    if (y < 0 || y >= ny || z < 0 || z >= nz) return;
    const int row = nx * (y + ny * z);
    const int a0 = row + (cx > 0 ? cx - 1 : 0), a1 = row + (cx + 2 < nx ? cx + 2 : nx);
    if (part == 0) {
      j0 = cstart[a0];
      len = cstart[a1] - j0;
    } else if (gstart) {
      j0 = b.n + gstart[a0];
      len = gstart[a1] - gstart[a0];
    }
    return;
  }
  if (y < 0) { y += ny; vy = 1; } else if (y >= ny) { y -= ny; vy = -1; }
  if (z < 0) { z += nz; vz = 1; } else if (z >= nz) { z -= nz; vz = -1; }
  const int row = nx * (y + ny * z);
  const bool xlo = cx == 0, xhi = cx == nx - 1;
  if (part == 0) {
    const int a0 = row + (xlo ? 0 : cx - 1), a1 = row + (xhi ? nx : cx + 2);
    j0 = cstart[a0];
    len = cstart[a1] - j0;
    img = img_bits(0, vy, vz);
  } else if (xlo || xhi) {
    const int b0 = xlo ? row + nx - 1 : row;
    j0 = cstart[b0];
    len = cstart[b0 + 1] - j0;
    img = img_bits(xlo ? 1 : -1, vy, vz);
  } else {
    j0 = 0;
    len = 0;
    img = 0u;
  }
}

constexpr int kSlots = 9;                  // candidates per pass: kSlots * 64
constexpr int kPassCand = kSlots * 64;
constexpr int kSegs = 2 * kWaveRows;         // row ranges: 9 direct-x + 9 wrapped-x

// One pass over the flattened candidates [c0, c0 + kPassCand) of the home
// cell against home atoms [h0, h1).  cand (LDS) holds the pass's
// candidate entries; STAGED (single-pass cells): each home atom's accepted
// entries are collected in buf (LDS) and written as whole 4-entry groups;
// otherwise they are stored one by one and counts carried in cnt_io.
// SLOTS: the list takes the candidates' task slots (cslot) instead of their
// entries, in 16-bit groups of 8, padded with pad_slot (the task's sentinel
// slot, staged far from every home atom).
template <bool STAGED, bool GHOST_OK, bool SLOTS>
__device__ void cell_wave_pass(const double4 *__restrict__ pos, const int4 *__restrict__ q,
                               const CellWaveBox &b, const BoxDev *__restrict__ bdev,
                               const uint32_t *cand, const uint16_t *cslot, int ncand, int c0,
                               int self_base, int self_j0, int pad_slot, int h0, int h1,
                               const int3 o, const double4 &po, uint32_t *__restrict__ nbr,
                               int *__restrict__ nnbr,
                               size_t stride, int cap, uint32_t *buf, int *cnt_io, float hxl,
                               float hyl, float hzl, int lane, bool first, int &kept_l, int &wmax,
                               int &went) {
  float px[kSlots], py[kSlots], pz[kSlots];
  float tin[GHOST_OK ? 1 : kSlots];  // per-slot inner threshold when images may lack ghosts
  // the value each candidate appends: its entry, or (SLOTS) its task slot,
  // two 16-bit slots to a register (slot k in half k % 2), the entry then
  // being re-read from cand on the rare exact path
  constexpr int kEntRegs = SLOTS ? (kSlots + 1) / 2 : kSlots;
  uint32_t ent[kEntRegs];
#pragma unroll
  for (int k = 0; k < kEntRegs; ++k) ent[k] = 0u;
  auto ent_of = [&](int k) -> uint32_t {
    return SLOTS ? ((k & 1) ? ent[k >> 1] >> 16 : ent[k >> 1] & 0xffffu) : ent[k];
  };
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    const int t = k * 64 + lane;
    px[k] = 1e30f;  // empty slot: never accepted, never in the band
    py[k] = 0.0f;
    pz[k] = 0.0f;
    if (!GHOST_OK) tin[k] = b.cthr_in;
    if (t < ncand) {
      const uint32_t e = cand[t];
      const int3 c = load_q3(q, e & kIndexMask);
      px[k] = q_rel(c.x, o.x, b.cunit[0]);
      py[k] = q_rel(c.y, o.y, b.cunit[1]);
      pz[k] = q_rel(c.z, o.z, b.cunit[2]);
      if (b.wide) {  // uniform
        const float3 r = pos_rel(pos, e & kIndexMask, e, po, *bdev);
        px[k] = r.x;
        py[k] = r.y;
        pz[k] = r.z;
      }
      // images whose ghost may not exist always take the exact path
      if (!GHOST_OK && (e >> 26) != kNoImage) tin[k] = -1.0f;
      if (SLOTS)
        ent[k >> 1] |= (uint32_t)cslot[t] << ((k & 1) * 16);
      else
        ent[k] = e;
    }
  }
  uint16_t *buf16 = reinterpret_cast<uint16_t *>(buf);
  uint16_t *nbr16 = reinterpret_cast<uint16_t *>(nbr);
  const uint32_t cin_m1 = __float_as_uint(b.cthr_in) - 1u;
  const float bmid = 0.5f * (b.cthr_in + b.cthr_out), bhw = b.cthr_out - b.cthr_in;
  // lane-private scratch word for lanes that accept nothing (STAGED)
  const int dummy_at = SLOTS ? kPassCand + 8 + lane : kPassCand + 4 + lane;
  for (int h = h0; h < h1; ++h) {
    // home atom coordinates: lane (h - h0) holds them
    const float hx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hxl), h - h0));
    const float hy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hyl), h - h0));
    const float hz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hzl), h - h0));
    // position of the home atom itself among the candidates
    const int sp = self_base + (h - self_j0) - c0;
    const int kself = sp >= 0 && sp < kPassCand ? sp >> 6 : -1;
    const unsigned long long selfbit = 1ull << (sp & 63);
    int cnt = (STAGED || first) ? 0 : cnt_io[h - h0];
    // Accepted entries go to the list in candidate order: a lane's entry
    // position is cnt + (accepted lanes below it).  The accept decisions
    // are 64-bit lane masks (scalar registers); lanes that accept nothing
    // write a scratch word, so the common path has no branches.
    auto append = [&](unsigned long long m, uint32_t e) {
      int at = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                              __builtin_amdgcn_mbcnt_lo((unsigned)m, (unsigned)cnt));
      const bool take = __builtin_amdgcn_inverse_ballot_w64(m);
      if (STAGED) {
        __asm__ volatile("" : "+v"(at));  // computed by every lane: a select, not a branch
        if (SLOTS)
          buf16[take ? at : dummy_at] = (uint16_t)e;
        else
          buf[take ? at : dummy_at] = e;
      } else if (take && at < cap) {
        if (SLOTS)
          nbr16[((size_t)(at >> 3) * stride + h) * 8 + (at & 7)] = (uint16_t)e;
        else
          nbr[((size_t)(at >> 2) * stride + h) * 4 + (at & 3)] = e;
      }
      cnt += __popcll(m);
    };
    unsigned long long anyband = 0ull;
    // GHOST_OK: per-lane pre-test of the band, min |r2 - mid| against its
    // full width (conservative: the band path re-tests exactly)
    float bm = 1e30f;
#pragma unroll
    for (int g = 0; g < kSlots; g += 3) {
      if (g * 64 >= ncand) continue;  // uniform; keeps the loops unrolled
#pragma unroll
      for (int k = g; k < (g + 3 < kSlots ? g + 3 : kSlots); ++k) {
        const float dx = px[k] - hx, dy = py[k] - hy, dz = pz[k] - hz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (GHOST_OK) {
          // 0 < r2 < cthr_in as one unsigned compare of the float bits
          // (r2 >= +0): the home atom itself (r2 = 0) fails it.  Vector
          // instructions only; the scalar unit is shared by four SIMDs.
          const unsigned long long m_in = __builtin_amdgcn_ballot_w64((__float_as_uint(r2) - 1u) < cin_m1);
          bm = fminf(bm, fabsf(r2 - bmid));
          append(m_in, ent_of(k));
        } else {
          const unsigned long long m_in = __builtin_amdgcn_ballot_w64(r2 < tin[GHOST_OK ? 0 : k]);
          const unsigned long long m_out = __builtin_amdgcn_ballot_w64(r2 < b.cthr_out);
          anyband |= m_out & ~m_in;  // the home atom itself is in m_in
          append(k == kself ? m_in & ~selfbit : m_in, ent_of(k));
        }
      }
    }
    if (GHOST_OK) anyband = __builtin_amdgcn_ballot_w64(bm < bhw);
    // Candidates in the band between the coarse thresholds (rare): the FP64
    // predicate decides, with the ghost-existence rule for images.  Kept
    // out of the loop above so that its global loads never make the common
    // path wait for outstanding list stores.
    if (__builtin_expect(anyband != 0ull, 0)) {
      const BoxDev &B = *bdev;
      const double *pq = reinterpret_cast<const double *>(pos + h);
      const double pix = pq[0], piy = pq[1], piz = pq[2];
#pragma unroll
      for (int k = 0; k < kSlots; ++k) {
        if (k * 64 >= ncand) continue;
        const float dx = px[k] - hx, dy = py[k] - hy, dz = pz[k] - hz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const bool band = (r2 < b.cthr_out) & !(r2 < (GHOST_OK ? b.cthr_in : tin[GHOST_OK ? 0 : k]));
        bool take = false;
        if (band) {
          uint32_t e = SLOTS ? cand[k * 64 + lane] : ent_of(k);
          __asm__ volatile("" : "+v"(e));  // decode here, not hoisted
          // x, y, z only: a loaded but unused w would leave a load pending
          // that the next reuse of its register has to wait for
          const double *pp = reinterpret_cast<const double *>(pos + (e & kIndexMask));
          const int v[3] = {(int)((e >> 26) & 3u) - 1, (int)((e >> 28) & 3u) - 1, (int)(e >> 30) - 1};
          const double pjc[3] = {pp[0], pp[1], pp[2]};
          // selects, not branches: every loaded value is used on every path,
          // so no load is left outstanding past this block
          double gc[3];
          bool exists = true;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const int flag = (pjc[d] < B.lo[d] + B.ghost_cut) ? -1 : ((pjc[d] > B.hi[d] - B.ghost_cut) ? 1 : 0);
            exists = exists & ((v[d] == 0) | (flag == v[d]));
            gc[d] = v[d] != 0 ? __dsub_rn(pjc[d], (double)v[d] * B.L[d]) : pjc[d];
          }
          take = list_pair(pix, piy, piz, gc[0], gc[1], gc[2], B.ce_sq) & exists;
        }
        append(__builtin_amdgcn_ballot_w64(take), ent_of(k));
      }
    }
    if (STAGED) {
      // whole groups (16 bytes), the last one padded with the atom itself
      const int kept = min(cnt, cap);
      if (SLOTS) {
        if (lane < 7) buf16[kept + lane] = (uint16_t)pad_slot;
        for (int g = lane; g * 8 < kept; g += 64)
          reinterpret_cast<u32x4 *>(nbr)[(size_t)g * stride + h] = reinterpret_cast<const u32x4 *>(buf)[g];
      } else {
        if (lane < 3) buf[kept + lane] = (uint32_t)h | (kNoImage << 26);
        for (int g = lane; g * 4 < kept; g += 64)
          reinterpret_cast<u32x4 *>(nbr)[(size_t)g * stride + h] = reinterpret_cast<const u32x4 *>(buf)[g];
      }
      kept_l = lane == h - h0 ? kept : kept_l;  // stored per chunk
      wmax = max(wmax, cnt);
      went += kept;
    } else if (lane == 0) {
      cnt_io[h - h0] = cnt;
    }
  }
}

// One home cell c (wavefront wv of its block); returns (longest list,
// entries kept).
// Slot offset (slot - atom index) of the atoms of segment (r, part) of home
// cell cx in its task's staging order: the piece of row r that holds the
// segment's first x-cell.
__device__ inline int seg_slot_delta(const int4 *__restrict__ tp, int nx, int cx, int r, int part,
                                     int open) {
  if (open) {  // one piece per part, holding the segment's cells
    const int4 a = tp[2 * r + part];
    return a.y - a.x;
  }
  const int xs = part == 0 ? (cx == 0 ? 0 : cx - 1) : (cx == 0 ? nx - 1 : 0);
  const int4 a = tp[2 * r];
  const int4 pc = (xs >= a.z && xs < a.w) ? a : tp[2 * r + 1];
  return pc.y - pc.x;
}

template <bool GHOST_OK, bool SLOTS>
__device__ int2 cell_wave(const double4 *__restrict__ pos, const int4 *__restrict__ q,
                          const CellWaveBox &b, const BoxDev *__restrict__ bdev,
                          const int *__restrict__ cstart, const int *__restrict__ gstart,
                          uint32_t *__restrict__ nbr,
                          int *__restrict__ nnbr, int stride, int cap, int c, int wv, int lane,
                          uint32_t (*lds)[kPassCand + 4 + 64], uint16_t *cslot,
                          int (*cnt_sh)[kMaxHomeCell], const int *__restrict__ ctask,
                          const int4 *__restrict__ tpiece, const int4 *__restrict__ tinfo) {
  const int nx = b.nc[0], ny = b.nc[1];
  const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
  const int h0 = cstart[c], h1 = cstart[c + 1];
  if (h0 == h1) return make_int2(0, 0);
  uint32_t *cand = lds[wv];
  // SLOTS: this cell's task, its piece table and sentinel slot
  const int tsl = SLOTS ? ctask[c] : 0;
  const int4 *tp = SLOTS ? tpiece + (size_t)tsl * kTaskPieces : nullptr;
  const int pad_slot = SLOTS ? tinfo[tsl].z : 0;
  // segment table, one row range per lane (lanes < kSegs): first entry,
  // length, image bits, (SLOTS) slot offset and flattened start.  The
  // loads of all ranges are issued together as vector loads; the passes
  // read the table back with readlane.
  int sj0 = 0, slen = 0, sdelta = 0;
  uint32_t simg = 0u;
  if (lane < kSegs) {
    const int r = lane < kWaveRows ? lane : lane - kWaveRows, part = lane < kWaveRows ? 0 : 1;
    row_range(cstart, gstart, b, cx, cy, cz, r, part, sj0, slen, simg);
    if (SLOTS && slen > 0) sdelta = seg_slot_delta(tp, nx, cx, r, part, b.open);
  }
  int sat = slen;  // inclusive prefix sum over lanes (kSegs <= 32)
#pragma unroll
  for (int d = 1; d < 32; d <<= 1) {
    const int v = __shfl_up(sat, d);
    sat += lane >= d ? v : 0;
  }
  const int total = __builtin_amdgcn_readlane(sat, kSegs - 1);
  sat -= slen;
  const int self_base = __builtin_amdgcn_readlane(sat, kWaveRows / 2);  // dy = dz = 0, direct x
  const int self_j0 = __builtin_amdgcn_readlane(sj0, kWaveRows / 2);
  const int3 o = load_q3(q, h0);  // coordinates are relative to the first home atom
  double4 po = make_double4(0.0, 0.0, 0.0, 0.0);
  if (b.wide) {
    const double *pp = reinterpret_cast<const double *>(pos + h0);
    po = make_double4(pp[0], pp[1], pp[2], 0.0);
  }
  int wmax = 0, went = 0;
  const bool staged = total <= kPassCand;
  for (int h = h0; h < h1; h += kMaxHomeCell) {
    const int he = min(h1, h + kMaxHomeCell);
    // lane l holds home atom h + l (relative coordinates) and its kept count
    float hxl = 0.0f, hyl = 0.0f, hzl = 0.0f;
    int kept_l = 0;
    if (h + lane < he) {
      const int3 c = load_q3(q, h + lane);
      hxl = q_rel(c.x, o.x, b.cunit[0]);
      hyl = q_rel(c.y, o.y, b.cunit[1]);
      hzl = q_rel(c.z, o.z, b.cunit[2]);
      if (b.wide) {
        const float3 r = pos_rel(pos, h + lane, kNoImage << 26, po, *bdev);
        hxl = r.x;
        hyl = r.y;
        hzl = r.z;
      }
    }
    for (int c0 = 0; c0 < total; c0 += kPassCand) {
      // gather this pass's candidate entries into LDS
      for (int sg = 0; sg < kSegs; ++sg) {
        const int len = __builtin_amdgcn_readlane(slen, sg);
        const int at = __builtin_amdgcn_readlane(sat, sg);
        if (len == 0 || at + len <= c0 || at >= c0 + kPassCand) continue;
        const int j0 = __builtin_amdgcn_readlane(sj0, sg);
        const uint32_t img = (uint32_t)__builtin_amdgcn_readlane((int)simg, sg);
        const int delta = SLOTS ? __builtin_amdgcn_readlane(sdelta, sg) : 0;
        for (int t = lane; t < len; t += 64) {
          const int p = at + t - c0;
          if (p >= 0 && p < kPassCand) {
            cand[p] = (uint32_t)(j0 + t) | img;
            if (SLOTS) cslot[p] = (uint16_t)(j0 + t + delta);
          }
        }
      }
      const int ncand = min(kPassCand, total - c0);
      if (staged)
        cell_wave_pass<true, GHOST_OK, SLOTS>(pos, q, b, bdev, cand, cslot, ncand, c0, self_base,
                                              self_j0, pad_slot, h, he, o, po, nbr, nnbr,
                                              (size_t)stride, cap,
                                              SLOTS ? reinterpret_cast<uint32_t *>(cslot) : cand,
                                              cnt_sh[wv], hxl, hyl, hzl, lane, c0 == 0, kept_l,
                                              wmax, went);
      else
        cell_wave_pass<false, GHOST_OK, SLOTS>(pos, q, b, bdev, cand, cslot, ncand, c0, self_base,
                                               self_j0, pad_slot, h, he, o, po, nbr, nnbr,
                                               (size_t)stride, cap, nullptr, cnt_sh[wv], hxl,
                                               hyl, hzl, lane, c0 == 0, kept_l, wmax, went);
    }
    if (staged && h + lane < he) nnbr[h + lane] = kept_l;
    if (!staged) {
      // pad the last group with the atom itself, store lengths
      for (int hh = h; hh < he; ++hh) {
        const int cnt = cnt_sh[wv][hh - h];
        const int kept = min(cnt, cap);
        const int padded = SLOTS ? (kept + 7) & ~7 : (kept + 3) & ~3;
        if (SLOTS && lane < padded - kept)
          reinterpret_cast<uint16_t *>(nbr)[((size_t)((kept + lane) >> 3) * stride + hh) * 8 + ((kept + lane) & 7)] =
              (uint16_t)pad_slot;
        if (!SLOTS && lane < padded - kept)
          nbr[((size_t)((kept + lane) >> 2) * stride + hh) * 4 + ((kept + lane) & 3)] =
              (uint32_t)hh | (kNoImage << 26);
        if (lane == 0) nnbr[hh] = kept;
        wmax = max(wmax, cnt);
        went += kept;
      }
    }
  }
  return make_int2(wmax, went);
}

template <bool GHOST_OK, bool SLOTS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
    k_nlist_cells(const double4 *__restrict__ pos, const int4 *__restrict__ q, CellWaveBox b,
                  const BoxDev *__restrict__ bdev, const int *__restrict__ cstart,
                  const int *__restrict__ gstart, uint32_t *__restrict__ nbr, int *__restrict__ nnbr, int stride, int cap,
                  const int *__restrict__ flags, int nblk_real, int2 *__restrict__ bstat,
                  const int *__restrict__ ctask, const int4 *__restrict__ tpiece,
                  const int4 *__restrict__ tinfo, int ncopy, CopyBack cpb) {
  if (!build_pending(flags)) return;
  const int lgrid = (int)gridDim.x - ncopy;  // list work-groups, a multiple of 8
  if ((int)blockIdx.x >= lgrid) {
    // the last ncopy work-groups copy the new order back to the primary
    // buffers (this kernel reads only q and the new-order positions)
    for (int k = ((int)blockIdx.x - lgrid) * blockDim.x + threadIdx.x; k < cpb.n;
         k += ncopy * blockDim.x)
      copy_back_atom(k, cpb);
    return;
  }
  // consecutive cells (which share most candidates) on the same XCD
  const int per = lgrid >> 3;
  const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (blk >= nblk_real) return;
  // candidates of a pass, then (staged) one home atom's entries + 3 pads;
  // 64 scratch words
  __shared__ __attribute__((aligned(16))) uint32_t lds[4][kPassCand + 4 + 64];
  __shared__ int cnt_sh[4][kMaxHomeCell];
  __shared__ int2 st_sh[4];
  // candidates' task slots; then (staged) one home atom's slots + 7 pads,
  // 64 scratch entries
  __shared__ __attribute__((aligned(16))) uint16_t cslot_sh[4][SLOTS ? kPassCand + 8 + 64 : 8];
  const int wv = (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // the home cell is uniform across the wavefront
  const int c = __builtin_amdgcn_readfirstlane(blk * 4 + wv);
  int2 st = make_int2(0, 0);
  if (c < b.nc[0] * b.nc[1] * b.nc[2])
    st = cell_wave<GHOST_OK, SLOTS>(pos, q, b, bdev, cstart, gstart, nbr, nnbr, stride, cap, c, wv,
                                    lane, lds, cslot_sh[wv], cnt_sh, ctask, tpiece, tinfo);
  if (lane == 0) st_sh[wv] = st;
  __syncthreads();
  if (threadIdx.x == 0) {  // per-block statistics, reduced by k_build_stats
    int2 t = st_sh[0];
    for (int w = 1; w < 4; ++w) t = make_int2(max(t.x, st_sh[w].x), t.y + st_sh[w].y);
    bstat[blk] = t;
  }
}

// Close a build: reduce the per-block statistics of the list kernel (a few
// blocks, one atomic each into flags[F_MAXNBR] / flags[F_ENTRIES], which
// k_wrap_bin cleared); the last block to finish checks the capacity and
// clears the request.
__global__ void __launch_bounds__(256)
    k_build_stats(int *__restrict__ flags, const int2 *__restrict__ wstat, int nw, int cap,
                  int step) {
  // the step index of the run, kept while the run goes on (the last kernel
  // of the build chain runs on every step)
  if (blockIdx.x == 0 && threadIdx.x == 0 && flags[F_HALT] == 0) flags[F_STEP] = step;
  if (!build_pending(flags)) return;
  __shared__ int smax[4], ssum[4];
  int m = 0, t = 0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < nw; k += gridDim.x * 256) {
    const int2 v = wstat[k];
    m = max(m, v.x);
    t += v.y;
  }
  for (int o = 32; o > 0; o >>= 1) {
    m = max(m, __shfl_xor(m, o, 64));
    t += __shfl_xor(t, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    smax[threadIdx.x >> 6] = m;
    ssum[threadIdx.x >> 6] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    m = max(max(smax[0], smax[1]), max(smax[2], smax[3]));
    t = ssum[0] + ssum[1] + ssum[2] + ssum[3];
    if (m > 0) atomicMax(&flags[F_MAXNBR], m);
    if (t > 0) atomicAdd(&flags[F_ENTRIES], t);
    __threadfence();
    if (atomicAdd(&flags[F_STATBLK], 1) == (int)gridDim.x - 1) {
      // every block's atomics are done
      flags[F_STATBLK] = 0;
      if (atomicAdd(&flags[F_MAXNBR], 0) > cap) {
        // some list was cut at its capacity: halt before any force uses it,
        // the request stays set so that the host can grow the lists and
        // run the build again
        flags[F_OVERFLOW] = 1;
        flags[F_HALT] = 1;
      } else {
        flags[F_REBUILD] = 0;
        flags[F_REBUILDS] += 1;
      }
    }
  }
}

// Force tasks of x-row `row` (one wavefront; k_scatter, after the cell
// starts of a build are known).  P and H (LDS, nx + 1 entries each) receive
// the exclusive prefix sums over x of the row's 9-row column counts and of
// its own counts; from them the home atoms and the staged neighbourhood of
// any run of cells [x0, x1) are O(1).  The row is cut greedily: each task
// takes the longest run from x0 that keeps at most kTaskBlock home atoms and
// a neighbourhood (plus the sentinel) within kStage, at least one cell; the
// last slot of the row takes whatever remains.  Each task slot gets its
// piece table (first atom, slot base, x-cells of every piece) and tinfo
// (home atoms [h0, h1), staged count S, x0 | x1 << 16), each cell its slot.
__device__ void task_row_partition(int row, int lane, int *P, int *H,
                                   const int *__restrict__ count, const CellStarts &cs,
                                   const TaskGeom &g, int4 *__restrict__ tpiece,
                                   int4 *__restrict__ tinfo, int *__restrict__ ctask,
                                   int *__restrict__ rowtasks, int *__restrict__ flags,
                                   const int *__restrict__ gstart, int n) {
  if (row >= g.ny * g.nz) return;  // whole wavefronts
  const int nx = g.nx, cy = row % g.ny, cz = row / g.ny;
  const int home = task_row_base(g, cy, cz, 4);  // dy = dz = 0
  int carry_p = 0, carry_h = 0;
  for (int c0 = 0; c0 < nx; c0 += 64) {
    const int x = c0 + lane;
    int col = 0, hm = 0;
    if (x < nx) {
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        const int rb = task_row_base(g, cy, cz, r);
        if (rb < 0) continue;  // open grids: outside the grid
        col += count[rb + x];
        if (g.open && gstart) col += gstart[rb + x + 1] - gstart[rb + x];
      }
      hm = count[home + x];
    }
    int ic = col, ih = hm;  // inclusive scans over the wavefront
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int vc = __shfl_up(ic, d, 64), vh = __shfl_up(ih, d, 64);
      if (lane >= d) {
        ic += vc;
        ih += vh;
      }
    }
    if (x < nx) {
      P[x + 1] = carry_p + ic;
      H[x + 1] = carry_h + ih;
    }
    carry_p += __shfl(ic, 63, 64);
    carry_h += __shfl(ih, 63, 64);
  }
  if (lane == 0) {
    P[0] = 0;
    H[0] = 0;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // staged atoms of [x0, x1): x-cells x0-1 .. x1 (periodic) of the 9 rows,
  // the whole row once when that covers it (task_piece_cells)
  auto staged = [&](int x0, int x1) {
    if (g.open) return P[x1 < nx ? x1 + 1 : nx] - P[x0 > 0 ? x0 - 1 : 0];
    if (x1 - x0 + 2 >= nx) return P[nx];
    if (x0 == 0) return P[x1 + 1] + (P[nx] - P[nx - 1]);
    if (x1 == nx) return P[nx] - P[x0 - 1] + P[1];
    return P[x1 + 1] - P[x0 - 1];
  };
  int x0 = 0, t = 0;
  bool ovf = false;
  // open grids: a row without owned atoms (the ghost margin) has no task
  if (g.open && H[nx] == 0) x0 = nx;
  while (x0 < nx) {  // uniform over the wavefront
    int x1 = x0 + 1;
    if (t + 1 < g.mpr) {
      // both limits grow with x1, so the admissible ends are a prefix
      for (int e0 = x0 + 1; e0 <= nx; e0 += 64) {
        const int e = e0 + lane;
        const bool ok = e <= nx && H[e] - H[x0] <= kTaskBlock && staged(x0, e) + 1 <= kStage;
        const unsigned long long m = __ballot(ok);
        if (m) x1 = e0 + 63 - __clzll(m);
        if (m != ~0ull) break;
      }
    } else {
      x1 = nx;
      ovf = ovf || H[nx] - H[x0] > kTaskBlock || staged(x0, nx) + 1 > kStage;
    }
    const int sl = row * g.mpr + t;
    // pieces: lane 2r + p holds piece p of stencil row r
    int j0 = 0, len = 0, a0 = 0, a1 = 0;
    if (lane < kTaskPieces) {
      const int rb = task_row_base(g, cy, cz, lane >> 1);
      if (!g.open) {
        task_piece_cells(nx, x0, x1, lane & 1, a0, a1);
        j0 = cs(rb + a0);
        len = a1 > a0 ? cs(rb + a1) - j0 : 0;
      } else if (rb >= 0) {
        // piece 2r: owned atoms, 2r + 1: ghosts (sorted section after the
        // owned atoms), of the same x-cells
        task_piece_cells_open(nx, x0, x1, a0, a1);
        if ((lane & 1) == 0) {
          j0 = cs(rb + a0);
          len = cs(rb + a1) - j0;
        } else if (gstart) {
          j0 = n + gstart[rb + a0];
          len = gstart[rb + a1] - gstart[rb + a0];
        }
      }
    }
    int base = len;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
      const int v = __shfl_up(base, d, 64);
      if (lane >= d) base += v;
    }
    base -= len;  // exclusive
    if (lane < kTaskPieces) tpiece[(size_t)sl * kTaskPieces + lane] = make_int4(j0, base, a0, a1);
    if (lane == kTaskPieces - 1) {
      const int S = base + len;
      tinfo[sl] = make_int4(cs(home + x0), cs(home + x1), S, x0 | (x1 << 16));
      if (S > 65535) atomicOr(&flags[F_STAGE], 1);
    }
    for (int x = x0 + lane; x < x1; x += 64) ctask[home + x] = sl;
    x0 = x1;
    ++t;
  }
  if (lane == 0) {
    rowtasks[row] = t;
    if (ovf) atomicOr(&flags[F_TASKOVF], 1);
  }
}

__global__ void k_build_begin(int *__restrict__ flags, int *__restrict__ cell_count, int ncell,
                              int *__restrict__ gcount) {
  if (!build_pending(flags)) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncell) {
    cell_count[c] = 0;
    if (gcount) gcount[c] = 0;
  }
}

__global__ void k_pack_q(const double4 *__restrict__ pos, int4 *__restrict__ q, int ntot, BoxDev b,
                         int *__restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntot) return;
  const double4 p = pos[i];
  if (isnan(p.x) | isnan(p.y) | isnan(p.z)) raise_fatal(flags, F_NAN);
  q[i] = qpack(p, b);
}

__global__ void k_set_flag(int *flags, int which, int value) { flags[which] = value; }

__global__ void k_fill_int(int *p, int n, int value) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

// ---- ensemble of independent periodic systems ----
//
// The same steps as the single-box build, restricted to the systems whose
// flag is set; each system has its own box and cell grid (per-atom list
// build over cells at least ce/2 wide).

// "all" rebuild policy: a request from any system rebuilds every system.
__global__ void k_ens_policy(EnsDev e, const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < e.nsys; s += gridDim.x * blockDim.x)
    e.sflag[s] = 1;
}

// Clear the flagged systems' cell counts; the last work-group also lists
// the flagged systems (EnsDev::act) in system order.
constexpr int kEnsBegin = 1024;
constexpr int kEnsListGrid = 8192;  // work-groups of the ensemble's wave build
__global__ void __launch_bounds__(kEnsBegin)
    k_ens_begin(EnsDev e, int *__restrict__ cell_count, int ncell, const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  if ((int)blockIdx.x < (int)gridDim.x - 1) {
    const int c = blockIdx.x * kEnsBegin + threadIdx.x;
    if (c < ncell && e.sflag[e.cell_sys[c]]) cell_count[c] = 0;
    return;
  }
  __shared__ int wsum[4][kEnsBegin / 64];
  __shared__ int carry[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x < 4) carry[threadIdx.x] = 0;
  __syncthreads();
  int *pre = e.act + 2, *first = e.act + 3 + e.nsys;
  // the flagged systems on the wave build and their cells (k_ens_nlist_cells)
  const int cb = ens_act_cells(e.nsys);
  int *cpre = e.act + cb + 2, *cfirst = e.act + cb + 3 + e.nsys;
  for (int s0 = 0; s0 < e.nsys; s0 += kEnsBegin) {
    const int s = s0 + (int)threadIdx.x;
    const bool on = s < e.nsys && e.sflag[s] != 0;
    const int na = on ? e.atom_off[s + 1] - e.atom_off[s] : 0;
    const bool onc = on && e.box[s].coarse;
    const int nc = onc ? e.box[s].nc[0] * e.box[s].nc[1] * e.box[s].nc[2] : 0;
    // inclusive scans: flagged systems, their atoms; wave-build ones, their cells
    int x = on ? 1 : 0, y = na, xc = onc ? 1 : 0, z = nc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int xv = __shfl_up(x, d, 64), yv = __shfl_up(y, d, 64);
      const int xcv = __shfl_up(xc, d, 64), zv = __shfl_up(z, d, 64);
      x += lane >= d ? xv : 0;
      y += lane >= d ? yv : 0;
      xc += lane >= d ? xcv : 0;
      z += lane >= d ? zv : 0;
    }
    if (lane == 63) {
      wsum[0][w] = x;
      wsum[1][w] = y;
      wsum[2][w] = xc;
      wsum[3][w] = z;
    }
    __syncthreads();
    int bx = carry[0], by = carry[1], bxc = carry[2], bz = carry[3];
    for (int k = 0; k < w; ++k) {
      bx += wsum[0][k];
      by += wsum[1][k];
      bxc += wsum[2][k];
      bz += wsum[3][k];
    }
    if (on) {  // the (bx + x - 1)-th flagged system
      pre[bx + x - 1] = by + y - na;
      first[bx + x - 1] = e.atom_off[s];
    }
    if (onc) {
      cpre[bxc + xc - 1] = bz + z - nc;
      cfirst[bxc + xc - 1] = e.cell_off[s];
    }
    __syncthreads();
    if (threadIdx.x == kEnsBegin - 1) {
      carry[0] = bx + x;
      carry[1] = by + y;
      carry[2] = bxc + xc;
      carry[3] = bz + z;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    e.act[0] = carry[0];
    e.act[1] = carry[1];
    pre[carry[0]] = carry[1];
    e.act[cb] = carry[2];
    e.act[cb + 1] = carry[3];
    cpre[carry[2]] = carry[3];
  }
}

// Atom t (t < total) of the flagged systems of a build: a binary search of
// the active list's atom prefix, staged in LDS when it fits.
constexpr int kActLds = 1024;
struct ActMap {
  int m, total;
  const int *pre, *first;
  __device__ int atom(int t) const {
    int lo = 0, hi = m - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    return first[lo] + (t - pre[lo]);
  }
};

// Every thread of the block calls this (it synchronizes the block).
__device__ inline ActMap act_map(const EnsDev &e, int *sh) {
  ActMap a;
  a.m = e.act[0];
  a.total = e.act[1];
  a.first = e.act + 3 + e.nsys;
  a.pre = e.act + 2;
  if (a.m + 1 <= kActLds) {
    for (int k = threadIdx.x; k <= a.m; k += blockDim.x) sh[k] = a.pre[k];
    __syncthreads();
    a.pre = sh;
  }
  return a;
}

// Wrap and bin the atoms of flagged systems; runs of equal cells in
// consecutive lanes take their slots with one atomic add.
// q_from (when given): the coordinates the fused force pass of the previous
// step wrote; they become q here, for every atom and before the early exit
// (the chain and the next force pass read q).
__global__ void k_ens_wrap_bin(double4 *__restrict__ pos, int n, EnsDev e, int *__restrict__ cell,
                               int *__restrict__ slot, int *__restrict__ cell_count,
                               const int *__restrict__ flags, int4 *__restrict__ q,
                               const int4 *__restrict__ q_from) {
  const int gsz = gridDim.x * blockDim.x;
  if (q_from)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gsz) q[i] = q_from[i];
  if (flags[F_REBUILD] == 0) return;
  __shared__ int act_sh[kActLds + 1];
  const ActMap am = act_map(e, act_sh);
  // block-uniform loop: the lanes of a wavefront stay together for the ballots
  for (int t0 = blockIdx.x * blockDim.x; t0 < am.total; t0 += gsz) {
  const int t = t0 + (int)threadIdx.x;
  const int lane = threadIdx.x & 63;
  int ci = -1;
  const bool valid = t < am.total;
  const int i = valid ? am.atom(t) : 0;
  const int sy = e.sys[i];
  if (valid) {
    const BoxDev &b = e.box[sy];
    double4 p = pos[i];
    double *c[3] = {&p.x, &p.y, &p.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (*c[d] < b.lo[d])
        *c[d] += b.L[d];
      else if (*c[d] > b.hi[d])
        *c[d] -= b.L[d];
    }
    pos[i] = p;
    ci = e.cell_off[sy] + cell_of(p, b);
    cell[i] = ci;
  }
  const int prev = __shfl_up(ci, 1, 64);
  const bool lead = valid && (lane == 0 || prev != ci);
  const unsigned long long lm = __ballot(lead);
  const unsigned long long ends = __ballot(lead || !valid);  // a run ends where the next starts
  const unsigned long long upto = lane == 63 ? lm : lm & ((2ull << lane) - 1ull);
  const int leader = upto ? 63 - __clzll(upto) : 0;
  int base = 0;
  if (lead) {
    const unsigned long long above = lane == 63 ? 0ull : ends & ~((2ull << lane) - 1ull);
    const int next = above ? __ffsll((long long)above) - 1 : 64;
    base = atomicAdd(&cell_count[ci], next - lane);
  }
  base = __shfl(base, leader, 64);
  if (valid) slot[i] = base + lane - leader;
  }
}

__global__ void k_ens_scatter(const int *__restrict__ cell, const int *__restrict__ slot,
                              const int *__restrict__ start, int *__restrict__ perm, int n, EnsDev e,
                              const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  __shared__ int act_sh[kActLds + 1];
  const ActMap am = act_map(e, act_sh);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < am.total; t += gridDim.x * blockDim.x) {
    const int i = am.atom(t);
    perm[start[cell[i]] + slot[i]] = i;
  }
}

__global__ void k_ens_rank(const int *__restrict__ cell, const int *__restrict__ start,
                           const int *__restrict__ perm, const int *__restrict__ tag,
                           int *__restrict__ order, int n, EnsDev e, const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  __shared__ int act_sh[kActLds + 1];
  const ActMap am = act_map(e, act_sh);
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < am.total; u += gridDim.x * blockDim.x) {
    const int i = am.atom(u);
    const int c = cell[i];
    const int a = start[c], end = start[c + 1];
    const int t = tag[i];
    int rank = 0;
    int k = a;
    for (; k + 8 <= end; k += 8) {  // eight members per round (independent loads)
      int v[8];
#pragma unroll
      for (int w = 0; w < 8; ++w) v[w] = perm[k + w];
#pragma unroll
      for (int w = 0; w < 8; ++w) v[w] = tag[v[w]];
#pragma unroll
      for (int w = 0; w < 8; ++w) rank += v[w] < t;
    }
    for (; k < end; ++k) rank += tag[perm[k]] < t;
    order[a + rank] = i;
  }
}

// Gather the flagged systems' atoms into their new order (k and order[k]
// are in the same system: atoms never change system).
__global__ void k_ens_permute(const int *__restrict__ order, int n, EnsDev e,
                              const double4 *__restrict__ pos, double4 *__restrict__ pos_out,
                              double4 *__restrict__ pos_ref, const double *__restrict__ vel,
                              double *__restrict__ vel_out, const double *__restrict__ acc,
                              double *__restrict__ acc_out, const int *__restrict__ tag,
                              int *__restrict__ tag_out, const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  __shared__ int act_sh[kActLds + 1];
  const ActMap am = act_map(e, act_sh);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < am.total; t += gridDim.x * blockDim.x) {
    const int k = am.atom(t);
    const int o = order[k];
    const double4 p = pos[o];
    pos_out[k] = p;
    pos_ref[k] = p;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      vel_out[d * n + k] = vel[d * n + o];
      acc_out[d * n + k] = acc[d * n + o];
    }
    tag_out[k] = tag[o];
  }
}

__global__ void k_ens_copy_back(int n, EnsDev e, const double4 *__restrict__ pos_alt,
                                double4 *__restrict__ pos, const double *__restrict__ vel_alt,
                                double *__restrict__ vel, const double *__restrict__ acc_alt,
                                double *__restrict__ acc, const int *__restrict__ tag_alt,
                                int *__restrict__ tag, int4 *__restrict__ q,
                                const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  __shared__ int act_sh[kActLds + 1];
  const ActMap am = act_map(e, act_sh);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < am.total; t += gridDim.x * blockDim.x) {
    const int k = am.atom(t);
    const double4 p = pos_alt[k];
    pos[k] = p;
    q[k] = qpack(p, e.box[e.sys[k]]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      vel[d * n + k] = vel_alt[d * n + k];
      if (acc) acc[d * n + k] = acc_alt[d * n + k];
    }
    tag[k] = tag_alt[k];
  }
}

// Per-atom Verlet list of the flagged systems (the FP64 predicate of the
// single-box per-atom build, with each system's box and cell grid).
__global__ void k_ens_nlist(const double4 *__restrict__ pos, int n, EnsDev e,
                            const int *__restrict__ cstart, uint32_t *__restrict__ nbr,
                            int *__restrict__ nnbr, int stride, int cap,
                            const int *__restrict__ flags, int2 *__restrict__ bstat) {
  if (flags[F_REBUILD] == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int cnt = 0, kept = 0;
  if (i < n) {
    const int sy = e.sys[i];
    if (e.sflag[sy] && !e.box[sy].coarse)  // wave-build systems: k_ens_nlist_cells
      nlist_atom(pos, n, e.box[sy], cstart + e.cell_off[sy], nullptr, nullptr, nbr, nnbr, stride,
                 cap, 0, i, cnt, kept);
  }
  block_stat(bstat, cnt, kept);
}

// Wave-per-cell Verlet list of the flagged systems that have the coarse
// filter (k_nlist_cells with each system's box, cell grid and fixed-point
// unit; every image has a ghost, ghost cutoff = list cutoff).  One
// wavefront per cell of the whole ensemble; cells of other systems return.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
    k_ens_nlist_cells(const double4 *__restrict__ pos, const int4 *__restrict__ q, EnsDev e,
                      int ncell, const int *__restrict__ cstart, uint32_t *__restrict__ nbr,
                      int *__restrict__ nnbr, int stride, int cap, const int *__restrict__ flags,
                      int2 *__restrict__ bstat) {
  if (flags[F_REBUILD] == 0) return;
  __shared__ __attribute__((aligned(16))) uint32_t lds[4][kPassCand + 4 + 64];
  __shared__ int cnt_sh[4][kMaxHomeCell];
  __shared__ int2 st_sh[4];
  __shared__ int cpre_sh[kActLds + 1];
  // the cells of the flagged wave-build systems (k_ens_begin), in groups of
  // four (one per wavefront); each XCD takes an eighth of the groups and its
  // work-groups stride through them, so that a step whose few systems
  // rebuild launches no idle work-groups for the others
  const int cb = ens_act_cells(e.nsys);
  const int cm = e.act[cb], ctot = e.act[cb + 1];
  const int *cpre = e.act + cb + 2, *cfirst = e.act + cb + 3 + e.nsys;
  if (cm + 1 <= kActLds) {
    for (int k = threadIdx.x; k <= cm; k += blockDim.x) cpre_sh[k] = cpre[k];
    __syncthreads();
    cpre = cpre_sh;
  }
  const int wv = (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int ng = (ctot + 3) / 4;
  const int xcd = (int)(blockIdx.x & 7), per = (int)(gridDim.x >> 3);
  const int g1 = (int)((long long)ng * (xcd + 1) / 8);
  int2 st = make_int2(0, 0);
  for (int g = (int)((long long)ng * xcd / 8) + (int)(blockIdx.x >> 3); g < g1; g += per) {
    const int t = g * 4 + wv;  // uniform over the wavefront
    if (t >= ctot) continue;
    int a = 0, z = cm - 1;  // the system of active cell t
    while (a < z) {
      const int mid = (a + z + 1) >> 1;
      if (cpre[mid] <= t)
        a = mid;
      else
        z = mid - 1;
    }
    const int c = __builtin_amdgcn_readfirstlane(cfirst[a] + (t - cpre[a]));
    const int sy = __builtin_amdgcn_readfirstlane(e.cell_sys[c]);
    const BoxDev *bd = e.box + sy;
    CellWaveBox cbx{};
    for (int d = 0; d < 3; ++d) {
      cbx.nc[d] = bd->nc[d];
      cbx.cunit[d] = bd->cunit[d];
      if (bd->nc[d] < 5) cbx.wide |= 1 << d;
    }
    cbx.cthr_in = bd->cthr_in;
    cbx.cthr_out = bd->cthr_out;
    const int off = e.cell_off[sy];
    const int2 r = cell_wave<true, false>(pos, q, cbx, bd, cstart + off, nullptr, nbr, nnbr, stride,
                                          cap, c - off, wv, lane, lds, nullptr, cnt_sh, nullptr,
                                          nullptr, nullptr);
    st = make_int2(max(st.x, r.x), st.y + r.y);
  }
  if (lane == 0) st_sh[wv] = st;
  __syncthreads();
  if (threadIdx.x == 0) {
    int2 t = st_sh[0];
    for (int w = 1; w < 4; ++w) t = make_int2(max(t.x, st_sh[w].x), t.y + st_sh[w].y);
    bstat[blockIdx.x] = t;
  }
}

// Close an ensemble build: longest list and overflow from the per-block
// statistics, per-system build counts, clear the requests.
__global__ void __launch_bounds__(1024)
    k_ens_done(int *__restrict__ flags, const int2 *__restrict__ bstat, int nw, int cap, EnsDev e) {
  if (flags[F_REBUILD] == 0) return;
  __shared__ int smax[1024];
  int m = 0;
  for (int k = threadIdx.x; k < nw; k += 1024) m = max(m, bstat[k].x);
  smax[threadIdx.x] = m;
  for (int s = threadIdx.x; s < e.nsys; s += 1024) {
    e.srebuilds[s] += e.sflag[s];
    e.sflag[s] = 0;
  }
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    flags[F_MAXNBR] = max(flags[F_MAXNBR], smax[0]);
    if (smax[0] > cap) flags[F_OVERFLOW] = 1;
    flags[F_REBUILD] = 0;
    flags[F_REBUILDS] += 1;
  }
}

inline int nblk(int n, int b) { return (n + b - 1) / b; }

void launch_scan(const int *count, int *start, int *tile_sum, int ncell, const int *flags,
                 hipStream_t st) {
  const int tiles = nblk(ncell, kScanTile);
  k_scan_tiles<<<tiles, kScanTile, 0, st>>>(count, start, tile_sum, ncell, flags);
  k_scan_add<<<tiles, kScanTile, 0, st>>>(start, tile_sum, ncell, tiles, flags);
}

}  // namespace

void launch_fill_int(int *p, int n, int value, hipStream_t st) {
  if (n > 0) k_fill_int<<<nblk(n, 256), 256, 0, st>>>(p, n, value);
}

void launch_build_ens(DevState &s, const EnsDev &e, int ncell, int block, bool policy_all,
                      bool per_atom, hipStream_t st, const int4 *q_from) {
  const int n = s.n;
  if (n == 0) return;
  if (policy_all) k_ens_policy<<<nblk(e.nsys, 256), 256, 0, st>>>(e, s.flags);
  k_ens_begin<<<nblk(ncell, kEnsBegin) + 1, kEnsBegin, 0, st>>>(e, s.cell_count, ncell, s.flags);
  // the per-atom kernels walk the flagged systems' atoms in a grid-stride loop
  const int ga = std::min(nblk(n, 256), 2048);
  k_ens_wrap_bin<<<ga, 256, 0, st>>>(s.pos, n, e, s.cell, s.slot, s.cell_count, s.flags,
                                                s.q, q_from);
  launch_scan(s.cell_count, s.cell_start, s.tile_sum, ncell, s.flags, st);
  k_ens_scatter<<<ga, 256, 0, st>>>(s.cell, s.slot, s.cell_start, s.perm, n, e, s.flags);
  k_ens_rank<<<ga, 256, 0, st>>>(s.cell, s.cell_start, s.perm, s.tag, s.order, n, e,
                                            s.flags);
  k_ens_permute<<<ga, 256, 0, st>>>(s.order, n, e, s.pos, s.pos_alt, s.pos_ref, s.vel,
                                               s.vel_alt, s.acc, s.acc_alt, s.tag, s.tag_alt,
                                               s.flags);
  k_ens_copy_back<<<ga, 256, 0, st>>>(n, e, s.pos_alt, s.pos, s.vel_alt, s.vel,
                                                 s.acc_alt, s.acc, s.tag_alt, s.tag, s.q, s.flags);
  // per-atom build only where some system has it (boxes the coarse filter
  // cannot serve)
  const int nb = per_atom ? nblk(n, block) : 0;
  if (per_atom)
    k_ens_nlist<<<nb, block, 0, st>>>(s.pos, n, e, s.cell_start, s.nbr, s.nnbr, s.stride, s.cap,
                                       s.flags, s.wstat);
  // a fixed grid (captured in the step graph) striding over the active cells
  const int gridc = std::min((nblk(ncell, 4) + 7) / 8 * 8, kEnsListGrid);
  k_ens_nlist_cells<<<gridc, 256, 0, st>>>(s.pos, s.q, e, ncell, s.cell_start, s.nbr, s.nnbr,
                                           s.stride, s.cap, s.flags, s.wstat + nb);
  k_ens_done<<<1, 1024, 0, st>>>(s.flags, s.wstat, nb + gridc, s.cap, e);
}

void launch_set_flag(int *flags, int which, int value, hipStream_t st) {
  k_set_flag<<<1, 1, 0, st>>>(flags, which, value);
}

void launch_pack_q(const DevState &s, const BoxDev &b, hipStream_t st) {
  const int ntot = s.n + s.ng;
  if (ntot == 0) return;
  k_pack_q<<<nblk(ntot, 256), 256, 0, st>>>(s.pos, s.q, ntot, b, s.flags);
}

void launch_build(DevState &s, const BoxDev &b, hipStream_t st, bool host_halo, bool keep_acc,
                  int step) {
  const int n = s.n;
  const int nc = s.ncell;
  const bool ghosts = host_halo && s.ng > 0;
  // Every kernel of the chain returns at once unless a rebuild is pending,
  // so the chain is kept short: one scan kernel, the cell starts and the
  // task table written by k_scatter, the fixed-point packing in k_rank_move
  // and (without host ghosts) the copy back in the list kernel.
  // Cell counts start cleared (allocation, then k_rank_move); the
  // host-halo ghost counts are cleared here.
  if (ghosts)
    k_build_begin<<<nblk(nc, 256), 256, 0, st>>>(s.flags, s.cell_count, nc, s.gcell_count);
  const bool coarse = n > 0 && b.coarse;
  const bool copy_in_list = coarse && !ghosts;
  // open wave builds sort the ghosts into the cell grid first: the task
  // partition counts them, and the list kernel reads them in cell order
  const bool gsort = ghosts && coarse && b.open;
  const CopyBack cpb{s.pos_alt,  s.pos,     s.vel_alt, s.vel, keep_acc ? s.acc_alt : nullptr,
                     keep_acc ? s.acc : nullptr, s.tag_alt, s.tag, s.id_alt,  s.id,  n,
                     s.gpos_alt, s.gid_alt, s.gtag_alt, s.gtag, s.gslot, gsort ? s.ng : 0};
  auto ghost_cells = [&]() {
    k_ghost_bin<<<nblk(s.ng, 256), 256, 0, st>>>(s.pos, s.pos_ref, n, s.ng, b, s.cell + n,
                                                  s.slot + n, s.gcell_count, s.flags);
    launch_scan(s.gcell_count, s.gcell_start, s.tile_sum, nc, s.flags, st);
    k_ghost_scatter<<<nblk(s.ng, 256), 256, 0, st>>>(s.cell + n, s.slot + n, s.gcell_start,
                                                      s.gcell_items, s.ng, s.flags);
  };
  if (gsort) {
    ghost_cells();
    k_ghost_move<<<nblk(s.ng, 256), 256, 0, st>>>(s.gcell_items, s.cell + n, s.gcell_start, n, s.ng,
                                                   s.pos, s.id, s.gtag, s.gpos_alt, s.gid_alt,
                                                   s.gtag_alt, s.pos_ref, s.q, b, s.flags);
  }
  if (n > 0) {
    k_wrap_bin<<<nblk(n, 256), 256, 0, st>>>(s.pos, n, b, s.cell, s.slot, s.cell_count, s.flags);
    k_scan_local<<<nblk(nc, kScanTile), kScanTile, 0, st>>>(s.cell_count, s.cell_local, s.tile_sum,
                                                           s.cell_start, nc, s.flags);
    const CellStarts cs{s.cell_local, s.tile_sum, s.cell_start, nc};
    const int nbs = nblk(n, 256), nfix = nblk(nc, 256);
    const bool tasks = coarse && s.slots;
    const int rows = tasks ? s.tg.ny * s.tg.nz : 0;
    k_scatter<<<nbs + nfix + nblk(rows, kRowWaves), 256, 0, st>>>(
        s.cell, s.slot, cs, s.cell_start, s.perm, n, s.flags, nbs, nfix, s.tg, s.cell_count,
        s.tpiece, s.tinfo, s.ctask, s.rowtasks, gsort ? s.gcell_start : nullptr);
    // accelerations are carried along only when no force pass follows; the
    // extra last work-groups (tasks) copy the task tables to compact order
    const TaskMapArgs tm{s.rowtasks, s.tinfo, s.tpiece, s.ctinfo, s.ctpiece, rows, s.tg.mpr,
                         compact_parts(rows)};
    k_rank_move<<<nblk(n, 256) + tm.parts, 256, 0, st>>>(
        s.cell, s.cell_start, s.perm, n, s.pos, s.pos_alt, s.pos_ref, s.vel, s.vel_alt,
        keep_acc ? s.acc : nullptr, s.acc_alt, s.tag, s.tag_alt, s.id, s.id_alt, s.cell_count, nc,
        (s.write_q || b.coarse) ? s.q : nullptr, b, s.flags, tm);
    // The host cannot know whether the device ran the build, so the new
    // order is copied back instead of swapping buffer pointers.
    if (!copy_in_list) k_copy_back<<<nblk(n + cpb.ng, 256), 256, 0, st>>>(cpb, s.flags);
  }
  if (ghosts && !gsort) {  // per-atom build: the ghost lists of the cells, by device index
    ghost_cells();
    k_ghost_sort_cells<<<nblk(nc, 256), 256, 0, st>>>(s.gcell_items, s.gcell_start, nc, s.flags);
  }
  int nw = 0;  // per-work-group statistics written by the list kernel
  if (coarse) {
    CellWaveBox cb{};
    for (int d = 0; d < 3; ++d) {
      cb.nc[d] = b.nc[d];
      cb.cunit[d] = b.cunit[d];
    }
    cb.cthr_in = b.cthr_in;
    cb.cthr_out = b.cthr_out;
    cb.open = b.open;
    cb.n = n;
    const int real = nblk(nc, 4), grid = (real + 7) / 8 * 8;
    // (the copy-back work-groups stride over the atoms: a fixed 1024 of them
    // keep the idle launch of steps without a rebuild short)
    const int ncopy = copy_in_list ? std::min((nblk(n, 256) + 7) / 8 * 8, 1024) : 0;
    // the list kernel reads the new-order positions where they were moved
    const double4 *pos_new = copy_in_list ? s.pos_alt : s.pos;
#define CV2_NLC(G, S)                                                                            \
  if (b.ghost_ok == G && s.slots == S)                                                           \
    k_nlist_cells<G, S><<<ncopy + grid, 256, 0, st>>>(pos_new, s.q, cb, s.box_dev, s.cell_start, \
                                                      gsort ? s.gcell_start : nullptr, s.nbr,    \
                                                      s.nnbr, s.stride,                          \
                                                      s.cap, s.flags,                            \
                                                      real, s.wstat, s.ctask, s.tpiece, s.tinfo, \
                                                      ncopy, cpb);
    CV2_NLC(true, true)
    CV2_NLC(true, false)
    CV2_NLC(false, true)
    CV2_NLC(false, false)
#undef CV2_NLC
    nw = real;
  } else if (n > 0) {
    k_nlist<<<nblk(n, 256), 256, 0, st>>>(s.pos, n, b, s.cell_start, s.gcell_start,
                                          s.gcell_items, s.nbr, s.nnbr, s.stride, s.cap, s.flags,
                                          ghosts ? 1 : 0, s.wstat);
    nw = nblk(n, 256);
  }
  k_build_stats<<<std::max(1, std::min(nblk(nw, 256), 64)), 256, 0, st>>>(s.flags, s.wstat, nw,
                                                                           s.cap, step);
}

TaskGeom task_geom(int n, const BoxDev &b, int min_mpr) {
  TaskGeom g{};
  g.nx = b.nc[0];
  g.ny = b.nc[1];
  g.nz = b.nc[2];
  // a row of the mean density needs about row / 230 tasks (at most
  // kTaskBlock home atoms, neighbourhood within the LDS stage); rows of a
  // lattice or a dense region need more, so the slots allow 1.5x that + 2
  const double row = (double)n / ((double)g.ny * g.nz);
  const int want = (int)std::ceil(1.5 * row / 230.0) + 2;
  g.mpr = std::max(1, std::min(g.nx, std::max(min_mpr, want)));
  g.nslot = g.ny * g.nz * g.mpr;
  g.open = b.open;
  return g;
}

}  // namespace cv2
