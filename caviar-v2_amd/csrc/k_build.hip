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

// Wrap, bin, count.  One thread per owned atom.
__global__ void k_wrap_bin(double4 *__restrict__ pos, int n, BoxDev b, int *__restrict__ cell,
                           int *__restrict__ slot, int *__restrict__ cell_count,
                           const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 p = pos[i];
  double *c[3] = {&p.x, &p.y, &p.z};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (b.bc[d] != 1) continue;
    if (*c[d] < b.lo[d])
      *c[d] += b.L[d];
    else if (*c[d] > b.hi[d])
      *c[d] -= b.L[d];
  }
  pos[i] = p;
  const int ci = cell_of(p, b);
  cell[i] = ci;
  slot[i] = atomicAdd(&cell_count[ci], 1);
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
  if (flags[F_REBUILD] == 0) return;
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
  if (flags[F_REBUILD] == 0) return;
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

__global__ void k_scatter(const int *__restrict__ cell, const int *__restrict__ slot,
                          const int *__restrict__ start, int *__restrict__ perm, int n,
                          const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  perm[start[cell[i]] + slot[i]] = i;
}

// Order atoms inside each cell by upload index so that the sorted order,
// and with it every list and sum, is independent of atomic arrival order.
// One wavefront per cell: each lane holds one atom and finds its rank by
// comparing its key with every other lane's (keys are unique); cells with
// more than 64 atoms fall back to an insertion sort by one lane.
__global__ void __launch_bounds__(256)
    k_sort_cells(int *__restrict__ perm, const int *__restrict__ start,
                 const int *__restrict__ tag, int ncell, const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= ncell) return;
  const int a = start[c], m = start[c + 1] - a;
  if (m <= 1) return;
  if (m <= 64) {
    const int v = lane < m ? perm[a + lane] : 0;
    const int key = lane < m ? tag[v] : 0x7fffffff;
    int rank = 0;
    for (int l = 0; l < m; ++l) rank += __shfl(key, l, 64) < key;
    if (lane < m) perm[a + rank] = v;
    return;
  }
  if (lane != 0) return;
  for (int k = a + 1; k < a + m; ++k) {
    const int v = perm[k];
    const int tv = tag[v];
    int q = k - 1;
    while (q >= a && tag[perm[q]] > tv) {
      perm[q + 1] = perm[q];
      --q;
    }
    perm[q + 1] = v;
  }
}

// Gather every per-atom array into the new order; remember build positions.
__global__ void k_permute(const int *__restrict__ perm, int n, const double4 *__restrict__ pos,
                          double4 *__restrict__ pos_out, double4 *__restrict__ pos_ref,
                          const double *__restrict__ vel, double *__restrict__ vel_out,
                          const double *__restrict__ acc, double *__restrict__ acc_out,
                          const int *__restrict__ tag, int *__restrict__ tag_out,
                          const int *__restrict__ id, int *__restrict__ id_out,
                          const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int o = perm[k];
  const double4 p = pos[o];
  pos_out[k] = p;
  pos_ref[k] = p;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    vel_out[d * n + k] = vel[d * n + o];
    acc_out[d * n + k] = acc[d * n + o];
  }
  tag_out[k] = tag[o];
  id_out[k] = id[o];
}

// Fixed-point copy of positions for the FP32 pair kernel: (x - lo) * 2^32 / L
// reduced modulo 2^32, so that int32 differences are minimum-image
// separations on periodic axes.
__device__ inline int4 qpack(const double4 &p, const BoxDev &b) {
  return pack_q(p.x, p.y, p.z, (int)p.w, b.lo, b.qscale);
}

__global__ void k_copy_back(int n, const double4 *__restrict__ pos_alt, double4 *__restrict__ pos,
                            const double *__restrict__ vel_alt, double *__restrict__ vel,
                            const double *__restrict__ acc_alt, double *__restrict__ acc,
                            const int *__restrict__ tag_alt, int *__restrict__ tag,
                            const int *__restrict__ id_alt, int *__restrict__ id,
                            int4 *__restrict__ q, BoxDev b, const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double4 p = pos_alt[k];
  pos[k] = p;
  if (q) q[k] = qpack(p, b);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    vel[d * n + k] = vel_alt[d * n + k];
    acc[d * n + k] = acc_alt[d * n + k];
  }
  tag[k] = tag_alt[k];
  id[k] = id_alt[k];
}

// Host-halo ghosts keep host order; they get their own CSR cell list and
// their build positions.
__global__ void k_ghost_bin(const double4 *__restrict__ pos, double4 *__restrict__ pos_ref,
                            int n, int ng, BoxDev b, int *__restrict__ cell,
                            int *__restrict__ slot, int *__restrict__ cell_count,
                            const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
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
  if (flags[F_REBUILD] == 0) return;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  items[start[cell[g]] + slot[g]] = g;
}

__global__ void k_ghost_sort_cells(int *__restrict__ items, const int *__restrict__ start,
                                   int ncell, const int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
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

// Candidate cells along one axis: c-1, c, c+1 (periodic wrap, or clamped
// for open axes), duplicates removed when the grid has fewer than 3 cells.
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

__global__ void __launch_bounds__(256)
    k_nlist(const double4 *__restrict__ pos, int n, BoxDev b, const int *__restrict__ cstart,
            const int *__restrict__ gstart, const int *__restrict__ gitems,
            uint32_t *__restrict__ nbr, int *__restrict__ nnbr, int stride, int cap,
            int *__restrict__ flags, int host_halo) {
  if (flags[F_REBUILD] == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
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
  if (out.cnt > cap) atomicOr(&flags[F_OVERFLOW], 1);
  atomicMax(&flags[F_MAXNBR], out.cnt);
  atomicAdd(&flags[F_ENTRIES], kept);
}

__global__ void k_build_done(int *__restrict__ flags) {
  if (flags[F_REBUILD] == 0) return;
  flags[F_REBUILD] = 0;
  flags[F_REBUILDS] += 1;
}

__global__ void k_build_begin(int *__restrict__ flags, int *__restrict__ cell_count, int ncell,
                              int *__restrict__ gcount) {
  if (flags[F_REBUILD] == 0) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncell) {
    cell_count[c] = 0;
    if (gcount) gcount[c] = 0;
  }
  if (c == 0) {
    flags[F_MAXNBR] = 0;
    flags[F_ENTRIES] = 0;
  }
}

__global__ void k_pack_q(const double4 *__restrict__ pos, int4 *__restrict__ q, int ntot, BoxDev b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntot) return;
  q[i] = qpack(pos[i], b);
}

__global__ void k_set_flag(int *flags, int which, int value) { flags[which] = value; }

inline int nblk(int n, int b) { return (n + b - 1) / b; }

void launch_scan(const int *count, int *start, int *tile_sum, int ncell, const int *flags,
                 hipStream_t st) {
  const int tiles = nblk(ncell, kScanTile);
  k_scan_tiles<<<tiles, kScanTile, 0, st>>>(count, start, tile_sum, ncell, flags);
  k_scan_add<<<tiles, kScanTile, 0, st>>>(start, tile_sum, ncell, tiles, flags);
}

}  // namespace

void launch_set_flag(int *flags, int which, int value, hipStream_t st) {
  k_set_flag<<<1, 1, 0, st>>>(flags, which, value);
}

void launch_pack_q(const DevState &s, const BoxDev &b, hipStream_t st) {
  const int ntot = s.n + s.ng;
  if (ntot == 0) return;
  k_pack_q<<<nblk(ntot, 256), 256, 0, st>>>(s.pos, s.q, ntot, b);
}

void launch_build(DevState &s, const BoxDev &b, hipStream_t st, bool host_halo) {
  const int n = s.n;
  const int nc = s.ncell;
  const bool ghosts = host_halo && s.ng > 0;
  k_build_begin<<<nblk(nc, 256), 256, 0, st>>>(s.flags, s.cell_count, nc,
                                                ghosts ? s.gcell_count : nullptr);
  if (n > 0) {
    k_wrap_bin<<<nblk(n, 256), 256, 0, st>>>(s.pos, n, b, s.cell, s.slot, s.cell_count, s.flags);
    launch_scan(s.cell_count, s.cell_start, s.tile_sum, nc, s.flags, st);
    k_scatter<<<nblk(n, 256), 256, 0, st>>>(s.cell, s.slot, s.cell_start, s.perm, n, s.flags);
    k_sort_cells<<<nblk(nc, 4), 256, 0, st>>>(s.perm, s.cell_start, s.tag, nc, s.flags);
    k_permute<<<nblk(n, 256), 256, 0, st>>>(s.perm, n, s.pos, s.pos_alt, s.pos_ref, s.vel,
                                            s.vel_alt, s.acc, s.acc_alt, s.tag, s.tag_alt, s.id,
                                            s.id_alt, s.flags);
    // The host cannot know whether the device ran the build, so the new
    // order is copied back instead of swapping buffer pointers.
    k_copy_back<<<nblk(n, 256), 256, 0, st>>>(n, s.pos_alt, s.pos, s.vel_alt, s.vel, s.acc_alt,
                                              s.acc, s.tag_alt, s.tag, s.id_alt, s.id,
                                              s.write_q ? s.q : nullptr, b, s.flags);
  }
  if (ghosts) {
    k_ghost_bin<<<nblk(s.ng, 256), 256, 0, st>>>(s.pos, s.pos_ref, n, s.ng, b, s.cell + n,
                                                  s.slot + n, s.gcell_count, s.flags);
    launch_scan(s.gcell_count, s.gcell_start, s.tile_sum, nc, s.flags, st);
    k_ghost_scatter<<<nblk(s.ng, 256), 256, 0, st>>>(s.cell + n, s.slot + n, s.gcell_start,
                                                      s.gcell_items, s.ng, s.flags);
    k_ghost_sort_cells<<<nblk(nc, 256), 256, 0, st>>>(s.gcell_items, s.gcell_start, nc, s.flags);
  }
  if (n > 0)
    k_nlist<<<nblk(n, 256), 256, 0, st>>>(s.pos, n, b, s.cell_start, s.gcell_start,
                                          s.gcell_items, s.nbr, s.nnbr, s.stride, s.cap, s.flags,
                                          ghosts ? 1 : 0);
  k_build_done<<<1, 1, 0, st>>>(s.flags);
}

}  // namespace cv2
