// k_halo.hip -- the device side of a domain's MPI ghost exchange.
//
// Atom_data::exchange_ghost_mpi_shared_atoms (exchange_ghost_mpi_shared_atoms.
// cpp:113-325) classifies every owned atom by the faces of its domain it is
// within the ghost cutoff of (-1 below lower_local + g, +1 above
// upper_local - g, zeroed at a non-periodic grid edge), sends it to each of
// the up to seven directions its nonzero components span (x, y, z, xy, xz,
// yz, xyz), and fills one packet per direction, atoms in owned-index order:
// [N ids][N types][3N positions][3N velocities], positions shifted by the
// global box length where they cross a periodic grid edge.  Here the owned
// index is the upload index (the device keeps atoms in cell order and
// `tag` maps back), so the packets are the reference's, bit for bit.
//
// Packing runs over tiles of kHaloTile upload indices:
//   k_halo_dirs   per device atom: its direction code, stored by upload index
//   k_halo_count  per tile: atoms per direction (LDS counters)
//   (host)        tile offsets per direction, packet bases
//   k_halo_write  per tile: each atom's rank in each of its directions
//                 (wavefront ballots, then a prefix over the tile's
//                 wavefronts), and its packet entries
// Unpacking writes the ghosts of the received packets, in the order given,
// after the owned atoms (the reference appends them likewise, :405-470).
#include "cv2_internal.h"

namespace cv2 {

namespace {

constexpr int kHaloTile = 1024;

// Direction code of a position: (v_a + 1) in 2 bits per axis, 0x15 = none.
__device__ inline int halo_code(const double4 &p, const HaloGeomDev &g) {
  const double c[3] = {p.x, p.y, p.z};
  int code = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int v = 0;
    if (c[a] < g.lo[a] + g.cut) {
      v = -1;
      if (g.gi[a] == 0) v *= g.bc[a];
    } else if (c[a] > g.hi[a] - g.cut) {
      v = +1;
      if (g.gi[a] == g.np[a] - 1) v *= g.bc[a];
    }
    code |= (v + 1) << (2 * a);
  }
  return code;
}

// Whether an atom of direction code `code` goes to direction d = i*9+j*3+k:
// on every axis the direction's component is 1 (axis not used) or the
// atom's own value, and d is not (1,1,1).
__device__ inline bool halo_in(int code, int d) {
  const int c[3] = {d / 9, (d / 3) % 3, d % 3};
  bool in = d != 13;
#pragma unroll
  for (int a = 0; a < 3; ++a) in = in & ((c[a] == 1) | (c[a] == ((code >> (2 * a)) & 3)));
  return in;
}

__global__ void k_halo_dirs(const double4 *__restrict__ pos, const int *__restrict__ tag, int n,
                            HaloGeomDev g, int *__restrict__ code_by_tag,
                            int *__restrict__ src_by_tag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int t = tag[k];
  code_by_tag[t] = halo_code(pos[k], g);
  src_by_tag[t] = k;
}

__global__ void __launch_bounds__(kHaloTile)
    k_halo_count(const int *__restrict__ code_by_tag, int n, int *__restrict__ tile_counts) {
  __shared__ int cnt[27];
  if (threadIdx.x < 27) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int t = blockIdx.x * kHaloTile + threadIdx.x;
  const int code = t < n ? code_by_tag[t] : 0x15;
  if (code != 0x15)
    for (int d = 0; d < 27; ++d)
      if (halo_in(code, d)) atomicAdd(&cnt[d], 1);
  __syncthreads();
  if (threadIdx.x < 27) tile_counts[blockIdx.x * 27 + threadIdx.x] = cnt[threadIdx.x];
}

// tile_off[tile * 27 + d]: entries of direction d in the tiles before;
// dir[d] = (packet base in doubles, N).
__global__ void __launch_bounds__(kHaloTile)
    k_halo_write(const double4 *__restrict__ pos, const double *__restrict__ vel,
                 const int *__restrict__ id, const int *__restrict__ code_by_tag,
                 const int *__restrict__ src_by_tag, int n, HaloGeomDev g,
                 const int *__restrict__ tile_off, const long long *__restrict__ dir_base,
                 const int *__restrict__ dir_n, double *__restrict__ out) {
  constexpr int kWaves = kHaloTile / 64;
  __shared__ int wcnt[kWaves][27];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t = blockIdx.x * kHaloTile + threadIdx.x;
  const int code = t < n ? code_by_tag[t] : 0x15;
  // each atom's rank among the tile's atoms of each of its directions,
  // in upload order: within the wavefront by ballot, then by wavefront
  int rank[27];
#pragma unroll
  for (int d = 0; d < 27; ++d) {
    const bool in = code != 0x15 && halo_in(code, d);
    const unsigned long long m = __ballot(in);
    rank[d] = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    if (lane == 0) wcnt[w][d] = __popcll(m);
  }
  __syncthreads();
  if (code == 0x15) return;
  const int k = src_by_tag[t];
  const double4 p = pos[k];
  const double pc[3] = {p.x, p.y, p.z};
  const double v[3] = {vel[k], vel[n + k], vel[2 * n + k]};
#pragma unroll
  for (int d = 0; d < 27; ++d) {
    if (!halo_in(code, d)) continue;
    int r = tile_off[blockIdx.x * 27 + d] + rank[d];
    for (int u = 0; u < w; ++u) r += wcnt[u][d];
    const long long N = dir_n[d];
    double *pk = out + dir_base[d];
    const int c[3] = {d / 9, (d / 3) % 3, d % 3};
    pk[r] = (double)id[k];
    pk[N + r] = p.w;  // the type
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double x = pc[a];
      if (g.bc[a] == 1) {
        if (g.gi[a] == 0 && c[a] == 0) x += g.size[a];
        if (g.gi[a] == g.np[a] - 1 && c[a] == 2) x -= g.size[a];
      }
      pk[2 * N + 3 * r + a] = x;
      if (g.with_vel) pk[5 * N + 3 * r + a] = v[a];
    }
  }
}

// Tile offsets, per-direction totals and packet bases on the device: one
// work-group per direction d.  It scans direction d's counts over the tiles
// into tile_off, stores its total N_d (dir_n, res[d]), and sums the counts
// of the directions before it for its packet base (the packets follow each
// other in direction order; dir_base in doubles, res[27] = their total).
// skip13: direction 13 holds no packet (the owned exchange counts its
// staying atoms there) and is left out of the bases and the total.
__global__ void __launch_bounds__(1024)
    k_halo_scan(const int *__restrict__ tile_counts, int ntiles, int per_atom, int skip13,
                int *__restrict__ tile_off, int *__restrict__ dir_n, long long *__restrict__ dir_base,
                long long *__restrict__ res) {
  __shared__ int sh[1024];
  __shared__ long long red[16];
  const int t = threadIdx.x, d = blockIdx.x;
  int carry = 0;
  for (int t0 = 0; t0 < ntiles; t0 += 1024) {
    const int v = t0 + t < ntiles ? tile_counts[(t0 + t) * 27 + d] : 0;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int a = t >= off ? sh[t - off] : 0;
      __syncthreads();
      sh[t] += a;
      __syncthreads();
    }
    if (t0 + t < ntiles) tile_off[(t0 + t) * 27 + d] = carry + sh[t] - v;
    carry += sh[1023];
    __syncthreads();
  }
  // atoms of the directions before d
  long long before = 0;
  for (int k = t; k < ntiles; k += 1024)
    for (int e = 0; e < d; ++e)
      if (!(skip13 && e == 13)) before += tile_counts[k * 27 + e];
  for (int o = 32; o > 0; o >>= 1) before += __shfl_xor(before, o, 64);
  if ((t & 63) == 0) red[t >> 6] = before;
  __syncthreads();
  if (t == 0) {
    long long b = 0;
    for (int w = 0; w < 16; ++w) b += red[w];
    dir_n[d] = carry;
    dir_base[d] = (long long)per_atom * b;
    res[d] = carry;
    if (d == 26) res[27] = (long long)per_atom * (b + carry);
    if (skip13 && d == 13) dir_base[d] = 0;
  }
}

// Ghost g of the concatenated received packets (pk_first[p] = first ghost of
// packet p, pk_base[p] = its first double, pk_n[p] = its atoms).  Where the
// list build keeps the ghosts in cell order (gslot given, the ghost count
// unchanged) ghost g goes to device slot gslot[g], so that each ghost is
// compared with its own build position; otherwise to slot g, and gtag /
// gslot (when allocated) restart as the identity.
__global__ void k_halo_unpack(const double *__restrict__ in, int npk,
                              const long long *__restrict__ pk_first,
                              const long long *__restrict__ pk_base,
                              const long long *__restrict__ pk_n, int ng, int n,
                              double4 *__restrict__ pos, int *__restrict__ id,
                              int *__restrict__ gtag, int *__restrict__ gslot, int remap) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= ng) return;
  int p = 0;
  while (p + 1 < npk && pk_first[p + 1] <= gi) ++p;
  const long long N = pk_n[p], c = gi - pk_first[p];
  const double *pk = in + pk_base[p];
  int k = gi;
  if (remap) {
    k = gslot[gi];
  } else if (gtag) {
    gtag[gi] = gi;
    gslot[gi] = gi;
  }
  pos[n + k] = make_double4(pk[2 * N + 3 * c], pk[2 * N + 3 * c + 1], pk[2 * N + 3 * c + 2],
                            pk[N + c]);
  id[n + k] = (int)pk[c];
}

// The same placement for ghosts uploaded as arrays (cv2_set_ghosts), staged
// in gpos_in / gid_in.
__global__ void k_ghost_place(const double4 *__restrict__ gpos_in, const int *__restrict__ gid_in,
                              int ng, int n, double4 *__restrict__ pos, int *__restrict__ id,
                              int *__restrict__ gtag, int *__restrict__ gslot, int remap) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= ng) return;
  int k = gi;
  if (remap) {
    k = gslot[gi];
  } else {
    gtag[gi] = gi;
    gslot[gi] = gi;
  }
  pos[n + k] = gpos_in[gi];
  id[n + k] = gid_in[gi];
}

}  // namespace

int halo_tiles(int n) { return (n + kHaloTile - 1) / kHaloTile; }

void launch_halo_classify(const DevState &s, const HaloGeomDev &g, int *code_by_tag,
                          int *src_by_tag, int *tile_counts, hipStream_t st) {
  if (s.n == 0) return;
  k_halo_dirs<<<(s.n + 255) / 256, 256, 0, st>>>(s.pos, s.tag, s.n, g, code_by_tag, src_by_tag);
  k_halo_count<<<halo_tiles(s.n), kHaloTile, 0, st>>>(code_by_tag, s.n, tile_counts);
}

void launch_halo_scan(int ntiles, int per_atom, const int *tile_counts, int *tile_off, int *dir_n,
                      long long *dir_base, long long *res, hipStream_t st, bool skip13) {
  k_halo_scan<<<27, 1024, 0, st>>>(tile_counts, ntiles, per_atom, skip13 ? 1 : 0, tile_off, dir_n,
                                   dir_base, res);
}

void launch_halo_write(const DevState &s, const HaloGeomDev &g, const int *code_by_tag,
                       const int *src_by_tag, const int *tile_off, const long long *dir_base,
                       const int *dir_n, double *out, hipStream_t st) {
  if (s.n == 0) return;
  k_halo_write<<<halo_tiles(s.n), kHaloTile, 0, st>>>(s.pos, s.vel, s.id, code_by_tag, src_by_tag,
                                                      s.n, g, tile_off, dir_base, dir_n, out);
}

void launch_halo_unpack(DevState &s, const double *in, int npk, const long long *pk_first,
                        const long long *pk_base, const long long *pk_n, int ng, bool remap,
                        hipStream_t st) {
  if (ng == 0) return;
  k_halo_unpack<<<(ng + 255) / 256, 256, 0, st>>>(in, npk, pk_first, pk_base, pk_n, ng, s.n, s.pos,
                                                   s.id, s.gtag, s.gslot, remap ? 1 : 0);
}

void launch_ghost_place(DevState &s, const double4 *gpos_in, const int *gid_in, int ng, bool remap,
                        hipStream_t st) {
  if (ng == 0) return;
  k_ghost_place<<<(ng + 255) / 256, 256, 0, st>>>(gpos_in, gid_in, ng, s.n, s.pos, s.id, s.gtag,
                                                   s.gslot, remap ? 1 : 0);
}

}  // namespace cv2
