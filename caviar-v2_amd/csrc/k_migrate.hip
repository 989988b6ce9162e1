// k_migrate.hip -- the device side of a domain's owned-atom exchange.
//
// Atom_data::exchange_owned_mpi_shared_atoms (exchange_owned_mpi_shared_atoms.
// cpp:44-602) moves every owned atom that has left its domain to the domain
// it entered: an atom strictly below lower_local on an axis goes -1 there,
// strictly above upper_local +1, and the component is zeroed at a
// non-periodic grid edge (:136-190); the atom goes to the one direction
// d = i*9 + j*3 + k (i = x_val + 1, ...) its three components name.  Where
// that neighbour is the domain itself (one process on every axis the
// direction moves along) the atom stays and is shifted by the local box
// length instead (:571-583); otherwise it leaves in direction d's packet,
// atoms in owned order, [N ids][3N positions][3N velocities]
// [3N accelerations][N molecule indices][N bond counts] as doubles
// (:330-369; no optional old-state or msd blocks), its position shifted by
// the global box length where it crosses a periodic grid edge (:402-431).
//
// Packing runs over the same tiles of 1024 upload indices as the ghost
// halo (k_halo.hip): direction codes by upload index, per-tile counts, the
// device scan (staying atoms are counted as direction 13 and left out of
// the packet bases), then each tile writes its leavers' packet entries and
// its staying atoms, compacted in owned order, into the alt arrays.  The
// received atoms are merged in by id (k_mig_place_recv).
#include "cv2_internal.h"

namespace cv2 {

namespace {

constexpr int kMigTile = 1024;  // = the halo tile (halo_tiles)
constexpr int kMigPerAtom = 12;

// Direction of an owned position (:136-190), 13 = inside the domain.
__device__ inline int mig_dir(const double4 &p, const HaloGeomDev &g, int v[3]) {
  const double c[3] = {p.x, p.y, p.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    v[a] = 0;
    if (c[a] < g.lo[a]) {
      v[a] = -1;
      if (g.gi[a] == 0) v[a] *= g.bc[a];
    } else if (c[a] > g.hi[a]) {
      v[a] = +1;
      if (g.gi[a] == g.np[a] - 1) v[a] *= g.bc[a];
    }
  }
  return (v[0] + 1) * 9 + (v[1] + 1) * 3 + (v[2] + 1);
}

// The position an owned atom keeps: shifted by the local box length where
// its direction leads back to this domain (bit d of selfmask, :571-583).
__device__ inline bool mig_self_shift(double4 &p, const HaloGeomDev &g, unsigned selfmask) {
  int v[3];
  const int d = mig_dir(p, g, v);
  if (d == 13 || !((selfmask >> d) & 1u)) return false;
  p.x -= v[0] * (g.hi[0] - g.lo[0]);
  p.y -= v[1] * (g.hi[1] - g.lo[1]);
  p.z -= v[2] * (g.hi[2] - g.lo[2]);
  return true;
}

// Per device atom: its direction by upload index (13: it stays, inside or
// shifted back; *moved is raised, once per wavefront, when some atom is
// shifted).  No state changes here, so a call that only sizes the packets
// leaves the atoms as they were.
__global__ void k_mig_classify(const double4 *__restrict__ pos, const int *__restrict__ tag, int n,
                               HaloGeomDev g, unsigned selfmask, int *__restrict__ code_by_tag,
                               int *__restrict__ src_by_tag, int *__restrict__ moved) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  bool shifted = false;
  if (k < n) {
    const double4 p = pos[k];
    int v[3];
    int d = mig_dir(p, g, v);
    if (d != 13 && ((selfmask >> d) & 1u)) {
      d = 13;
      shifted = true;
    }
    const int t = tag[k];
    code_by_tag[t] = d;
    src_by_tag[t] = k;
  }
  if (__ballot(shifted) != 0ull && (threadIdx.x & 63) == 0) atomicOr(moved, 1);
}

// The shifts alone, in place (no atom leaves: order and count unchanged).
__global__ void k_mig_shift(double4 *__restrict__ pos, int n, HaloGeomDev g, unsigned selfmask) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double4 p = pos[k];
  if (mig_self_shift(p, g, selfmask)) pos[k] = p;
}

__global__ void __launch_bounds__(kMigTile)
    k_mig_count(const int *__restrict__ code_by_tag, int n, int *__restrict__ tile_counts) {
  __shared__ int cnt[27];
  if (threadIdx.x < 27) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int t = blockIdx.x * kMigTile + threadIdx.x;
  if (t < n) atomicAdd(&cnt[code_by_tag[t]], 1);
  __syncthreads();
  if (threadIdx.x < 27) tile_counts[blockIdx.x * 27 + threadIdx.x] = cnt[threadIdx.x];
}

// Each atom's rank in its direction among the tile's atoms, in owned order
// (wavefront ballots, then a prefix over the tile's wavefronts), plus the
// tile's offset: a leaver writes its packet entries, a staying atom its
// state at its compacted index in the alt arrays (vel_alt / acc_alt with
// stride n_stay, the owned count after the exchange).
__global__ void __launch_bounds__(kMigTile)
    k_mig_write(const double4 *__restrict__ pos, const double *__restrict__ vel,
                const double *__restrict__ acc, const int *__restrict__ id,
                const int *__restrict__ code_by_tag, const int *__restrict__ src_by_tag, int n,
                HaloGeomDev g, unsigned selfmask, const int *__restrict__ tile_off,
                const long long *__restrict__ dir_base, const int *__restrict__ dir_n,
                double *__restrict__ out, double4 *__restrict__ pos_alt,
                double *__restrict__ vel_alt, double *__restrict__ acc_alt,
                int *__restrict__ id_alt) {
  constexpr int kWaves = kMigTile / 64;
  __shared__ int wcnt[kWaves][27];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t = blockIdx.x * kMigTile + threadIdx.x;
  const int code = t < n ? code_by_tag[t] : -1;
  int rank = 0;
  for (int d = 0; d < 27; ++d) {
    const unsigned long long m = __ballot(code == d);
    if (code == d)
      rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    if (lane == 0) wcnt[w][d] = __popcll(m);
  }
  __syncthreads();
  if (code < 0) return;
  int r = tile_off[blockIdx.x * 27 + code] + rank;
  for (int u = 0; u < w; ++u) r += wcnt[u][code];
  const int k = src_by_tag[t];
  double4 p = pos[k];
  const double v[3] = {vel[k], vel[n + k], vel[2 * n + k]};
  const double a[3] = {acc[k], acc[n + k], acc[2 * n + k]};
  if (code == 13) {
    const int ns = dir_n[13];
    mig_self_shift(p, g, selfmask);
    pos_alt[r] = p;
    id_alt[r] = id[k];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      vel_alt[c * ns + r] = v[c];
      acc_alt[c * ns + r] = a[c];
    }
    return;
  }
  const long long N = dir_n[code];
  double *pk = out + dir_base[code];
  const int dc[3] = {code / 9, (code / 3) % 3, code % 3};
  const double pc[3] = {p.x, p.y, p.z};
  pk[r] = (double)id[k];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double x = pc[c];
    if (g.bc[c] == 1) {  // :402-431
      if (g.gi[c] == 0 && dc[c] == 0) x += g.size[c];
      if (g.gi[c] == g.np[c] - 1 && dc[c] == 2) x -= g.size[c];
    }
    pk[N + 3 * r + c] = x;
    pk[4 * N + 3 * r + c] = v[c];
    pk[7 * N + 3 * r + c] = a[c];
  }
  pk[10 * N + r] = -1.0;  // molecule index: atoms outside molecules only
  pk[11 * N + r] = 0.0;   // atomic bond count
}

// The staying atoms, compacted in the alt arrays, back into the state
// arrays in owned order (SoA stride n), with tag = owned index.
__global__ void k_mig_adopt(int n, const double4 *__restrict__ pos_alt,
                            const double *__restrict__ vel_alt, const double *__restrict__ acc_alt,
                            const int *__restrict__ id_alt, double4 *__restrict__ pos,
                            double *__restrict__ vel, double *__restrict__ acc, int *__restrict__ id,
                            int *__restrict__ tag) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  pos[r] = pos_alt[r];
  id[r] = id_alt[r];
  tag[r] = r;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    vel[c * n + r] = vel_alt[c * n + r];
    acc[c * n + r] = acc_alt[c * n + r];
  }
}

// The state in owned order (device atom k to owned index tag[k]) in the
// alt arrays, for the merge that follows (the device keeps atoms in cell
// order between exchanges).
__global__ void k_mig_to_owned(int n, const int *__restrict__ tag, const double4 *__restrict__ pos,
                               const double *__restrict__ vel, const double *__restrict__ acc,
                               const int *__restrict__ id, double4 *__restrict__ pos_alt,
                               double *__restrict__ vel_alt, double *__restrict__ acc_alt,
                               int *__restrict__ id_alt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int t = tag[k];
  pos_alt[t] = pos[k];
  id_alt[t] = id[k];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    vel_alt[c * n + t] = vel[c * n + k];
    acc_alt[c * n + t] = acc[c * n + k];
  }
}

// Merge of the received atoms into the owned order.  The nr received
// atoms come sorted by id (rid ascending, packet entry src[k]); with
// `merge` the owned atoms are in ascending id order and each side's
// position in the merged order is its index plus the count of the other
// side's smaller ids (received after owned on equal ids); without it the
// received atoms follow the owned ones.  Owned atom c moves from the state
// arrays (stride ns) to the alt arrays (stride n = ns + nr).
__device__ inline int count_below(const int *__restrict__ a, int len, int key, bool or_equal) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key || (or_equal && a[mid] == key))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void k_mig_merge_owned(int ns, int nr, bool merge, const int *__restrict__ rid,
                                  const double4 *__restrict__ pos, const double *__restrict__ vel,
                                  const double *__restrict__ acc, const int *__restrict__ id,
                                  double4 *__restrict__ pos_alt, double *__restrict__ vel_alt,
                                  double *__restrict__ acc_alt, int *__restrict__ id_alt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ns) return;
  const int n = ns + nr;
  const int i = id[c];
  const int t = merge ? c + count_below(rid, nr, i, false) : c;
  pos_alt[t] = pos[c];
  id_alt[t] = i;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    vel_alt[d * n + t] = vel[d * ns + c];
    acc_alt[d * n + t] = acc[d * ns + c];
  }
}

// Received atom k (k-th smallest id) from its packet entry: packet p at
// in + pbase[p] with pn[p] atoms, entry e; its type from the id table.
__global__ void k_mig_place_recv(int ns, int nr, bool merge, const int *__restrict__ rid,
                                 const int2 *__restrict__ src, const double *__restrict__ in,
                                 const long long *__restrict__ pbase,
                                 const long long *__restrict__ pn,
                                 const int *__restrict__ owned_id,
                                 const int *__restrict__ type_by_id, double4 *__restrict__ pos_alt,
                                 double *__restrict__ vel_alt, double *__restrict__ acc_alt,
                                 int *__restrict__ id_alt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nr) return;
  const int n = ns + nr;
  const int i = rid[k];
  const int t = merge ? k + count_below(owned_id, ns, i, true) : ns + k;
  const int2 s = src[k];
  const long long N = pn[s.x], e = s.y;
  const double *pk = in + pbase[s.x];
  pos_alt[t] = make_double4(pk[N + 3 * e], pk[N + 3 * e + 1], pk[N + 3 * e + 2],
                            (double)type_by_id[i]);
  id_alt[t] = i;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    vel_alt[d * n + t] = pk[4 * N + 3 * e + d];
    acc_alt[d * n + t] = pk[7 * N + 3 * e + d];
  }
}

}  // namespace

void launch_mig_classify(DevState &s, const HaloGeomDev &g, unsigned selfmask, int *code_by_tag,
                         int *src_by_tag, int *tile_counts, int *moved, hipStream_t st) {
  if (s.n == 0) return;
  k_mig_classify<<<(s.n + 255) / 256, 256, 0, st>>>(s.pos, s.tag, s.n, g, selfmask, code_by_tag,
                                                     src_by_tag, moved);
  k_mig_count<<<halo_tiles(s.n), kMigTile, 0, st>>>(code_by_tag, s.n, tile_counts);
}

void launch_mig_write(DevState &s, const HaloGeomDev &g, unsigned selfmask, const int *code_by_tag,
                      const int *src_by_tag, const int *tile_off, const long long *dir_base,
                      const int *dir_n, double *out, hipStream_t st) {
  if (s.n == 0) return;
  k_mig_write<<<halo_tiles(s.n), kMigTile, 0, st>>>(s.pos, s.vel, s.acc, s.id, code_by_tag,
                                                    src_by_tag, s.n, g, selfmask, tile_off,
                                                    dir_base, dir_n, out, s.pos_alt, s.vel_alt,
                                                    s.acc_alt, s.id_alt);
}

void launch_mig_shift(DevState &s, const HaloGeomDev &g, unsigned selfmask, hipStream_t st) {
  if (s.n == 0) return;
  k_mig_shift<<<(s.n + 255) / 256, 256, 0, st>>>(s.pos, s.n, g, selfmask);
}

void launch_mig_adopt(DevState &s, int n, hipStream_t st) {
  if (n == 0) return;
  k_mig_adopt<<<(n + 255) / 256, 256, 0, st>>>(n, s.pos_alt, s.vel_alt, s.acc_alt, s.id_alt, s.pos,
                                               s.vel, s.acc, s.id, s.tag);
}

void launch_mig_to_owned(DevState &s, hipStream_t st) {
  if (s.n == 0) return;
  k_mig_to_owned<<<(s.n + 255) / 256, 256, 0, st>>>(s.n, s.tag, s.pos, s.vel, s.acc, s.id, s.pos_alt,
                                                     s.vel_alt, s.acc_alt, s.id_alt);
  k_mig_adopt<<<(s.n + 255) / 256, 256, 0, st>>>(s.n, s.pos_alt, s.vel_alt, s.acc_alt, s.id_alt,
                                                  s.pos, s.vel, s.acc, s.id, s.tag);
}

void launch_mig_merge(DevState &s, int ns, int nr, bool merge, const int *rid, const int2 *src,
                      const double *in, const long long *pbase, const long long *pn,
                      const int *type_by_id, hipStream_t st) {
  if (ns > 0)
    k_mig_merge_owned<<<(ns + 255) / 256, 256, 0, st>>>(ns, nr, merge, rid, s.pos, s.vel, s.acc,
                                                         s.id, s.pos_alt, s.vel_alt, s.acc_alt,
                                                         s.id_alt);
  if (nr > 0)
    k_mig_place_recv<<<(nr + 255) / 256, 256, 0, st>>>(ns, nr, merge, rid, src, in, pbase, pn, s.id,
                                                        type_by_id, s.pos_alt, s.vel_alt, s.acc_alt,
                                                        s.id_alt);
}

int mig_per_atom() { return kMigPerAtom; }

}  // namespace cv2
