// k_verify.hip -- an independent check of every list build (CV2_NL_VERIFY).
//
// After each build the device recomputes the pair set from scratch and
// compares it with the list the build wrote.  The check visits every
// ordered pair of owned atoms (i, j), j != i, at the one periodic image of j
// that can lie within the list cutoff (the minimum image: the box exceeds
// twice the cutoff on periodic axes), keeps the image only where the
// reference's single-domain ghost of it exists (across a face only when its
// owner lies within the ghost cutoff of that face,
// exchange_ghost_single_md_domain.cpp:70-100), and applies the FP64
// predicate (dx*dx + dy*dy) + dz*dz < ce^2 of neighborlist.cpp:261-266.  No
// cell grid, coarse filter or fixed-point coordinate is involved.  Both
// sides sum a 64-bit mix of (id_i, id_j, image) and count their pairs; a
// build whose sums or counts differ is counted as a mismatch.
//
// The state array vs (verify_state_words() words, device, per context): the
// build count last checked, whether the current launch checks, the two sums
// and counts of the last check, the builds checked and the mismatches.  The kernels run after
// every queued build and do nothing unless the device completed a new build.
// The all-pairs pass skips, per wavefront, the 256-atom tiles whose bounding
// box is farther than the cutoff from the box of the wavefront's own atoms
// under every periodic shift (a test on boxes, exact in its direction: a
// skipped tile holds no pair within the cutoff).
#include "cv2_internal.h"

namespace cv2 {

namespace {

enum : int {
  kVsLast = 0,     // flags[F_REBUILDS] at the last check
  kVsActive = 1,   // this launch checks a new build
  kVsListSum = 2,
  kVsListCnt = 3,
  kVsBruteSum = 4,
  kVsBruteCnt = 5,
  kVsChecked = 6,  // builds checked
  kVsBad = 7,      // builds whose pair set differed
};

__device__ inline unsigned long long pair_mix(unsigned long long a, unsigned long long b,
                                              unsigned long long img) {
  // splitmix64 finaliser of the packed key
  unsigned long long z = (a << 38) ^ (b << 6) ^ img;
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ inline void wave_add(unsigned long long s, unsigned long long c,
                                unsigned long long *vs) {
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, 64);
    c += __shfl_xor(c, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (s) atomicAdd(&vs[0], s);
    if (c) atomicAdd(&vs[1], c);
  }
}

__global__ void k_verify_open(const int *__restrict__ flags, unsigned long long *__restrict__ vs) {
  const unsigned long long builds = (unsigned long long)flags[F_REBUILDS];
  const bool active = flags[F_HALT] == 0 && builds != vs[kVsLast];
  vs[kVsActive] = active ? 1ull : 0ull;
  if (active)  // otherwise the sums of the last check stay readable
    for (int k = kVsListSum; k <= kVsBruteCnt; ++k) vs[k] = 0ull;
}

// The built list (32-bit entries: j | image << 26).
__global__ void k_verify_list(const uint32_t *__restrict__ nbr, const int *__restrict__ nnbr,
                              int stride, int cap, int n, const int *__restrict__ id,
                              unsigned long long *__restrict__ vs) {
  if (vs[kVsActive] == 0ull) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long s = 0ull, c = 0ull;
  if (i < n) {
    const unsigned long long idi = (unsigned long long)id[i];
    const int m = min(nnbr[i], cap);  // (a build cut at the capacity halts and is not checked)
    for (int k = 0; k < m; ++k) {
      const uint32_t e = nbr[((size_t)(k >> 2) * stride + i) * 4 + (k & 3)];
      s += pair_mix(idi, (unsigned long long)id[e & kIndexMask], e >> 26);
      ++c;
    }
  }
  wave_add(s, c, vs + kVsListSum);
}

__device__ inline double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Every pair from scratch: home atom i against all atoms, staged through
// LDS 256 at a time.
__global__ void __launch_bounds__(256)
    k_verify_brute(const double4 *__restrict__ pos, const int *__restrict__ id, int n, BoxDev b,
                   unsigned long long *__restrict__ vs) {
  if (vs[kVsActive] == 0ull) return;
  __shared__ double4 tile[256];
  __shared__ int tid[256];
  __shared__ double tbox[4][6];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const bool home = i < n;
  const double4 pi = home ? pos[i] : make_double4(0.0, 0.0, 0.0, 0.0);
  const double pic[3] = {pi.x, pi.y, pi.z};
  const unsigned long long idi = home ? (unsigned long long)id[i] : 0ull;
  const double ce = sqrt(b.ce_sq) * (1.0 + 1e-12);  // per-axis pre-tests, conservative
  // bounding box of this wavefront's home atoms
  double hlo[3], hhi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    hlo[d] = wave_min(home ? pic[d] : 1e300);
    hhi[d] = wave_max(home ? pic[d] : -1e300);
  }
  unsigned long long s = 0ull, c = 0ull;
  for (int t0 = 0; t0 < n; t0 += 256) {
    __syncthreads();
    const bool in = t0 + (int)threadIdx.x < n;
    double4 pj4 = make_double4(0.0, 0.0, 0.0, 0.0);
    if (in) {
      pj4 = pos[t0 + threadIdx.x];
      tile[threadIdx.x] = pj4;
      tid[threadIdx.x] = id[t0 + threadIdx.x];
    }
    {
      const double pjc[3] = {pj4.x, pj4.y, pj4.z};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double a = wave_min(in ? pjc[d] : 1e300), z = wave_max(in ? pjc[d] : -1e300);
        if (lane == 0) {
          tbox[wv][d] = a;
          tbox[wv][3 + d] = z;
        }
      }
    }
    __syncthreads();
    const int m = min(256, n - t0);
    // the tile's box against this wavefront's box, under the periodic shifts
    bool skip = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double a = 1e300, z = -1e300;
      for (int w = 0; w < 4; ++w) {
        a = fmin(a, tbox[w][d]);
        z = fmax(z, tbox[w][3 + d]);
      }
      double gap = 1e300;
      for (int sh = -1; sh <= 1; ++sh) {
        if (sh != 0 && !b.bc[d]) continue;
        const double lo_t = a + sh * b.L[d], hi_t = z + sh * b.L[d];
        gap = fmin(gap, fmax(0.0, fmax(lo_t - hhi[d], hlo[d] - hi_t)));
      }
      skip = skip || !(gap < ce);
    }
    if (!home || skip) continue;
    for (int u = 0; u < m; ++u) {
      const int j = t0 + u;
      const double4 pj = tile[u];
      const double pjc[3] = {pj.x, pj.y, pj.z};
      int v[3];
      bool near = j != i, exists = true;
      double g[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double dd = pjc[d] - pic[d];
        v[d] = 0;
        if (b.bc[d]) v[d] = dd > b.halfL[d] ? 1 : (dd < -b.halfL[d] ? -1 : 0);
        g[d] = v[d] != 0 ? __dsub_rn(pjc[d], (double)v[d] * b.L[d]) : pjc[d];
        near = near && fabs(g[d] - pic[d]) < ce;
        const int flag = pjc[d] < b.lo[d] + b.ghost_cut ? -1 : (pjc[d] > b.hi[d] - b.ghost_cut ? 1 : 0);
        exists = exists && (v[d] == 0 || flag == v[d]);
      }
      if (!near || !exists) continue;
      const double dx = __dsub_rn(g[0], pic[0]), dy = __dsub_rn(g[1], pic[1]),
                   dz = __dsub_rn(g[2], pic[2]);
      const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
      if (r2 < b.ce_sq) {
        const unsigned long long img =
            (unsigned long long)((v[0] + 1) | ((v[1] + 1) << 2) | ((v[2] + 1) << 4));
        s += pair_mix(idi, (unsigned long long)tid[u], img);
        ++c;
      }
    }
  }
  wave_add(s, c, vs + kVsBruteSum);
}

__global__ void k_verify_close(const int *__restrict__ flags, unsigned long long *__restrict__ vs) {
  if (vs[kVsActive] == 0ull) return;
  vs[kVsLast] = (unsigned long long)flags[F_REBUILDS];
  vs[kVsChecked] += 1ull;
  if (vs[kVsListSum] != vs[kVsBruteSum] || vs[kVsListCnt] != vs[kVsBruteCnt]) vs[kVsBad] += 1ull;
}

__global__ void k_verify_arm(const int *__restrict__ flags, unsigned long long *__restrict__ vs) {
  vs[kVsLast] = (unsigned long long)flags[F_REBUILDS];
}

}  // namespace

int verify_state_words() { return 8; }

// A fresh state checks from the next build on: the list standing now may
// have been built at positions since moved within the skin.
void launch_verify_arm(const int *flags, unsigned long long *vs, hipStream_t st) {
  k_verify_arm<<<1, 1, 0, st>>>(flags, vs);
}

void launch_verify(const DevState &s, const BoxDev &b, unsigned long long *vs, hipStream_t st) {
  k_verify_open<<<1, 1, 0, st>>>(s.flags, vs);
  if (s.n > 0) {
    const int g = (s.n + 255) / 256;
    k_verify_list<<<g, 256, 0, st>>>(s.nbr, s.nnbr, s.stride, s.cap, s.n, s.id, vs);
    k_verify_brute<<<g, 256, 0, st>>>(s.pos, s.id, s.n, b, vs);
  }
  k_verify_close<<<1, 1, 0, st>>>(s.flags, vs);
}

}  // namespace cv2
