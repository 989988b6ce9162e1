// k_integrate.hip -- velocity-Verlet and Langevin updates, rebuild check,
// thermodynamic reductions.  State is FP64 throughout.
//
// Velocity Verlet, md_simulator.cpp:617-729, is applied in its
// kick-drift-kick form (v += a dt/2; x += v dt; forces; v += a dt/2), the
// same trajectory as the reference's x += v dt + a dt^2/2, v += (a+a_old) dt/2
// without storing a_old.
//
// Langevin, md_simulator.cpp:731-869 with a, b, c of :398-405:
//   v += a dt/2 + b eta;  x += c v;  forces;  v = a_L v + b eta + a dt/2
// with the same eta in both halves.  eta is drawn from Philox4x32-10
// (Salmon, Moraes, Dror, Shaw, SC'11) keyed by (seed, atom id, step) and
// Box-Muller, so it needs no state and is independent of atom order.
//
// The rebuild check of Neighborlist::update_is_needed (neighborlist.cpp:
// 100-148) -- any atom displaced by more than half the skin since the last
// build -- is fused into the position update and raises the device flag
// that the build kernels test.
#include "cv2_internal.h"

namespace cv2 {

namespace {

constexpr int kBlock = 256;

inline int nblk(int n, int b) { return (n + b - 1) / b; }

__device__ inline void check_disp(const double4 &p, const double4 &r, double thresh_sq,
                                  int *flags) {
  const double dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
  if (dx * dx + dy * dy + dz * dz > thresh_sq) atomicOr(&flags[F_REBUILD], 1);
}

__global__ void __launch_bounds__(kBlock)
    k_kick_drift(double4 *__restrict__ pos, const double4 *__restrict__ pos_ref,
                 double *__restrict__ vel, const double *__restrict__ acc, int n, double dt,
                 double thresh_sq, int *__restrict__ flags) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double hdt = 0.5 * dt;
  double4 p = pos[i];
  const double vx = vel[i] + acc[i] * hdt;
  const double vy = vel[n + i] + acc[n + i] * hdt;
  const double vz = vel[2 * n + i] + acc[2 * n + i] * hdt;
  p.x += vx * dt;
  p.y += vy * dt;
  p.z += vz * dt;
  pos[i] = p;
  vel[i] = vx;
  vel[n + i] = vy;
  vel[2 * n + i] = vz;
  check_disp(p, pos_ref[i], thresh_sq, flags);
}

__global__ void __launch_bounds__(kBlock)
    k_kick(double *__restrict__ vel, const double *__restrict__ acc, int n, double dt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double hdt = 0.5 * dt;
  vel[i] += acc[i] * hdt;
  vel[n + i] += acc[n + i] * hdt;
  vel[2 * n + i] += acc[2 * n + i] * hdt;
}

// ---- Philox4x32-10 ----
struct U4 {
  uint32_t x, y, z, w;
};

__device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
    const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Three standard normals for (seed, id, step).
__device__ inline void gauss3(uint64_t seed, int id, uint64_t step, double &g0, double &g1,
                              double &g2) {
  const U4 r = philox4x32_10(U4{(uint32_t)step, (uint32_t)(step >> 32), (uint32_t)id, 0x4c414e47u},
                             (uint32_t)seed, (uint32_t)(seed >> 32));
  constexpr double k2m32 = 2.3283064365386963e-10;
  const double u0 = ((double)r.x + 0.5) * k2m32, u1 = (double)r.y * k2m32;
  const double u2 = ((double)r.z + 0.5) * k2m32, u3 = (double)r.w * k2m32;
  const double ra = sqrt(-2.0 * log(u0)), rb = sqrt(-2.0 * log(u2));
  double sa, ca, sb, cb;
  sincospi(2.0 * u1, &sa, &ca);
  sincospi(2.0 * u3, &sb, &cb);
  g0 = ra * ca;
  g1 = ra * sa;
  g2 = rb * cb;
  (void)sb;
}

__global__ void __launch_bounds__(kBlock)
    k_langevin_1(double4 *__restrict__ pos, const double4 *__restrict__ pos_ref,
                 double *__restrict__ vel, const double *__restrict__ acc,
                 const int *__restrict__ id, int n, double dt, double lb, double lc,
                 uint64_t seed, uint64_t step, double thresh_sq, int *__restrict__ flags) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double ex, ey, ez;
  gauss3(seed, id[i], step, ex, ey, ez);
  // md_simulator.cpp:778: v += 0.5*a*dt + b*eta
  const double vx = vel[i] + (acc[i] * 0.5 * dt + ex * lb);
  const double vy = vel[n + i] + (acc[n + i] * 0.5 * dt + ey * lb);
  const double vz = vel[2 * n + i] + (acc[2 * n + i] * 0.5 * dt + ez * lb);
  double4 p = pos[i];
  // :835: x += v*c
  p.x += vx * lc;
  p.y += vy * lc;
  p.z += vz * lc;
  pos[i] = p;
  vel[i] = vx;
  vel[n + i] = vy;
  vel[2 * n + i] = vz;
  check_disp(p, pos_ref[i], thresh_sq, flags);
}

__global__ void __launch_bounds__(kBlock)
    k_langevin_2(double *__restrict__ vel, const double *__restrict__ acc,
                 const int *__restrict__ id, int n, double dt, double la, double lb,
                 uint64_t seed, uint64_t step) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double ex, ey, ez;
  gauss3(seed, id[i], step, ex, ey, ez);
  // :860: v = a*v + b*eta + 0.5*acc*dt
  vel[i] = vel[i] * la + ex * lb + acc[i] * 0.5 * dt;
  vel[n + i] = vel[n + i] * la + ey * lb + acc[n + i] * 0.5 * dt;
  vel[2 * n + i] = vel[2 * n + i] * la + ez * lb + acc[2 * n + i] * 0.5 * dt;
}

// Per-block sums of (m v.v, m vx, m vy, m vz, m).
__global__ void __launch_bounds__(kBlock)
    k_thermo(const double4 *__restrict__ pos, const double *__restrict__ vel, int n,
             const LJDev *__restrict__ lj, double *__restrict__ partials) {
  __shared__ double sh[5][kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[5] = {0, 0, 0, 0, 0};
  if (i < n) {
    const double m = lj->mass[(int)pos[i].w];
    const double vx = vel[i], vy = vel[n + i], vz = vel[2 * n + i];
    v[0] = m * (vx * vx + vy * vy + vz * vz);
    v[1] = m * vx;
    v[2] = m * vy;
    v[3] = m * vz;
    v[4] = m;
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    double x = v[c];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (l == 0) sh[c][w] = x;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double s = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) s += sh[threadIdx.x][k];
    partials[(size_t)blockIdx.x * 5 + threadIdx.x] = s;
  }
}

__global__ void __launch_bounds__(1024) k_reduce5(const double *__restrict__ partials, int nblk,
                                                  double *__restrict__ out) {
  __shared__ double sh[1024];
  for (int c = 0; c < 5; ++c) {
    double s = 0.0;
    for (int k = threadIdx.x; k < nblk; k += 1024) s += partials[(size_t)k * 5 + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
    __syncthreads();
  }
}

}  // namespace

void launch_kick_drift(DevState &s, const BoxDev &b, double dt, hipStream_t st) {
  if (s.n == 0) return;
  k_kick_drift<<<nblk(s.n, kBlock), kBlock, 0, st>>>(s.pos, s.pos_ref, s.vel, s.acc, s.n, dt,
                                                      b.thresh_sq, s.flags);
}

void launch_kick(DevState &s, double dt, hipStream_t st) {
  if (s.n == 0) return;
  k_kick<<<nblk(s.n, kBlock), kBlock, 0, st>>>(s.vel, s.acc, s.n, dt);
}

void launch_langevin_1(DevState &s, const BoxDev &b, const LJDev *, double dt, double lb,
                       double lc, uint64_t seed, uint64_t step, hipStream_t st) {
  if (s.n == 0) return;
  k_langevin_1<<<nblk(s.n, kBlock), kBlock, 0, st>>>(s.pos, s.pos_ref, s.vel, s.acc, s.id, s.n,
                                                      dt, lb, lc, seed, step, b.thresh_sq,
                                                      s.flags);
}

void launch_langevin_2(DevState &s, const LJDev *, double dt, double la, double lb,
                       uint64_t seed, uint64_t step, hipStream_t st) {
  if (s.n == 0) return;
  k_langevin_2<<<nblk(s.n, kBlock), kBlock, 0, st>>>(s.vel, s.acc, s.id, s.n, dt, la, lb, seed,
                                                      step);
}

// Writes (sum m v^2, sum m v (3), sum m) to reduce_out[2..6].
void launch_thermo(DevState &s, const LJDev *lj_dev, hipStream_t st) {
  if (s.n == 0) return;
  const int nb = nblk(s.n, kBlock);
  k_thermo<<<nb, kBlock, 0, st>>>(s.pos, s.vel, s.n, lj_dev, s.partials);
  k_reduce5<<<1, 1024, 0, st>>>(s.partials, nb, s.reduce_out + 2);
}

}  // namespace cv2
