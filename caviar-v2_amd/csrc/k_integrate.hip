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
// One kernel, k_update, applies the end of step s (after its forces) and the
// start of step s+1 (up to its positions) in a single pass over the atoms,
// and also writes the fixed-point coordinates the FP32 force kernel reads.
// The rebuild check of Neighborlist::update_is_needed (neighborlist.cpp:
// 100-148) -- any atom displaced by more than half the skin since the last
// build -- is fused into the position update and raises the device flag
// that the build kernels test.
#include <algorithm>

#include "cv2_internal.h"

namespace cv2 {

namespace {

constexpr int kBlock = 256;

inline int nblk(int n, int b) { return (n + b - 1) / b; }

// One pass of the integrator over the owned atoms (integrate_values,
// integrate_store).
template <bool LANGEVIN, bool FINISH, bool START, bool WRITEQ>
__global__ void __launch_bounds__(kBlock)
    k_update(double4 *__restrict__ pos, const double4 *__restrict__ pos_ref,
             double *__restrict__ vel, const double *__restrict__ acc, const int *__restrict__ id,
             int4 *__restrict__ q, int n, StepParams p, int *__restrict__ flags) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || flags[F_HALT]) return;
  const double a[3] = {acc[i], acc[n + i], acc[2 * n + i]};
  double v[3] = {vel[i], vel[n + i], vel[2 * n + i]};
  double4 x = make_double4(0.0, 0.0, 0.0, 0.0), r = x;
  if (START) {
    x = pos[i];
    r = pos_ref[i];
  }
  integrate_values<LANGEVIN, FINISH, START, WRITEQ>(x, v, LANGEVIN ? id[i] : 0, a, p);
  integrate_store<START, WRITEQ>(i, n, x, v, r, pos, vel, q, p, flags);
}

// Displacement check alone (Neighborlist::update_is_needed), owned atoms
// and host-halo ghosts.
__global__ void __launch_bounds__(kBlock)
    k_check_disp(const double4 *__restrict__ pos, const double4 *__restrict__ pos_ref, int n,
                 int ntot, StepParams p, int *__restrict__ flags) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  bool t = false;
  if (i < ntot) {
    const double4 x = pos[i], r = pos_ref[i];
    const double dx = x.x - r.x, dy = x.y - r.y, dz = x.z - r.z;
    // the wrap / ghost-band triggers concern owned atoms
    t = dx * dx + dy * dy + dz * dz > p.thresh_sq || (p.ref_rebuild && i < n && ref_trigger(x, r, p));
  }
  // one atomic per wavefront: after a new halo every ghost may trigger
  if (__ballot(t) != 0ull && (threadIdx.x & 63) == 0) atomicOr(&flags[F_REBUILD], 1);
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

// Ensemble velocity Verlet: the same half steps as k_update, fixed-point
// coordinates and rebuild test per system.
template <bool FINISH, bool START>
__global__ void __launch_bounds__(kBlock)
    k_ens_update(double4 *__restrict__ pos, const double4 *__restrict__ pos_ref,
                 double *__restrict__ vel, const double *__restrict__ acc, int4 *__restrict__ q,
                 int n, EnsDev e, double dt, double thresh_sq, int *__restrict__ flags) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double a[3] = {acc[i], acc[n + i], acc[2 * n + i]};
  ens_step_atom<FINISH, START>(i, n, e, pos, pos_ref, vel, a, q, dt, thresh_sq, flags);
}

// Per-system (kinetic energy, potential energy, virial): one block per
// system, fixed order.
__global__ void __launch_bounds__(kBlock)
    k_ens_thermo(const double *__restrict__ vel, int n, EnsDev e, const LJDev *__restrict__ lj,
                 const long long *__restrict__ eatom, double *__restrict__ out) {
  __shared__ double sk[kBlock];
  __shared__ long long sv[kBlock], sp[kBlock];
  const int sy = blockIdx.x;
  const int a0 = e.atom_off[sy], a1 = e.atom_off[sy + 1];
  double k2 = 0.0;
  long long v = 0, p = 0;
  for (int i = a0 + threadIdx.x; i < a1; i += kBlock) {
    const double vx = vel[i], vy = vel[n + i], vz = vel[2 * n + i];
    k2 += vx * vx + vy * vy + vz * vz;
    v += eatom[2 * (size_t)i];
    p += eatom[2 * (size_t)i + 1];
  }
  sk[threadIdx.x] = k2;
  sv[threadIdx.x] = v;
  sp[threadIdx.x] = p;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sk[threadIdx.x] += sk[threadIdx.x + o];
      sv[threadIdx.x] += sv[threadIdx.x + o];
      sp[threadIdx.x] += sp[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[3 * sy] = 0.5 * lj->mass[0] * sk[0];
    out[3 * sy + 1] = (double)sp[0] / kEnergyScale;
    out[3 * sy + 2] = (double)sv[0] / kEnergyScale;
  }
}

template <bool L, bool F, bool S>
void update_q(DevState &s, const StepParams &p, bool writeq, hipStream_t st) {
  const int g = nblk(s.n, kBlock);
  if (writeq)
    k_update<L, F, S, true><<<g, kBlock, 0, st>>>(s.pos, s.pos_ref, s.vel, s.acc, s.id, s.q, s.n, p, s.flags);
  else
    k_update<L, F, S, false><<<g, kBlock, 0, st>>>(s.pos, s.pos_ref, s.vel, s.acc, s.id, s.q, s.n, p, s.flags);
}

}  // namespace

StepParams step_params(const BoxDev &b, const UpdateArgs &u) {
  StepParams p{};
  p.dt = u.dt;
  p.la = u.la;
  p.lb = u.lb;
  p.lc = u.lc;
  p.seed = u.seed;
  p.step = u.step;
  p.thresh_sq = b.thresh_sq;
  for (int d = 0; d < 3; ++d) {
    p.lo[d] = b.lo[d];
    p.qscale[d] = b.qscale[d];
    const bool per = b.bc[d] == 1;
    const double inf = __builtin_inf();
    p.wlo[d] = per ? b.lo[d] : -inf;
    p.whi[d] = per ? b.hi[d] : inf;
    p.blo[d] = per ? b.lo[d] + b.ghost_cut : -inf;
    p.bhi[d] = per ? b.hi[d] - b.ghost_cut : inf;
    p.hi[d] = b.hi[d];
    p.wL[d] = b.wrap[d] == 1 ? b.L[d] : 0.0;
  }
  p.ref_rebuild = u.ref_rebuild ? 1 : 0;
  p.wrap = u.wrap_owned ? 1 : 0;
  return p;
}

void launch_update(DevState &s, const BoxDev &b, const UpdateArgs &u, hipStream_t st) {
  if (s.n == 0 || (!u.finish && !u.start)) return;
  const StepParams p = step_params(b, u);
#define CV2_UPD(L, F, S)                             \
  if (u.langevin == L && u.finish == F && u.start == S) { \
    update_q<L, F, S>(s, p, u.writeq, st);           \
    return;                                          \
  }
  CV2_UPD(false, true, true)
  CV2_UPD(false, true, false)
  CV2_UPD(false, false, true)
  CV2_UPD(true, true, true)
  CV2_UPD(true, true, false)
  CV2_UPD(true, false, true)
#undef CV2_UPD
}

void launch_update_ens(DevState &s, const EnsDev &e, bool finish, bool start, double dt,
                       double thresh_sq, hipStream_t st) {
  if (s.n == 0 || (!finish && !start)) return;
  const int g = nblk(s.n, kBlock);
  if (finish && start)
    k_ens_update<true, true><<<g, kBlock, 0, st>>>(s.pos, s.pos_ref, s.vel, s.acc, s.q, s.n, e, dt, thresh_sq, s.flags);
  else if (finish)
    k_ens_update<true, false><<<g, kBlock, 0, st>>>(s.pos, s.pos_ref, s.vel, s.acc, s.q, s.n, e, dt, thresh_sq, s.flags);
  else
    k_ens_update<false, true><<<g, kBlock, 0, st>>>(s.pos, s.pos_ref, s.vel, s.acc, s.q, s.n, e, dt, thresh_sq, s.flags);
}

void launch_thermo_ens(const DevState &s, const EnsDev &e, const LJDev *lj_dev,
                       const long long *eatom, double *out, hipStream_t st) {
  if (e.nsys == 0) return;
  k_ens_thermo<<<e.nsys, kBlock, 0, st>>>(s.vel, s.n, e, lj_dev, eatom, out);
}

// Velocity rescale of the NVE constraint (Nve::apply_thermostat,
// nve.cpp:186-189: every owned velocity times lambda).
__global__ void __launch_bounds__(kBlock) k_scale_vel(double *__restrict__ vel, int64_t m, double lambda) {
  for (int64_t k = blockIdx.x * (int64_t)kBlock + threadIdx.x; k < m; k += (int64_t)gridDim.x * kBlock)
    vel[k] *= lambda;
}

void launch_scale_vel(DevState &s, double lambda, hipStream_t st) {
  const int64_t m = 3 * (int64_t)s.n;
  if (m == 0) return;
  const int grid = (int)std::min<int64_t>((m + kBlock - 1) / kBlock, 4096);
  k_scale_vel<<<grid, kBlock, 0, st>>>(s.vel, m, lambda);
}

void launch_check_disp(DevState &s, const BoxDev &b, bool ref_rebuild, hipStream_t st) {
  const int ntot = s.n + s.ng;
  if (ntot == 0) return;
  UpdateArgs u{};
  u.ref_rebuild = ref_rebuild;
  k_check_disp<<<nblk(ntot, kBlock), kBlock, 0, st>>>(s.pos, s.pos_ref, s.n, ntot, step_params(b, u),
                                                      s.flags);
}

// Writes (sum m v^2, sum m v (3), sum m) to reduce_out[2..6].
void launch_thermo(DevState &s, const LJDev *lj_dev, hipStream_t st) {
  if (s.n == 0) return;
  const int nb = nblk(s.n, kBlock);
  k_thermo<<<nb, kBlock, 0, st>>>(s.pos, s.vel, s.n, lj_dev, s.partials);
  k_reduce5<<<1, 1024, 0, st>>>(s.partials, nb, s.reduce_out + 2);
}

}  // namespace cv2
