// k_lj.hip -- Lennard-Jones accelerations, virial and energy on the device.
//
// Force law of force_field::Lj::calculate_acceleration (lj.cpp:393-517):
//   forceCoef = 4 eps (-12 rho^12/r^2 + 6 rho^6/r^2 + 12 rho_c^12/rc^2 - 6 rho_c^6/rc^2),
//   rho^2 = sigma^2/r^2, rho_c^2 = sigma^2/rc^2, pairs with r^2 > rc^2 skipped,
//   F_i = forceCoef * (r_j - r_i), a_i += F_i / m_i, W += -forceCoef r^2 once per pair.
// The potential energy is the U(r) whose U'(r)/r is forceCoef, with U(rc)=0
// (the reference computes no energy):
//   U = 4 eps [(rho^12 - rho^6) + (6 rho_c^12 - 3 rho_c^6) r^2/rc^2 - (7 rho_c^12 - 4 rho_c^6)].
//
// Both kernels walk the full Verlet list (each pair from both sides), one
// thread per owned atom, so every atom's sum is private to one thread: no
// atomics on forces.  Virial and energy count a pair once, from the side
// with the lower sorted index (host-halo ghosts: the lower id, as
// lj.cpp:487 does).
//
// k_lj_f64     FP64 pair arithmetic, FP64 accumulation.
// k_lj_f32fx   FP32 pair arithmetic on separations formed exactly from
//              fixed-point coordinates; per-pair force components are
//              converted to 64-bit fixed point (2^-32 units) inside one FP64
//              FMA and summed as integers, so the per-atom sum is exact and
//              independent of summation order.  Energy and virial use 2^-24
//              units and integer atomics for the global sums.
#include "cv2_internal.h"

namespace cv2 {

namespace {

constexpr int kBlock = 256;
// 1.5 * 2^52: adding it to a double of magnitude < 2^51 rounds to an integer
// held in the low mantissa bits, so bits(t) - bits(kMagic) is that integer.
constexpr double kMagic = 6755399441055744.0;
constexpr long long kMagicBits = 0x4338000000000000LL;

// Blocks are dealt round-robin to the 8 XCDs; remap so that each XCD gets a
// contiguous range of (spatially sorted) atoms and shares its L2 working set.
__device__ inline int xcd_block(int nblk_real) {
  const int nb = gridDim.x;  // multiple of 8
  const int per = nb >> 3;
  const int b = blockIdx.x;
  const int logical = (b & 7) * per + (b >> 3);
  return logical < nblk_real ? logical : -1;
}

__device__ inline double block_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r += sh[k];
  return r;
}

__device__ inline long long block_sum_i64(long long v, long long *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  long long r = 0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r += sh[k];
  return r;
}

// The list is stored in groups of four entries per atom,
// nbr[((k / 4) * stride + i) * 4 + k % 4], padded with the atom itself (r = 0,
// skipped by the r^2 > 0 test): a lane reads its next four entries as one
// 16-byte load and the loads of a wave are contiguous.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ inline u32x4 load_group(const uint32_t *__restrict__ nbr, size_t stride, int i, int k) {
  return __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(nbr) + (size_t)(k >> 2) * stride + i);
}

template <bool MULTI, bool ENERGY, bool HALO>
__global__ void __launch_bounds__(kBlock)
    k_lj_f64(const double4 *__restrict__ pos, const uint32_t *__restrict__ nbr,
             const int *__restrict__ nnbr, int stride, int n, const int *__restrict__ id,
             const LJDev *__restrict__ lj, BoxDev b, double *__restrict__ acc,
             double *__restrict__ partials, int *__restrict__ flags, int nblk_real) {
  __shared__ double sh[kBlock / 64];
  const int blk = xcd_block(nblk_real);
  if (blk < 0) return;
  const int i = blk * kBlock + threadIdx.x;
  double vir = 0.0, pe = 0.0;
  if (i < n) {
    const double4 pi = pos[i];
    const int ti = MULTI ? (int)pi.w : 0;
    const int nt = MULTI ? lj->nt : 1;
    // single-type parameters stay in registers for the whole loop
    const double rc2_0 = lj->rc2[0], sig2_0 = lj->sig2[0], eps4_0 = lj->eps4[0], fsh_0 = lj->fshift[0];
    const int cp = (nnbr[i] + 3) & ~3;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int id_i = (ENERGY && HALO) ? id[i] : 0;
    for (int k = 0; k < cp; k += 4) {
      const u32x4 g = load_group(nbr, (size_t)stride, i, k);
      const uint32_t e[4] = {g.x, g.y, g.z, g.w};
      double4 pj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) pj[u] = pos[e[u] & kIndexMask];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = (int)(e[u] & kIndexMask);
        double gx = pj[u].x, gy = pj[u].y, gz = pj[u].z;
        const uint32_t img = e[u] >> 26;
        if (img != kNoImage) {
          gx -= (double)((int)(img & 3u) - 1) * b.L[0];
          gy -= (double)((int)((img >> 2) & 3u) - 1) * b.L[1];
          gz -= (double)((int)(img >> 4) - 1) * b.L[2];
        }
        const double dx = gx - pi.x, dy = gy - pi.y, dz = gz - pi.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const int tp = MULTI ? ti * nt + (int)pj[u].w : 0;
        const double rc2 = MULTI ? lj->rc2[tp] : rc2_0;
        if (r2 <= rc2 && r2 > 0.0) {
          // 1/r^2: hardware reciprocal and two Newton steps (about 1 ulp)
          double r2i = __builtin_amdgcn_rcp(r2);
          r2i = fma(r2i, fma(-r2, r2i, 1.0), r2i);
          r2i = fma(r2i, fma(-r2, r2i, 1.0), r2i);
          const double s2 = (MULTI ? lj->sig2[tp] : sig2_0) * r2i;
          const double s6 = s2 * s2 * s2;
          const double eps4 = MULTI ? lj->eps4[tp] : eps4_0;
          const double fc = eps4 * (r2i * (s6 * (-12.0 * s6 + 6.0)) + (MULTI ? lj->fshift[tp] : fsh_0));
          fx += fc * dx;
          fy += fc * dy;
          fz += fc * dz;
          if (ENERGY) {
            const bool once = HALO && j >= n ? (id_i < id[j]) : (i < j);
            if (once) {
              vir -= fc * r2;
              pe += eps4 * ((s6 * s6 - s6) + lj->ecut[tp] * r2 - lj->ecut2[tp]);
            }
          }
        }
      }
    }
    const double mi = lj->mass_inv[ti];
    const double ax = fx * mi, ay = fy * mi, az = fz * mi;
    acc[i] = ax;
    acc[n + i] = ay;
    acc[2 * n + i] = az;
    if (isnan(ax) || isnan(ay) || isnan(az)) atomicOr(&flags[F_NAN], 1);
  }
  if (ENERGY) {
    const double v = block_sum(vir, sh);
    __syncthreads();
    const double u = block_sum(pe, sh);
    if (threadIdx.x == 0) {
      partials[2 * blk] = v;
      partials[2 * blk + 1] = u;
    }
  }
}

template <bool MULTI, bool ENERGY, bool HALO>
__global__ void __launch_bounds__(kBlock)
    k_lj_f32fx(const int4 *__restrict__ q, const uint32_t *__restrict__ nbr,
               const int *__restrict__ nnbr, int stride, int n, const int *__restrict__ id,
               const LJDev *__restrict__ lj, BoxDev b, double *__restrict__ acc,
               long long *__restrict__ partials, int *__restrict__ flags, int nblk_real) {
  __shared__ long long sh[kBlock / 64];
  const int blk = xcd_block(nblk_real);
  if (blk < 0) return;
  const int i = blk * kBlock + threadIdx.x;
  // length of one residual unit (2^-9 of the fixed-point quantum) per axis
  const double qu0 = b.qinv[0] / kResidScale, qu1 = b.qinv[1] / kResidScale,
               qu2 = b.qinv[2] / kResidScale;
  long long vir = 0, pe = 0;
  if (i < n) {
    const int4 qi = q[i];
    const int ti = MULTI ? (qi.w & 7) : 0;
    const int nt = MULTI ? lj->nt : 1;
    const double rc2_0 = lj->rc2[0], sig2_0 = lj->sig2[0], r2ovf_0 = lj->r2ovf[0];
    const float eps4_0 = lj->eps4f[0], fsh_0 = lj->fshiftf[0];
    const int rix = q_resid(qi.w, 0), riy = q_resid(qi.w, 1), riz = q_resid(qi.w, 2);
    const int cp = (nnbr[i] + 3) & ~3;
    unsigned long long sx = 0, sy = 0, sz = 0;
    int terms = 0;
    bool ovf = false;
    const int id_i = (ENERGY && HALO) ? id[i] : 0;
    for (int k = 0; k < cp; k += 4) {
      const u32x4 g = load_group(nbr, (size_t)stride, i, k);
      const uint32_t e[4] = {g.x, g.y, g.z, g.w};
      int4 qj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) qj[u] = q[e[u] & kIndexMask];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = (int)(e[u] & kIndexMask);
        // exact separation in residual units: quanta * 512 + residual difference
        const double dx = fma((double)(qj[u].x - qi.x), kResidScale,
                              (double)(q_resid(qj[u].w, 0) - rix)) * qu0;
        const double dy = fma((double)(qj[u].y - qi.y), kResidScale,
                              (double)(q_resid(qj[u].w, 1) - riy)) * qu1;
        const double dz = fma((double)(qj[u].z - qi.z), kResidScale,
                              (double)(q_resid(qj[u].w, 2) - riz)) * qu2;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const int tp = MULTI ? ti * nt + (qj[u].w & 7) : 0;
        if (r2 <= (MULTI ? lj->rc2[tp] : rc2_0) && r2 > 0.0) {
          // r^-2 and (sigma^2/r^2)^3 in FP64, each rounded once to FP32;
          // the force polynomial in FP32
          double r2i = __builtin_amdgcn_rcp(r2);
          r2i = fma(r2i, fma(-r2, r2i, 1.0), r2i);
          const double s2 = (MULTI ? lj->sig2[tp] : sig2_0) * r2i;
          const float s6 = (float)(s2 * s2 * s2);
          const float r2if = (float)r2i;
          const float eps4 = MULTI ? lj->eps4f[tp] : eps4_0;
          const float fc = eps4 * (r2if * (s6 * (-12.0f * s6 + 6.0f)) + (MULTI ? lj->fshiftf[tp] : fsh_0));
          // |F| = |fc| r stays below 2^19 for r^2 >= r2ovf (fixed-point range)
          ovf |= r2 < (MULTI ? lj->r2ovf[tp] : r2ovf_0);
          // F * 2^32 rounded to an integer inside one FP64 FMA
          const double fcd = (double)fc;
          const double fcs = fcd * kForceScale;
          sx += (unsigned long long)__double_as_longlong(fma(fcs, dx, kMagic));
          sy += (unsigned long long)__double_as_longlong(fma(fcs, dy, kMagic));
          sz += (unsigned long long)__double_as_longlong(fma(fcs, dz, kMagic));
          ++terms;
          if (ENERGY) {
            const bool once = HALO && j >= n ? (id_i < id[j]) : (i < j);
            if (once) {
              vir += __double_as_longlong(fma(-fcd * r2, kEnergyScale, kMagic)) - kMagicBits;
              const float u = eps4 * ((s6 * s6 - s6) + (float)(lj->ecut[tp] * r2) - (float)lj->ecut2[tp]);
              pe += __double_as_longlong(fma((double)u, kEnergyScale, kMagic)) - kMagicBits;
            }
          }
        }
      }
    }
    const unsigned long long base = (unsigned long long)terms * (unsigned long long)kMagicBits;
    const double mi = lj->mass_inv[ti] * (1.0 / kForceScale);
    acc[i] = (double)(long long)(sx - base) * mi;
    acc[n + i] = (double)(long long)(sy - base) * mi;
    acc[2 * n + i] = (double)(long long)(sz - base) * mi;
    if (ovf) atomicOr(&flags[F_FXOVF], 1);
  }
  if (ENERGY) {
    const long long v = block_sum_i64(vir, sh);
    __syncthreads();
    const long long u = block_sum_i64(pe, sh);
    if (threadIdx.x == 0) {  // per-block partials, summed by k_reduce_i64
      partials[2 * blk] = v;
      partials[2 * blk + 1] = u;
    }
  }
}

// Sum of per-block fixed-point partials (integer: exact in any order).
__global__ void __launch_bounds__(1024) k_reduce_i64(const long long *__restrict__ partials,
                                                     int nblk, long long *__restrict__ out) {
  __shared__ long long sh[1024];
  for (int c = 0; c < 2; ++c) {
    long long t = 0;
    for (int k = threadIdx.x; k < nblk; k += 1024) t += partials[(size_t)k * 2 + c];
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
    __syncthreads();
  }
}

// Ensemble form of k_lj_f32fx (single species): each atom reads the
// fixed-point scale of its own system; ENERGY writes per-atom fixed-point
// virial and energy (each pair once: i < j) for the per-system sums.
template <int BLOCK, bool ENERGY>
__global__ void __launch_bounds__(BLOCK)
    k_lj_ens(const int4 *__restrict__ q, const uint32_t *__restrict__ nbr,
             const int *__restrict__ nnbr, int stride, int n, EnsDev e,
             const LJDev *__restrict__ lj, double *__restrict__ acc, long long *__restrict__ eatom,
             int *__restrict__ flags) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const BoxDev &b = e.box[e.sys[i]];
  const double qu0 = b.qinv[0] / kResidScale, qu1 = b.qinv[1] / kResidScale,
               qu2 = b.qinv[2] / kResidScale;
  const int4 qi = q[i];
  const double rc2_0 = lj->rc2[0], sig2_0 = lj->sig2[0], r2ovf_0 = lj->r2ovf[0];
  const float eps4_0 = lj->eps4f[0], fsh_0 = lj->fshiftf[0];
  const double ecut = lj->ecut[0], ecut2 = lj->ecut2[0];
  const int rix = q_resid(qi.w, 0), riy = q_resid(qi.w, 1), riz = q_resid(qi.w, 2);
  const int cp = (nnbr[i] + 3) & ~3;
  unsigned long long sx = 0, sy = 0, sz = 0;
  long long vir = 0, pe = 0;
  int terms = 0;
  bool ovf = false;
  for (int k = 0; k < cp; k += 4) {
    const u32x4 g = load_group(nbr, (size_t)stride, i, k);
    const uint32_t ent[4] = {g.x, g.y, g.z, g.w};
    int4 qj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) qj[u] = q[ent[u] & kIndexMask];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = (int)(ent[u] & kIndexMask);
      const double dx = fma((double)(qj[u].x - qi.x), kResidScale, (double)(q_resid(qj[u].w, 0) - rix)) * qu0;
      const double dy = fma((double)(qj[u].y - qi.y), kResidScale, (double)(q_resid(qj[u].w, 1) - riy)) * qu1;
      const double dz = fma((double)(qj[u].z - qi.z), kResidScale, (double)(q_resid(qj[u].w, 2) - riz)) * qu2;
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 <= rc2_0 && r2 > 0.0) {
        double r2i = __builtin_amdgcn_rcp(r2);
        r2i = fma(r2i, fma(-r2, r2i, 1.0), r2i);
        const double s2 = sig2_0 * r2i;
        const float s6 = (float)(s2 * s2 * s2);
        const float r2if = (float)r2i;
        const float fc = eps4_0 * (r2if * (s6 * (-12.0f * s6 + 6.0f)) + fsh_0);
        ovf |= r2 < r2ovf_0;
        const double fcd = (double)fc;
        const double fcs = fcd * kForceScale;
        sx += (unsigned long long)__double_as_longlong(fma(fcs, dx, kMagic));
        sy += (unsigned long long)__double_as_longlong(fma(fcs, dy, kMagic));
        sz += (unsigned long long)__double_as_longlong(fma(fcs, dz, kMagic));
        ++terms;
        if (ENERGY && i < j) {
          vir += __double_as_longlong(fma(-fcd * r2, kEnergyScale, kMagic)) - kMagicBits;
          const float u = eps4_0 * ((s6 * s6 - s6) + (float)(ecut * r2) - (float)ecut2);
          pe += __double_as_longlong(fma((double)u, kEnergyScale, kMagic)) - kMagicBits;
        }
      }
    }
  }
  const unsigned long long base = (unsigned long long)terms * (unsigned long long)kMagicBits;
  const double mi = lj->mass_inv[0] * (1.0 / kForceScale);
  acc[i] = (double)(long long)(sx - base) * mi;
  acc[n + i] = (double)(long long)(sy - base) * mi;
  acc[2 * n + i] = (double)(long long)(sz - base) * mi;
  if (ovf) atomicOr(&flags[F_FXOVF], 1);
  if (ENERGY) {
    eatom[2 * (size_t)i] = vir;
    eatom[2 * (size_t)i + 1] = pe;
  }
}

// Fixed-order sum of per-block partials (deterministic).
__global__ void __launch_bounds__(1024) k_reduce_partials(const double *__restrict__ partials,
                                                         int nblk, int width,
                                                         double *__restrict__ out) {
  __shared__ double sh[1024];
  for (int c = 0; c < width; ++c) {
    double s = 0.0;
    for (int k = threadIdx.x; k < nblk; k += 1024) s += partials[(size_t)k * width + c];
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

template <bool MULTI, bool ENERGY, bool HALO>
void lj_dispatch(DevState &s, const BoxDev &b, const LJDev *lj_dev, int precision, hipStream_t st) {
  const int real = (s.n + kBlock - 1) / kBlock;
  const int grid = ((real + 7) / 8) * 8;
  if (precision == 0) {
    k_lj_f64<MULTI, ENERGY, HALO><<<grid, kBlock, 0, st>>>(s.pos, s.nbr, s.nnbr, s.stride, s.n,
                                                           s.id, lj_dev, b, s.acc, s.partials,
                                                           s.flags, real);
    if (ENERGY) k_reduce_partials<<<1, 1024, 0, st>>>(s.partials, real, 2, s.reduce_out);
  } else {
    long long *part = reinterpret_cast<long long *>(s.partials);
    k_lj_f32fx<MULTI, ENERGY, HALO><<<grid, kBlock, 0, st>>>(s.q, s.nbr, s.nnbr, s.stride, s.n,
                                                             s.id, lj_dev, b, s.acc, part,
                                                             s.flags, real);
    if (ENERGY) k_reduce_i64<<<1, 1024, 0, st>>>(part, real, s.fx_acc);
  }
}

}  // namespace

void launch_forces_ens(DevState &s, const EnsDev &e, const LJDev *lj_dev, int block, bool energy,
                       long long *eatom, hipStream_t st) {
  if (s.n == 0) return;
#define CV2_LJE(B)                                                                            \
  if (block == B) {                                                                           \
    const int g = (s.n + B - 1) / B;                                                          \
    if (energy)                                                                               \
      k_lj_ens<B, true><<<g, B, 0, st>>>(s.q, s.nbr, s.nnbr, s.stride, s.n, e, lj_dev, s.acc, \
                                          eatom, s.flags);                                    \
    else                                                                                      \
      k_lj_ens<B, false><<<g, B, 0, st>>>(s.q, s.nbr, s.nnbr, s.stride, s.n, e, lj_dev,       \
                                           s.acc, eatom, s.flags);                            \
    return;                                                                                   \
  }
  CV2_LJE(64)
  CV2_LJE(128)
  CV2_LJE(256)
#undef CV2_LJE
}

void launch_forces(DevState &s, const BoxDev &b, const LJDev *lj_dev, const LJDev &lj,
                   int precision, bool energy, hipStream_t st) {
  if (s.n == 0) return;
  const bool multi = lj.nt > 1;
  const bool halo = s.ng > 0;
#define CV2_LJ(M, E, H)                                      \
  if (multi == M && energy == E && halo == H) {              \
    lj_dispatch<M, E, H>(s, b, lj_dev, precision, st);       \
    return;                                                  \
  }
  CV2_LJ(false, false, false)
  CV2_LJ(false, true, false)
  CV2_LJ(true, false, false)
  CV2_LJ(true, true, false)
  CV2_LJ(false, false, true)
  CV2_LJ(false, true, true)
  CV2_LJ(true, false, true)
  CV2_LJ(true, true, true)
#undef CV2_LJ
}

}  // namespace cv2
