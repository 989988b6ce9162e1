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
             const int *__restrict__ tag, const int *__restrict__ mol,
             const LJDev *__restrict__ lj, BoxDev b, double *__restrict__ acc,
             double *__restrict__ partials, int *__restrict__ flags, int nblk_real) {
  __shared__ double sh[kBlock / 64];
  const int blk = xcd_block(nblk_real);
  if (blk < 0 || flags[F_HALT]) return;
  const int i = blk * kBlock + threadIdx.x;
  double vir = 0.0, pe = 0.0;
  if (i < n) {
    const double4 pi = pos[i];
    const int ti = MULTI ? (int)pi.w : 0;
    const int nt = MULTI ? lj->nt : 1;
    // single-type parameters stay in registers for the whole loop
    const double rc2_0 = lj->rc2[0], sig2_0 = lj->sig2[0], eps4_0 = lj->eps4[0], fsh_0 = lj->fshift[0];
    const double jt = MULTI ? lj->jump_tol : 0.0;  // jump_fix (species-table kernels only)
    const int cp = (nnbr[i] + 3) & ~3;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int id_i = (ENERGY && HALO) ? id[i] : 0;
    // ignore_intra_molecule: pairs of one molecule index are skipped,
    // atoms outside molecules (-1) included, as lj.cpp:446 compares them
    const int mol_i = (MULTI && mol) ? mol[tag[i]] : 0;
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
        double r2 = dx * dx + dy * dy + dz * dz;
        const int tp = MULTI ? ti * nt + (int)pj[u].w : 0;
        const double rc2 = MULTI ? lj->rc2[tp] : rc2_0;
        const bool same_mol = MULTI && mol && mol[tag[j]] == mol_i;
        if (r2 <= rc2 && r2 > 0.0 && !same_mol) {
          if (MULTI && r2 < jt) r2 = jt;  // lj.cpp:459-468
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
    if (isnan(ax) || isnan(ay) || isnan(az)) raise_fatal(flags, F_NAN);
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

// Constants of the FP32 pair arithmetic for one species pair.  With g = 1
// the separation is in length units (dx = dq * u_d); on a box with one
// fixed-point unit u on all axes (ISO) the separation is left in units of u
// (g = u), which saves three multiplications per pair, and the unit is
// folded into the constants:
//   forceCoef * 2^32 * g = r2i' s6 (e6 - e12 s6) + fsh     r2i' = 1 / r2', r2' = r2 / g^2
// with e6 = 6 (4 eps) 2^32 / g, e12 = 12 (4 eps) 2^32 / g, fsh = 4 eps fshift 2^32 g,
// s6 = (sigma^2 / r2)^3 = (sig2' r2i')^3, sig2' = sigma^2 / g^2.
struct PairF32 {
  double rc2, sig2, r2ovf;  // in units of g^2
  double jt;                // jump_fix clamp of r^2 (0: off), units of g^2
  float rc2if;              // 1 / rc2 rounded to FP32 (cutoff test of the branchless entry)
  float e6, e12, fsh;
  float eps4;
  double ecut, ecut2;       // energy: ecut in units of 1/g^2
  double evir;              // virial scale: g 2^-32 2^24 (fixed-point energy units)
};

// Species pair tp in length units (g = 1), from the precomputed FP32 tables.
__device__ inline PairF32 pair_f32_table(const LJDev *__restrict__ lj, int tp) {
  PairF32 c;
  c.rc2 = lj->rc2[tp];
  c.sig2 = lj->sig2[tp];
  c.r2ovf = lj->r2ovf[tp];
  c.jt = lj->jump_tol;
  c.rc2if = (float)(1.0 / c.rc2);
  c.e6 = lj->e6f[tp];
  c.e12 = lj->e12f[tp];
  c.fsh = lj->fshf[tp];
  c.eps4 = lj->eps4f[tp];
  c.ecut = lj->ecut[tp];
  c.ecut2 = lj->ecut2[tp];
  c.evir = kEnergyScale / kForceScale;
  return c;
}

__device__ inline PairF32 pair_f32(const LJDev *__restrict__ lj, int tp, double g) {
  PairF32 c;
  const double g2i = 1.0 / (g * g);
  c.rc2 = lj->rc2[tp] * g2i;
  c.sig2 = lj->sig2[tp] * g2i;
  c.r2ovf = lj->r2ovf[tp] * g2i;
  c.jt = lj->jump_tol * g2i;
  c.rc2if = (float)(1.0 / c.rc2);
  const double e4s = lj->eps4[tp] * kForceScale / g;
  c.e6 = (float)(6.0 * e4s);
  c.e12 = (float)(12.0 * e4s);
  c.fsh = (float)(lj->eps4[tp] * lj->fshift[tp] * kForceScale * g);
  c.eps4 = lj->eps4f[tp];
  c.ecut = lj->ecut[tp] * g * g;
  c.ecut2 = lj->ecut2[tp];
  c.evir = g * (kEnergyScale / kForceScale);
  return c;
}

// One list entry of the FP32 path: separation from the fixed-point
// coordinates (exact integer difference, one conversion), r^2 and the
// cutoff test in FP64, r^-2 and (sigma^2/r^2)^3 in FP64 each rounded once to
// FP32, the force polynomial in FP32, and the force component times 2^32
// rounded to an integer inside one FP64 FMA and summed as a 64-bit integer.
struct AccF32 {
  unsigned long long sx = 0, sy = 0, sz = 0;
  long long vir = 0, pe = 0;
  int terms = 0;
  bool ovf = false;
};

template <bool ISO, bool ENERGY, bool JUMP = false>
__device__ inline void pair_f32_entry(const int4 qi, const int4 qj, const double u0, const double u1,
                                      const double u2, const PairF32 &c, bool once, AccF32 &a) {
  double dx = (double)(qj.x - qi.x), dy = (double)(qj.y - qi.y), dz = (double)(qj.z - qi.z);
  if (!ISO) {
    dx *= u0;
    dy *= u1;
    dz *= u2;
  }
  double r2 = dx * dx + dy * dy + dz * dz;
  if (r2 <= c.rc2 && r2 > 0.0) {
    if (JUMP && r2 < c.jt) r2 = c.jt;  // jump_fix, lj.cpp:459-468
    double r2i = __builtin_amdgcn_rcp(r2);
    r2i = fma(r2i, fma(-r2, r2i, 1.0), r2i);
    const double s2 = c.sig2 * r2i;
    const float s6 = (float)(s2 * s2 * s2);
    const float r2if = (float)r2i;
    const float fc = fmaf(s6 * fmaf(-c.e12, s6, c.e6), r2if, c.fsh);
    // |F| stays below 2^19 for r^2 >= r2ovf (fixed-point range)
    a.ovf |= r2 < c.r2ovf;
    const double fcs = (double)fc;
    a.sx += (unsigned long long)__double_as_longlong(fma(fcs, dx, kMagic));
    a.sy += (unsigned long long)__double_as_longlong(fma(fcs, dy, kMagic));
    a.sz += (unsigned long long)__double_as_longlong(fma(fcs, dz, kMagic));
    ++a.terms;
    if (ENERGY && once) {
      // -forceCoef r^2 and U in 2^-24 units
      a.vir += __double_as_longlong(fma(-fcs * r2, c.evir, kMagic)) - kMagicBits;
      const float e = c.eps4 * ((s6 * s6 - s6) + (float)(c.ecut * r2) - (float)c.ecut2);
      a.pe += __double_as_longlong(fma((double)e, kEnergyScale, kMagic)) - kMagicBits;
    }
  }
}

__device__ inline double fx_value(unsigned long long s, int terms) {
  const unsigned long long base = (unsigned long long)terms * (unsigned long long)kMagicBits;
  return (double)(long long)(s - base);
}

template <bool MULTI, bool ENERGY, bool HALO, bool ISO>
__global__ void __launch_bounds__(kBlock)
    k_lj_f32fx(const int4 *__restrict__ q, const uint32_t *__restrict__ nbr,
               const int *__restrict__ nnbr, int stride, int n, const int *__restrict__ id,
               const int *__restrict__ tag, const int *__restrict__ mol,
               const LJDev *__restrict__ lj, BoxDev b, double *__restrict__ acc,
               long long *__restrict__ partials, int *__restrict__ flags, int nblk_real) {
  static_assert(!(MULTI && ISO), "species tables are in length units");
  __shared__ long long sh[kBlock / 64];
  const int blk = xcd_block(nblk_real);
  if (blk < 0 || flags[F_HALT]) return;
  const int i = blk * kBlock + threadIdx.x;
  const double u0 = b.qinv[0], u1 = b.qinv[1], u2 = b.qinv[2];
  AccF32 a;
  if (i < n) {
    const int4 qi = q[i];
    const int ti = MULTI ? (qi.w & 7) : 0;
    const int nt = MULTI ? lj->nt : 1;
    const PairF32 c0 = pair_f32(lj, 0, ISO ? u0 : 1.0);
    const int cp = (nnbr[i] + 3) & ~3;
    const int id_i = (ENERGY && HALO) ? id[i] : 0;
    const int mol_i = (MULTI && mol) ? mol[tag[i]] : 0;  // as in k_lj_f64
    for (int k = 0; k < cp; k += 4) {
      const u32x4 g = load_group(nbr, (size_t)stride, i, k);
      const uint32_t e[4] = {g.x, g.y, g.z, g.w};
      int4 qj[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) qj[v] = q[e[v] & kIndexMask];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = (int)(e[v] & kIndexMask);
        const bool once = ENERGY && (HALO && j >= n ? (id_i < id[j]) : (i < j));
        if (MULTI) {
          if (!(mol && mol[tag[j]] == mol_i))
            pair_f32_entry<false, ENERGY, true>(qi, qj[v], u0, u1, u2,
                                                pair_f32_table(lj, ti * nt + (qj[v].w & 7)), once, a);
        } else {
          pair_f32_entry<ISO, ENERGY>(qi, qj[v], u0, u1, u2, c0, once, a);
        }
      }
    }
    // fc * dx and fc' * dq are both the force component times 2^32
    const double mi = lj->mass_inv[ti] * (1.0 / kForceScale);
    acc[i] = fx_value(a.sx, a.terms) * mi;
    acc[n + i] = fx_value(a.sy, a.terms) * mi;
    acc[2 * n + i] = fx_value(a.sz, a.terms) * mi;
    if (a.ovf) raise_fatal(flags, F_FXOVF);
  }
  if (ENERGY) {
    const long long v = block_sum_i64(a.vir, sh);
    __syncthreads();
    const long long u = block_sum_i64(a.pe, sh);
    if (threadIdx.x == 0) {  // per-block partials, summed by k_reduce_i64
      partials[2 * blk] = v;
      partials[2 * blk + 1] = u;
    }
  }
}

// ---- FP32 forces on wave-build boxes: one block per force task ----
//
// The block stages the task's neighbourhood (kTaskPieces atom ranges, see
// TaskGeom) into LDS as int32 x, y, z arrays (+ type, + atom index for the
// energy pass), in piece order, plus a sentinel at slot S, and each thread
// takes one home atom of the task: its list holds 16-bit slots into the
// staging area, 8 per 16-byte group (nbr16[((k/8)*stride + i)*8 + k%8]),
// padded with the sentinel.  Neighbour reads are LDS reads instead of
// gathers through the vector memory pipeline, and the list is half the
// bytes of the 32-bit form.  A task whose neighbourhood exceeds the LDS
// staging capacity (dense regions) reads its neighbours from q instead.
//
// The sentinel is the first home atom moved by 2^31 units in y and z: home
// atoms share one cell row (|dy|, |dz| < w), and 2^31 units exceed 2 w on a
// wave-build box, so it is more than w >= list cutoff from all of them.
// The lists hold no self entries, so the pair arithmetic needs no r^2 > 0
// test and runs without branches: a pair beyond the cutoff adds a zero
// force, every entry (padding included) adds one magic bias.
constexpr int kSortBins = 64;  // list lengths in groups of 8 (longer lists share the last bin)
constexpr int kTailGrid = 64;  // work-groups of the tail launch (tasks beyond the main grid)

__device__ inline u32x4 load_group8(const uint32_t *__restrict__ nbr, size_t stride, int i, int k) {
  return __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(nbr) + (size_t)(k >> 3) * stride + i);
}

// Atom index of staging slot sl < S: the last piece whose slot base is <= sl
// (empty pieces share the base of the next one).
__device__ inline int slot_to_j(const int4 *pc, int sl) {
  int p = 0;
#pragma unroll
  for (int k = 1; k < kTaskPieces; ++k) p = pc[k].y <= sl ? k : p;
  return pc[p].x + (sl - pc[p].y);
}

// Branchless entry (single species): force component * 2^32 rounded inside
// one FP64 FMA.  fcmax tracks the largest |forceCoef'| of a pair within the
// cutoff: |fc' d| <= |fc'| rc on every axis, so fcmax * rc < 2^51 keeps each
// product in the range of the rounding FMA (checked once per atom).
template <bool ISO, bool SELF = false>
__device__ inline void pair_f32_flat(const int4 qi, const int4 qj, double u0, double u1, double u2,
                                     const PairF32 &c, unsigned long long &sx,
                                     unsigned long long &sy, unsigned long long &sz,
                                     float &fcmax) {
  double dx = (double)(qj.x - qi.x), dy = (double)(qj.y - qi.y), dz = (double)(qj.z - qi.z);
  if (!ISO) {
    dx *= u0;
    dy *= u1;
    dz *= u2;
  }
  const double r2 = dx * dx + dy * dy + dz * dz;
  double r2i = __builtin_amdgcn_rcp(r2);
  r2i = fma(r2i, fma(-r2, r2i, 1.0), r2i);
  const double s2 = c.sig2 * r2i;
  const float s6 = (float)(s2 * s2 * s2);
  const float r2if = (float)r2i;
  const float fc = fmaf(s6 * fmaf(-c.e12, s6, c.e6), r2if, c.fsh);
  // The cutoff test compares r^-2 in FP32 (one FP32 compare instead of an
  // FP64 one): it can differ from r2 <= rc2 only within an FP32 rounding of
  // the cutoff, where the shifted force is zero to first order
  // (forceCoef(rc) = 0, lj.cpp:475-484).  SELF: the list pads with the atom
  // itself (r^2 = 0), which adds nothing.
  // (an empty asm here had pinned the select before the conversion; it also
  // kept the scheduler from filling the compare's hazard slots: 21 s_nop per
  // 8 entries with it, 15 without, same results)
  const float fsel = (SELF ? (r2 <= c.rc2 && r2 > 0.0) : (r2if >= c.rc2if)) ? fc : 0.0f;
  fcmax = fmaxf(fcmax, fabsf(fsel));
  const double fcs = (double)fsel;
  sx += (unsigned long long)__double_as_longlong(fma(fcs, dx, kMagic));
  sy += (unsigned long long)__double_as_longlong(fma(fcs, dy, kMagic));
  sz += (unsigned long long)__double_as_longlong(fma(fcs, dz, kMagic));
}

// Coordinates (and type) of slot sl: LDS when staged, else q (the sentinel
// at slot S built from the first home atom).
template <bool STAGED, bool MULTI>
__device__ inline int4 task_coord(const int *sx, const int *sy, const int *sz,
                                  const unsigned char *stype, const int4 *__restrict__ q,
                                  const int4 *pc, int S, const int4 &sent, int sl) {
  if (STAGED) return make_int4(sx[sl], sy[sl], sz[sl], MULTI ? (int)stype[sl] : 0);
  const bool is_sent = sl == S;
  const int4 v = q[slot_to_j(pc, is_sent ? 0 : sl)];  // selects, not a branch
  return make_int4(is_sent ? sent.x : v.x, is_sent ? sent.y : v.y, is_sent ? sent.z : v.z,
                   is_sent ? sent.w : v.w);
}

// All entries of home atom i (8 per group: slots decoded and coordinates
// read for the whole group before the pair arithmetic).
template <bool MULTI, bool ENERGY, bool ISO, bool STAGED>
__device__ inline void task_atom(const int4 *__restrict__ q, const uint32_t *__restrict__ nbr,
                                 int stride, int i, int cp, const int4 qi, int ti, int nt,
                                 const LJDev *__restrict__ lj, const PairF32 &c0, double u0,
                                 double u1, double u2, const int *sx, const int *sy,
                                 const int *sz, const unsigned char *stype, const int *sj,
                                 const int4 *pc, int S, const int4 &sent, AccF32 &a,
                                 float &fcmax, const int *__restrict__ hid, int n, bool flat2,
                                 double &rc_flat) {
  if (cp == 0) return;
  // host-halo ghosts (hid set): a pair with a ghost counts once by id, as
  // lj.cpp:487 (the sentinel, j = 0x7fffffff, is never within the cutoff)
  const int id_i = (ENERGY && hid) ? hid[i] : 0;
  // flat2 (two species, no energy, no jump_fix): the branchless entry of the
  // single-species kernel, with the pair constants of the home atom's two
  // type pairs held in registers and selected per entry by the neighbour's
  // type instead of gathered from the table
  PairF32 cA{}, cB{};
  if (MULTI && flat2) {
    cA = pair_f32_table(lj, ti * 2);
    cB = pair_f32_table(lj, ti * 2 + 1);
    rc_flat = sqrt(fmax(cA.rc2, cB.rc2));
  }
  // the next group is loaded while this one is processed
  u32x4 gn = load_group8(nbr, (size_t)stride, i, 0);
  for (int k = 0; k < cp; k += 8) {
    const u32x4 g = gn;
    gn = load_group8(nbr, (size_t)stride, i, min(k + 8, cp - 8));
    __builtin_amdgcn_sched_barrier(0);  // keep the load ahead of this group's work
    const uint32_t w[4] = {g.x, g.y, g.z, g.w};
    int sl[8];
    int4 qj[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      sl[v] = (int)((w[v >> 1] >> (16 * (v & 1))) & 0xFFFFu);
      qj[v] = task_coord<STAGED, MULTI>(sx, sy, sz, stype, q, pc, S, sent, sl[v]);
    }
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      if (MULTI && !ENERGY && flat2) {  // uniform over the launch
        const bool b = (qj[v].w & 1) != 0;
        PairF32 c;
        c.sig2 = b ? cB.sig2 : cA.sig2;
        c.rc2if = b ? cB.rc2if : cA.rc2if;
        c.e6 = b ? cB.e6 : cA.e6;
        c.e12 = b ? cB.e12 : cA.e12;
        c.fsh = b ? cB.fsh : cA.fsh;
        pair_f32_flat<ISO>(qi, qj[v], u0, u1, u2, c, a.sx, a.sy, a.sz, fcmax);
      } else if (ENERGY || MULTI) {
        int j = 0;
        if (ENERGY) j = sl[v] == S ? 0x7fffffff : (STAGED ? sj[sl[v]] : slot_to_j(pc, sl[v]));
        const PairF32 c = MULTI ? pair_f32_table(lj, ti * nt + (qj[v].w & 7)) : c0;
        bool once = ENERGY && i < j;
        if (ENERGY && hid && j >= n && j != 0x7fffffff) once = id_i < hid[j];
        pair_f32_entry<ISO, ENERGY, MULTI>(qi, qj[v], u0, u1, u2, c, once, a);
      } else {
        pair_f32_flat<ISO>(qi, qj[v], u0, u1, u2, c0, a.sx, a.sy, a.sz, fcmax);
      }
    }
  }
}

// Integrator arrays of the fused form (FUSE): the epilogue applies finish +
// start of step p.step to each home atom instead of storing its acceleration.
struct FuseArgs {
  double4 *pos;
  const double4 *pos_ref;
  double *vel;
  const int *id;
  int4 *q_out;
  StepParams p;
  int langevin;
  int start;  // 0: the last step of a run, its finishing velocity half only
};

// One work-group per force task, in the compact order of the build
// (ctinfo / ctpiece, flags[F_NTASK] of them).  Work-groups are dealt
// round-robin to the 8 XCDs: XCD x takes the x-th eighth of the tasks
// (spatially contiguous rows, one L2 working set).  The main launch
// (TAIL = false, main_per work-groups per XCD) runs one task per work-group;
// a build with more tasks than that leaves the rest of each eighth to the
// tail launch (TAIL), whose few work-groups loop over them (and otherwise
// return at once).  A loop in the main launch would keep the pair constants
// live across the integrator epilogue: it measured 8% slower.
// FUSE: five wavefronts per SIMD (96 registers); the LDS stage allows five
// blocks per CU
template <bool MULTI, bool ENERGY, bool ISO, bool FUSE, bool TAIL>
__global__ void __launch_bounds__(kTaskBlock) __attribute__((amdgpu_waves_per_eu(FUSE ? 5 : 1)))
    k_lj_f32_task(const int4 *__restrict__ q, const uint32_t *__restrict__ nbr,
                  const int *__restrict__ nnbr, int stride, int n, const int4 *__restrict__ tinfo,
                  const int4 *__restrict__ tpiece,
                  int stage_limit, const LJDev *__restrict__ lj, BoxDev b,
                  double *__restrict__ acc, long long *__restrict__ partials,
                  int *__restrict__ flags, FuseArgs fz, int main_per) {
  static_assert(!(MULTI && ISO), "species tables are in length units");
  static_assert(!(FUSE && ENERGY), "the energy pass is not fused");
  __shared__ __attribute__((aligned(16))) int sx[kStage];
  __shared__ int sy[kStage], sz[kStage];
  __shared__ unsigned char stype[MULTI ? kStage : 1];
  __shared__ int sj[ENERGY ? kStage : 1];
  __shared__ int4 pc[kTaskPieces];
  __shared__ unsigned char pstart[kStage / 64 + 2];  // piece holding slot 64 c
  __shared__ long long sh[kTaskBlock / 64];
  __shared__ int lhist[kSortBins + 1];
  __shared__ unsigned char lorder[kTaskBlock];
  __shared__ int4 sent;  // the sentinel (read by the staging and the unstaged path only)
  if (flags[F_HALT]) return;  // uniform over the launch
  const int ntask = flags[F_NTASK];
  const int xcd = blockIdx.x & 7;  // grids: multiples of 8
  const int k0 = (int)((long long)ntask * xcd / 8), k1 = (int)((long long)ntask * (xcd + 1) / 8);
  long long vir = 0, pe = 0;
  // FUSE: accelerations of the (at most kTaskBlock) home atoms, in the
  // staging area once every neighbour has been read
  double *at_sh = reinterpret_cast<double *>(sx);
  static_assert(3 * kTaskBlock * sizeof(double) <= sizeof(int) * kStage, "staging area too small");
  auto fuse_step = [&](int ia, double4 x, double v[3], const double4 &r, int idv, const double *aa) {
    if (fz.start) {
      if (fz.langevin)
        integrate_values<true, true, true, true>(x, v, idv, aa, fz.p);
      else
        integrate_values<false, true, true, true>(x, v, idv, aa, fz.p);
      integrate_store<true, true, false>(ia, n, x, v, r, fz.pos, fz.vel, fz.q_out, fz.p, flags);
    } else {  // positions and q stay (other tasks may still stage them); the
              // accelerations are kept for the next run's first half step
      if (fz.langevin)
        integrate_values<true, true, false, true>(x, v, idv, aa, fz.p);
      else
        integrate_values<false, true, false, true>(x, v, idv, aa, fz.p);
      integrate_store<false, false, false>(ia, n, x, v, r, fz.pos, fz.vel, fz.q_out, fz.p, flags);
      acc[ia] = aa[0];
      acc[n + ia] = aa[1];
      acc[2 * n + ia] = aa[2];
    }
  };

  // two species without jump_fix: the branchless entry (task_atom)
  const bool flat2 = MULTI && !ENERGY && lj->nt == 2 && lj->jump_tol == 0.0;
  auto run_task = [&](int k, double u0, double u1, double u2, const PairF32 &c0, double fc_rc, int nt) {
    const int4 *tpk = tpiece + (size_t)k * kTaskPieces;
    if (threadIdx.x < kTaskPieces) pc[threadIdx.x] = tpk[threadIdx.x];
    // the piece of slot 64 c, counted from the piece bases (uniform loads),
    // so that staging starts each 64-slot chunk at its piece
    if (threadIdx.x <= kStage / 64 + 1) {
      const int sl0 = (int)threadIdx.x * 64;
      int p = 0;
#pragma unroll
      for (int j = 1; j < kTaskPieces; ++j) p += tpk[j].y <= sl0 ? 1 : 0;
      pstart[threadIdx.x] = (unsigned char)p;
    }
    const int4 ti4 = tinfo[k];
    const int h0 = ti4.x, h1 = ti4.y, S = ti4.z;
    const bool staged = S + 1 <= stage_limit;  // uniform over the block (sentinel included)
    const int nh = h1 - h0;
    const bool sorted = nh <= kTaskBlock;
    // the list length of this thread's home atom (sort key), loaded while
    // the neighbourhood is staged
    const int my_nn = (int)threadIdx.x < nh ? nnbr[h0 + threadIdx.x] : 0;
    if (threadIdx.x <= kSortBins) lhist[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
      int4 v = make_int4(0, 0, 0, 0);
      if (h0 < h1) {
        v = q[h0];
        v.y = (int)((unsigned)v.y + 0x80000000u);
        v.z = (int)((unsigned)v.z + 0x80000000u);
      }
      sent = v;
    }
    __syncthreads();
    if (staged) {
      // slots threadIdx.x + k * kTaskBlock; their pieces are non-decreasing in k
      constexpr int K = kStage / kTaskBlock;
      int jj[K];
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const int sl = kk * kTaskBlock + (int)threadIdx.x;
        // the chunk's first piece; most chunks lie in one piece (no search),
        // the rest cross few boundaries (uniform over the wavefront)
        int p = pstart[sl >> 6];
        if (p != (int)pstart[(sl >> 6) + 1])
          while (p + 1 < kTaskPieces && pc[p + 1].y <= sl) ++p;
        jj[kk] = pc[p].x + (sl - pc[p].y);
      }
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const int sl = kk * kTaskBlock + (int)threadIdx.x;
        if (sl <= S) {
          const int4 v = sl < S ? q[jj[kk]] : sent;
          sx[sl] = v.x;
          sy[sl] = v.y;
          sz[sl] = v.z;
          if (MULTI) stype[MULTI ? sl : 0] = (unsigned char)(v.w & 7);
          if (ENERGY) sj[ENERGY ? sl : 0] = sl < S ? jj[kk] : 0x7fffffff;
        }
      }
    }
    // A wavefront runs until its longest list is done: give each wavefront
    // home atoms of similar list length (counting sort on the group count,
    // longest first).  Which thread computes an atom does not change its
    // result.
    if (sorted) {
      int bin = 0, rank = 0;
      if ((int)threadIdx.x < nh) {
        bin = kSortBins - 1 - min((my_nn + 7) >> 3, kSortBins - 1);
        rank = atomicAdd(&lhist[bin + 1], 1);
      }
      __syncthreads();
      // bin offsets: an inclusive scan of the 64 counts by the first wavefront
      static_assert(kSortBins == 64, "one wavefront scans the bins");
      if (threadIdx.x < 64) {
        int c = lhist[threadIdx.x + 1];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const int v = __shfl_up(c, d, 64);
          c += (int)threadIdx.x >= d ? v : 0;
        }
        lhist[threadIdx.x + 1] = c;
      }
      __syncthreads();
      if ((int)threadIdx.x < nh) lorder[lhist[bin] + rank] = (unsigned char)threadIdx.x;
    }
    __syncthreads();
    double my_a[3] = {0.0, 0.0, 0.0};
    int my_t = -1;
    for (int t = (int)threadIdx.x; t < nh; t += kTaskBlock) {
      const int i = h0 + (sorted ? (int)lorder[t] : t);
      const int4 qi = q[i];
      const int ti = MULTI ? (qi.w & 7) : 0;
      const int cp = (nnbr[i] + 7) & ~7;
      AccF32 a;
      float fcmax = 0.0f;
      double rc_flat = fc_rc;
      if (staged)
        task_atom<MULTI, ENERGY, ISO, true>(q, nbr, stride, i, cp, qi, ti, nt, lj, c0, u0, u1, u2, sx,
                                            sy, sz, stype, sj, pc, S, sent, a, fcmax,
                                            FUSE ? nullptr : fz.id, n, flat2, rc_flat);
      else
        task_atom<MULTI, ENERGY, ISO, false>(q, nbr, stride, i, cp, qi, ti, nt, lj, c0, u0, u1, u2,
                                             sx, sy, sz, stype, sj, pc, S, sent, a, fcmax,
                                             FUSE ? nullptr : fz.id, n, flat2, rc_flat);
      // the flat form biases every entry, the branched form counts its terms
      const bool flat = !ENERGY && (!MULTI || flat2);
      const int terms = flat ? cp : a.terms;
      const bool ovf = flat ? !((double)fcmax * rc_flat < 0.9 * 2251799813685248.0) : a.ovf;
      const double mi = lj->mass_inv[ti] * (1.0 / kForceScale);
      const double ai[3] = {fx_value(a.sx, terms) * mi, fx_value(a.sy, terms) * mi,
                            fx_value(a.sz, terms) * mi};
      if (FUSE && sorted) {
        // one atom per thread: handed back in atom order through LDS (below)
        my_a[0] = ai[0];
        my_a[1] = ai[1];
        my_a[2] = ai[2];
        my_t = i - h0;
      } else if (FUSE) {
        double v[3] = {fz.vel[i], fz.vel[n + i], fz.vel[2 * n + i]};
        fuse_step(i, fz.pos[i], v, fz.pos_ref[i], fz.langevin ? fz.id[i] : 0, ai);
      } else {
        acc[i] = ai[0];
        acc[n + i] = ai[1];
        acc[2 * n + i] = ai[2];
      }
      if (ovf) raise_fatal(flags, F_FXOVF);
      vir += a.vir;
      pe += a.pe;
    }
    if (FUSE && sorted) {
      // the integrator in atom order: coalesced position / velocity traffic
      __syncthreads();  // every neighbour read of the staging area is done
      if (my_t >= 0) {
        at_sh[3 * my_t] = my_a[0];
        at_sh[3 * my_t + 1] = my_a[1];
        at_sh[3 * my_t + 2] = my_a[2];
      }
      __syncthreads();
      // (the integrator's loads are issued here rather than before the force
      // loop: held through it they cost a wavefront per SIMD of occupancy)
      const int t = (int)threadIdx.x;
      if (t < nh) {
        const int ia = h0 + t;
        double v[3] = {fz.vel[ia], fz.vel[n + ia], fz.vel[2 * n + ia]};
        const double ai[3] = {at_sh[3 * t], at_sh[3 * t + 1], at_sh[3 * t + 2]};
        fuse_step(ia, fz.pos[ia], v, fz.pos_ref[ia], fz.langevin ? fz.id[ia] : 0, ai);
      }
    }
  };

  if (!TAIL) {
    const double u0 = b.qinv[0], u1 = b.qinv[1], u2 = b.qinv[2];
    const PairF32 c0 = pair_f32(lj, 0, ISO ? u0 : 1.0);
    const double fc_rc = sqrt(c0.rc2);  // cutoff in the units of the separations
    const int k = k0 + (int)(blockIdx.x >> 3);
    if (k < min(k1, k0 + main_per)) run_task(k, u0, u1, u2, c0, fc_rc, MULTI ? lj->nt : 1);
  } else {
    for (int k = k0 + main_per + (int)(blockIdx.x >> 3); k < k1; k += (int)(gridDim.x >> 3)) {
      // the constants are formed per task (not hoisted out of the loop)
      __asm__ volatile("" ::: "memory");
      const double u0 = b.qinv[0], u1 = b.qinv[1], u2 = b.qinv[2];
      const PairF32 c0 = pair_f32(lj, 0, ISO ? u0 : 1.0);
      run_task(k, u0, u1, u2, c0, sqrt(c0.rc2), MULTI ? lj->nt : 1);
      __syncthreads();  // the LDS of this task is reused by the next
    }
  }
  if (ENERGY) {
    const long long v = block_sum_i64(vir, sh);
    __syncthreads();
    const long long u = block_sum_i64(pe, sh);
    if (threadIdx.x == 0) {  // per-block partials (tail: after the main launch's), summed by k_reduce_i64
      const int blk = TAIL ? 8 * main_per + (int)blockIdx.x : (int)blockIdx.x;
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
// fixed-point unit of its own system (ISO: every system has one unit on all
// axes); ENERGY writes per-atom fixed-point virial and energy (each pair
// once: i < j) for the per-system sums.
// FUSE: the epilogue runs the integrator pass of the next step
// (ens_step_atom) on the accelerations just summed, writing the next
// fixed-point coordinates to q_out (other threads still read q).
template <int BLOCK, bool ENERGY, bool ISO, bool FUSE>
__global__ void __launch_bounds__(BLOCK)
    k_lj_ens(const int4 *__restrict__ q, const uint32_t *__restrict__ nbr,
             const int *__restrict__ nnbr, int stride, int n, EnsDev e,
             const LJDev *__restrict__ lj, double *__restrict__ acc, long long *__restrict__ eatom,
             int *__restrict__ flags, double4 *__restrict__ pos, const double4 *__restrict__ pos_ref,
             double *__restrict__ vel, int4 *__restrict__ q_out, double dt, double thresh_sq) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const BoxDev &b = e.box[e.sys[i]];
  const double u0 = b.qinv[0], u1 = b.qinv[1], u2 = b.qinv[2];
  const int4 qi = q[i];
  const PairF32 c = pair_f32(lj, 0, ISO ? u0 : 1.0);
  const int cp = (nnbr[i] + 3) & ~3;
  AccF32 a;
  if (ENERGY) {
    for (int k = 0; k < cp; k += 4) {
      const u32x4 g = load_group(nbr, (size_t)stride, i, k);
      const uint32_t ent[4] = {g.x, g.y, g.z, g.w};
      int4 qj[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) qj[v] = q[ent[v] & kIndexMask];
#pragma unroll
      for (int v = 0; v < 4; ++v)
        pair_f32_entry<ISO, ENERGY>(qi, qj[v], u0, u1, u2, c, i < (int)(ent[v] & kIndexMask), a);
    }
  } else if (cp > 0) {
    // branchless entries (every entry adds one rounding bias), the next
    // group's entries loaded while this one runs
    float fcmax = 0.0f;
    u32x4 gn = load_group(nbr, (size_t)stride, i, 0);
    for (int k = 0; k < cp; k += 4) {
      const u32x4 g = gn;
      gn = load_group(nbr, (size_t)stride, i, min(k + 4, cp - 4));
      __builtin_amdgcn_sched_barrier(0);
      const uint32_t ent[4] = {g.x, g.y, g.z, g.w};
      int4 qj[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) qj[v] = q[ent[v] & kIndexMask];
#pragma unroll
      for (int v = 0; v < 4; ++v) pair_f32_flat<ISO, true>(qi, qj[v], u0, u1, u2, c, a.sx, a.sy, a.sz, fcmax);
    }
    a.terms = cp;
    a.ovf = !((double)fcmax * sqrt(c.rc2) < 0.9 * 2251799813685248.0);
  }
  const double mi = lj->mass_inv[0] * (1.0 / kForceScale);
  const double ai[3] = {fx_value(a.sx, a.terms) * mi, fx_value(a.sy, a.terms) * mi,
                        fx_value(a.sz, a.terms) * mi};
  if (FUSE) {
    ens_step_atom<true, true>(i, n, e, pos, pos_ref, vel, ai, q_out, dt, thresh_sq, flags);
  } else {
    acc[i] = ai[0];
    acc[n + i] = ai[1];
    acc[2 * n + i] = ai[2];
  }
  if (a.ovf) atomicOr(&flags[F_FXOVF], 1);
  if (ENERGY) {
    eatom[2 * (size_t)i] = a.vir;
    eatom[2 * (size_t)i + 1] = a.pe;
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
                                                           s.id, s.tag, s.mol, lj_dev, b, s.acc,
                                                           s.partials, s.flags, real);
    if (ENERGY) k_reduce_partials<<<1, 1024, 0, st>>>(s.partials, real, 2, s.reduce_out);
  } else {
    long long *part = reinterpret_cast<long long *>(s.partials);
    if (s.slots) {
      const int tgrid = s.task_grid;
      const bool tail = tgrid < (s.tg.nslot + 7) / 8 * 8;  // a capped grid: the rest in a tail launch
      // unfused launches carry the ids of the host-halo ghosts' pair rule
      FuseArgs fa{};
      fa.id = (ENERGY && HALO) ? s.id : nullptr;
#define CV2_LJT(I)                                                                              \
  do {                                                                                           \
  k_lj_f32_task<MULTI, ENERGY, I, false, false><<<tgrid, kTaskBlock, 0, st>>>(                   \
      s.q, s.nbr, s.nnbr, s.stride, s.n, s.ctinfo, s.ctpiece, s.stage_limit, lj_dev, b, s.acc,    \
      part, s.flags, fa, tgrid / 8);                                                             \
  if (tail)                                                                                      \
    k_lj_f32_task<MULTI, ENERGY, I, false, true><<<kTailGrid, kTaskBlock, 0, st>>>(              \
        s.q, s.nbr, s.nnbr, s.stride, s.n, s.ctinfo, s.ctpiece, s.stage_limit, lj_dev, b, s.acc,  \
        part, s.flags, fa, tgrid / 8);                                                           \
  } while (0)
      if (!MULTI && b.qiso)
        CV2_LJT(!MULTI);
      else
        CV2_LJT(false);
#undef CV2_LJT
      if (ENERGY) k_reduce_i64<<<1, 1024, 0, st>>>(part, tgrid + (tail ? kTailGrid : 0), s.fx_acc);
      return;
    }
    if (!MULTI && b.qiso)
      k_lj_f32fx<false, ENERGY, HALO, true><<<grid, kBlock, 0, st>>>(
          s.q, s.nbr, s.nnbr, s.stride, s.n, s.id, s.tag, s.mol, lj_dev, b, s.acc, part, s.flags,
          real);
    else
      k_lj_f32fx<MULTI, ENERGY, HALO, false><<<grid, kBlock, 0, st>>>(
          s.q, s.nbr, s.nnbr, s.stride, s.n, s.id, s.tag, s.mol, lj_dev, b, s.acc, part, s.flags,
          real);
    if (ENERGY) k_reduce_i64<<<1, 1024, 0, st>>>(part, real, s.fx_acc);
  }
}

}  // namespace

template <int B, bool ISO>
void lj_ens_dispatch(DevState &s, const EnsDev &e, const LJDev *lj_dev, bool energy,
                     long long *eatom, double fuse_dt, double thresh_sq, hipStream_t st) {
  const int g = (s.n + B - 1) / B;
  if (energy)
    k_lj_ens<B, true, ISO, false><<<g, B, 0, st>>>(s.q, s.nbr, s.nnbr, s.stride, s.n, e, lj_dev,
                                                   s.acc, eatom, s.flags, s.pos, s.pos_ref, s.vel,
                                                   s.q_alt, 0.0, 0.0);
  else if (fuse_dt > 0.0)
    k_lj_ens<B, false, ISO, true><<<g, B, 0, st>>>(s.q, s.nbr, s.nnbr, s.stride, s.n, e, lj_dev,
                                                   s.acc, eatom, s.flags, s.pos, s.pos_ref, s.vel,
                                                   s.q_alt, fuse_dt, thresh_sq);
  else
    k_lj_ens<B, false, ISO, false><<<g, B, 0, st>>>(s.q, s.nbr, s.nnbr, s.stride, s.n, e, lj_dev,
                                                    s.acc, eatom, s.flags, s.pos, s.pos_ref, s.vel,
                                                    s.q_alt, 0.0, 0.0);
}

void launch_forces_ens(DevState &s, const EnsDev &e, const LJDev *lj_dev, int block, bool iso,
                       bool energy, long long *eatom, hipStream_t st, double fuse_dt,
                       double thresh_sq) {
  if (s.n == 0) return;
#define CV2_LJE(B)                                                    \
  if (block == B) {                                                   \
    if (iso)                                                          \
      lj_ens_dispatch<B, true>(s, e, lj_dev, energy, eatom, fuse_dt, thresh_sq, st);      \
    else                                                              \
      lj_ens_dispatch<B, false>(s, e, lj_dev, energy, eatom, fuse_dt, thresh_sq, st);     \
    return;                                                           \
  }
  CV2_LJE(64)
  CV2_LJE(128)
  CV2_LJE(256)
#undef CV2_LJE
}

bool launch_forces_fused(DevState &s, const BoxDev &b, const LJDev *lj_dev, const LJDev &lj,
                         const UpdateArgs &u, hipStream_t st) {
  // (the fused epilogue has no owned wrap and no reference rebuild triggers)
  if (!s.slots || s.n == 0 || s.ng > 0 || !s.q_alt || !u.finish) return false;
  if (u.ref_rebuild || u.wrap_owned) return false;
  FuseArgs fz{s.pos, s.pos_ref, s.vel, s.id, s.q_alt, step_params(b, u), u.langevin ? 1 : 0,
              u.start ? 1 : 0};
  long long *part = reinterpret_cast<long long *>(s.partials);
  const int tgrid = s.task_grid;
  const bool tail = tgrid < (s.tg.nslot + 7) / 8 * 8;
#define CV2_LJF(M, I)                                                                           \
  do {                                                                                           \
  k_lj_f32_task<M, false, I, true, false><<<tgrid, kTaskBlock, 0, st>>>(                          \
      s.q, s.nbr, s.nnbr, s.stride, s.n, s.ctinfo, s.ctpiece, s.stage_limit, lj_dev, b, s.acc,    \
      part, s.flags, fz, tgrid / 8);                                                             \
  if (tail)                                                                                      \
  k_lj_f32_task<M, false, I, true, true><<<kTailGrid, kTaskBlock, 0, st>>>(                       \
      s.q, s.nbr, s.nnbr, s.stride, s.n, s.ctinfo, s.ctpiece, s.stage_limit, lj_dev, b, s.acc,    \
      part, s.flags, fz, tgrid / 8);                                                             \
  } while (0)
  if (lj.nt > 1 || lj.jump_tol > 0.0)
    CV2_LJF(true, false);
  else if (b.qiso)
    CV2_LJF(false, true);
  else
    CV2_LJF(false, false);
#undef CV2_LJF
  return true;
}

void launch_forces(DevState &s, const BoxDev &b, const LJDev *lj_dev, const LJDev &lj,
                   int precision, bool energy, hipStream_t st) {
  if (s.n == 0) return;
  // jump_fix and ignore_intra_molecule run on the species-table kernels
  const bool multi = lj.nt > 1 || lj.jump_tol > 0.0 || s.mol != nullptr;
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
