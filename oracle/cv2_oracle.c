/*
 * cv2_oracle.c -- CPU restatement of the CAVIAR2 MD step (test infrastructure).
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the HIP path is checked against, and
 * the CPU baseline that bench.py times.  The product never links it.
 *
 * Build with -ffp-contract=off and without -march flags: the reference is
 * built the same way on x86-64 (SSE2, no FMA), and every expression below
 * keeps the reference's association order so that results are bit-identical.
 * Each function cites the reference file:line it restates.
 *
 * Parity pins (tests/test_oracle.py): bit-exact against the compiled
 * reference on committed fixtures (wrap, ghosts, both neighbour builds, LJ
 * accelerations and virial, VV and Langevin trajectories, kinetic energy);
 * closed forms for the dimer force, FCC shell counts and the potential
 * energy (finite differences of orc_lj_pair_energy against the force).
 */
#include "cv2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* wrap: Domain::fix_position, domain.cpp:443-493                      */
/* ------------------------------------------------------------------ */
int orc_wrap(int64_t n, double *pos, const orc_box *b, int32_t *msd) {
  int any = 0;
  double size[3];
  for (int d = 0; d < 3; ++d) size[d] = b->hi[d] - b->lo[d]; /* domain.cpp:498 */
  for (int64_t i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) {
      if (b->bc[d] != 1) continue;
      double p = pos[3 * i + d];
      if (p < b->lo[d]) {
        p += size[d];
        any = 1;
        if (msd) msd[3 * i + d] -= 1;
      } else if (p > b->hi[d]) {
        p -= size[d];
        any = 1;
        if (msd) msd[3 * i + d] += 1;
      }
      pos[3 * i + d] = p;
    }
  }
  return any;
}

/* ------------------------------------------------------------------ */
/* ghosts: exchange_ghost_single_md_domain.cpp:24-216                  */
/* ------------------------------------------------------------------ */
static void emit_ghost(int64_t *k, int64_t cap, double *gpos, int64_t *owner, int32_t *code,
                       int64_t i, double x, double y, double z, int vx, int vy, int vz) {
  if (*k < cap) {
    gpos[3 * *k + 0] = x;
    gpos[3 * *k + 1] = y;
    gpos[3 * *k + 2] = z;
    if (owner) owner[*k] = i;
    if (code) code[*k] = (vx + 1) + 3 * (vy + 1) + 9 * (vz + 1);
  }
  ++*k;
}

/* exchange_ghost_mpi_shared_atoms.cpp:113-203 (classification), :222-241
 * (layout), :258-325 (fill). */
int64_t orc_halo_pack(int64_t n, const double *pos, const double *vel, const int64_t *id,
                      const int32_t *type, const int32_t gi[3], const int32_t np[3],
                      const double lo[3], const double hi[3], const double size_global[3],
                      const int32_t bc[3], double g, int32_t with_vel, int64_t counts[27],
                      int64_t cap, double *out) {
  const int64_t total = with_vel ? 8 : 5;
  int32_t *dval = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(n > 0 ? n : 1));
  for (int d = 0; d < 27; ++d) counts[d] = 0;
  for (int64_t i = 0; i < n; ++i) {
    int v[3];
    for (int a = 0; a < 3; ++a) {
      const double x = pos[3 * i + a];
      v[a] = 0;
      if (x < lo[a] + g) {
        v[a] = -1;
        if (gi[a] == 0) v[a] *= bc[a];
      } else if (x > hi[a] - g) {
        v[a] = +1;
        if (gi[a] == np[a] - 1) v[a] *= bc[a];
      }
      dval[3 * i + a] = v[a];
    }
    /* x, y, z, xy, xz, yz, xyz: each combination of the nonzero components */
    for (int m = 1; m < 8; ++m) {
      int ok = 1, c[3];
      for (int a = 0; a < 3; ++a) {
        const int use = (m >> a) & 1;
        if (use && v[a] == 0) ok = 0;
        c[a] = use ? v[a] + 1 : 1;
      }
      if (ok) counts[c[0] * 9 + c[1] * 3 + c[2]]++;
    }
  }
  int64_t need = 0, base[27];
  for (int d = 0; d < 27; ++d) {
    base[d] = need;
    need += total * counts[d];
  }
  if (need <= cap) {
    int64_t fill[27] = {0};
    for (int64_t i = 0; i < n; ++i) {
      const int *v = dval + 3 * i;
      for (int m = 1; m < 8; ++m) {
        int ok = 1, c[3];
        for (int a = 0; a < 3; ++a) {
          const int use = (m >> a) & 1;
          if (use && v[a] == 0) ok = 0;
          c[a] = use ? v[a] + 1 : 1;
        }
        if (!ok) continue;
        const int d = c[0] * 9 + c[1] * 3 + c[2];
        const int64_t N = counts[d], k = fill[d]++;
        double *pk = out + base[d];
        double p[3];
        for (int a = 0; a < 3; ++a) {
          p[a] = pos[3 * i + a];
          if (bc[a] == 1) {
            if (gi[a] == 0 && c[a] == 0) p[a] += size_global[a];
            if (gi[a] == np[a] - 1 && c[a] == 2) p[a] -= size_global[a];
          }
        }
        pk[k] = (double)id[i];
        pk[N + k] = (double)type[i];
        for (int a = 0; a < 3; ++a) pk[2 * N + 3 * k + a] = p[a];
        if (with_vel)
          for (int a = 0; a < 3; ++a) pk[5 * N + 3 * k + a] = vel[3 * i + a];
      }
    }
  }
  free(dval);
  return need;
}

int64_t orc_owned_pack(int64_t n, const double *pos, const double *vel, const double *acc,
                       const int64_t *id, const int32_t gi[3], const int32_t np[3],
                       const double lo[3], const double hi[3], const double size_global[3],
                       const int32_t bc[3], int64_t counts[27], int64_t cap, double *out,
                       int32_t *stay, double *pos_after) {
  /* :131-206: one direction per atom outside the domain; me == all[i][j][k]
   * exactly when every axis the direction moves along has one process */
  int32_t *dir = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int self[27];
  for (int d = 0; d < 27; ++d) {
    const int c[3] = {d / 9, (d / 3) % 3, d % 3};
    self[d] = 1;
    for (int a = 0; a < 3; ++a)
      if (c[a] != 1 && np[a] != 1) self[d] = 0;
    counts[d] = 0;
  }
  for (int64_t i = 0; i < n; ++i) {
    int v[3];
    for (int a = 0; a < 3; ++a) {
      const double x = pos[3 * i + a];
      v[a] = 0;
      if (x < lo[a]) {
        v[a] = -1;
        if (gi[a] == 0) v[a] *= bc[a];
      } else if (x > hi[a]) {
        v[a] = +1;
        if (gi[a] == np[a] - 1) v[a] *= bc[a];
      }
    }
    const int d = (v[0] + 1) * 9 + (v[1] + 1) * 3 + (v[2] + 1);
    dir[i] = d;
    if (d != 13 && !self[d]) counts[d]++; /* :199-204 */
  }
  /* :308-333 layout: id, position, velocity, acceleration, molecule index,
   * bond count; packets follow each other in direction order */
  const int64_t total = 12;
  int64_t need = 0, base[27];
  for (int d = 0; d < 27; ++d) {
    base[d] = need;
    need += total * counts[d];
  }
  if (need <= cap) {
    int64_t fill[27] = {0};
    for (int64_t i = 0; i < n; ++i) {
      const int d = dir[i];
      const double *p = pos + 3 * i;
      if (stay) stay[i] = d == 13 || self[d];
      if (d == 13 || self[d]) { /* :571-583: back into this domain */
        if (pos_after) {
          const int c[3] = {d / 9 - 1, (d / 3) % 3 - 1, d % 3 - 1};
          for (int a = 0; a < 3; ++a) pos_after[3 * i + a] = p[a] - c[a] * (hi[a] - lo[a]);
        }
        continue;
      }
      const int c[3] = {d / 9, (d / 3) % 3, d % 3};
      const int64_t N = counts[d], k = fill[d]++;
      double *pk = out + base[d];
      pk[k] = (double)id[i];
      for (int a = 0; a < 3; ++a) {
        double x = p[a];
        if (bc[a] == 1) { /* :402-431 */
          if (gi[a] == 0 && c[a] == 0) x += size_global[a];
          if (gi[a] == np[a] - 1 && c[a] == 2) x -= size_global[a];
        }
        pk[N + 3 * k + a] = x;
        pk[4 * N + 3 * k + a] = vel[3 * i + a];
        pk[7 * N + 3 * k + a] = acc[3 * i + a];
      }
      pk[10 * N + k] = -1.0; /* molecule index (atoms outside molecules) */
      pk[11 * N + k] = 0.0;  /* atomic bond count */
      if (pos_after)
        for (int a = 0; a < 3; ++a) pos_after[3 * i + a] = p[a];
    }
  }
  free(dir);
  return need;
}

int64_t orc_ghosts(int64_t n, const double *pos, const orc_box *b, double g, int64_t cap,
                   double *gpos, int64_t *owner, int32_t *code) {
  const double llow[3] = {b->lo[0] + g, b->lo[1] + g, b->lo[2] + g};
  const double lupp[3] = {b->hi[0] - g, b->hi[1] - g, b->hi[2] - g};
  const double w[3] = {b->hi[0] - b->lo[0], b->hi[1] - b->lo[1], b->hi[2] - b->lo[2]};
  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    int v[3];
    const double p[3] = {px, py, pz};
    for (int d = 0; d < 3; ++d) {
      if (p[d] < llow[d])
        v[d] = -1;
      else if (p[d] > lupp[d])
        v[d] = +1;
      else
        v[d] = 0;
      v[d] *= b->bc[d];
    }
    /* image = pos - val*width, emitted in the order x, y, z, xy, xz, yz, xyz */
    const double sx = px - (double)v[0] * w[0];
    const double sy = py - (double)v[1] * w[1];
    const double sz = pz - (double)v[2] * w[2];
    if (v[0] != 0) emit_ghost(&k, cap, gpos, owner, code, i, sx, py, pz, v[0], 0, 0);
    if (v[1] != 0) emit_ghost(&k, cap, gpos, owner, code, i, px, sy, pz, 0, v[1], 0);
    if (v[2] != 0) emit_ghost(&k, cap, gpos, owner, code, i, px, py, sz, 0, 0, v[2]);
    if (v[0] != 0 && v[1] != 0) emit_ghost(&k, cap, gpos, owner, code, i, sx, sy, pz, v[0], v[1], 0);
    if (v[0] != 0 && v[2] != 0) emit_ghost(&k, cap, gpos, owner, code, i, sx, py, sz, v[0], 0, v[2]);
    if (v[1] != 0 && v[2] != 0) emit_ghost(&k, cap, gpos, owner, code, i, px, sy, sz, 0, v[1], v[2]);
    if (v[0] != 0 && v[1] != 0 && v[2] != 0)
      emit_ghost(&k, cap, gpos, owner, code, i, sx, sy, sz, v[0], v[1], v[2]);
  }
  return k;
}

/* dot product order of vector3d.hpp:65-69: (x*x + y*y) + z*z */
static inline double dot3(double x, double y, double z) { return x * x + y * y + z * z; }

/* ------------------------------------------------------------------ */
/* O(N^2) Verlet list: neighborlist.cpp:247-297, cutoff_extra :317      */
/* ------------------------------------------------------------------ */
int64_t orc_nlist_brute(int64_t n, const double *pos, int64_t ng, const double *gpos,
                        double cutoff, double coef, int64_t cap, int64_t *offs,
                        uint32_t *nbr) {
  const double ce = cutoff * coef;
  const double ce_sq = ce * ce;
  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i) {
    offs[i] = k;
    const double *pi = pos + 3 * i;
    for (int64_t j = i + 1; j < n; ++j) {
      const double *pj = pos + 3 * j;
      if (dot3(pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]) < ce_sq) {
        if (k < cap) nbr[k] = (uint32_t)j;
        ++k;
      }
    }
    for (int64_t j = 0; j < ng; ++j) {
      const double *pj = gpos + 3 * j;
      if (dot3(pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]) < ce_sq) {
        if (k < cap) nbr[k] = (uint32_t)(j + n);
        ++k;
      }
    }
  }
  offs[n] = k;
  return k;
}

/* ------------------------------------------------------------------ */
/* cell-list Verlet list: neighborlist.cpp:303-455, 496-672             */
/* ------------------------------------------------------------------ */
typedef struct {
  int64_t *start, *count, *items; /* bins as CSR over (x + nx*y + nx*ny*z) */
} orc_bins;

int64_t orc_nlist_cell(int64_t n, const double *pos, int64_t ng, const double *gpos,
                       const orc_box *b, double cutoff, double coef, int64_t cap,
                       int64_t *offs, uint32_t *nbr) {
  const double ce = cutoff * coef; /* :334 */
  const double ce_sq = ce * ce;
  const double cinv = 1.0 / ce; /* :336 */
  /* lower_local/upper_local with a 1x1x1 process grid, domain.cpp:322-326 */
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = b->lo[d] + (b->hi[d] - b->lo[d]) * 0 / 1;
    hi[d] = b->lo[d] + (b->hi[d] - b->lo[d]) * 1 / 1;
  }
  int nb[3];
  for (int d = 0; d < 3; ++d) nb[d] = (int)ceil((hi[d] - lo[d]) * cinv); /* :340-342 */
  const int64_t ncell = (int64_t)nb[0] * nb[1] * nb[2];
  const int64_t ntot = n + ng;
  int64_t *cell = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ntot > 0 ? ntot : 1));
  int64_t *cnt = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
  int64_t *start = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
  int64_t *items = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ntot > 0 ? ntot : 1));
  /* binlist_index :392-414 (floor, clamp ghosts into edge bins) */
  for (int64_t a = 0; a < ntot; ++a) {
    const double *p = a < n ? pos + 3 * a : gpos + 3 * (a - n);
    int ind[3];
    for (int d = 0; d < 3; ++d) {
      int v = (int)floor((p[d] - lo[d]) * cinv);
      if (v < 0) v = 0;
      if (v >= nb[d]) v = nb[d] - 1;
      ind[d] = v;
    }
    cell[a] = ind[0] + (int64_t)nb[0] * ind[1] + (int64_t)nb[0] * nb[1] * ind[2];
    cnt[cell[a]]++;
  }
  for (int64_t c = 0; c < ncell; ++c) start[c + 1] = start[c] + cnt[c];
  memset(cnt, 0, sizeof(int64_t) * (size_t)ncell);
  /* update_cell_list :422-455: owned in index order, then ghosts as i+n */
  for (int64_t a = 0; a < ntot; ++a) items[start[cell[a]] + cnt[cell[a]]++] = a;

  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i) {
    offs[i] = k;
    const double *pi = pos + 3 * i;
    int ci[3];
    {
      int64_t c = cell[i];
      ci[0] = (int)(c % nb[0]);
      ci[1] = (int)((c / nb[0]) % nb[1]);
      ci[2] = (int)(c / ((int64_t)nb[0] * nb[1]));
    }
    /* make_neigh_bin :560-672: stencil x outer, z inner; clamp for open bc,
     * wrap for periodic bc; remove later duplicates (first occurrence kept). */
    int mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
      mn[d] = ci[d] - 1;
      mx[d] = ci[d] + 1;
      if (b->bc[d] == 0) {
        if (mn[d] == -1) mn[d] = 0;
        if (mx[d] == nb[d]) mx[d] = nb[d] - 1;
      }
    }
    int64_t sten[27];
    int ns = 0;
    for (int x = mn[0]; x <= mx[0]; ++x)
      for (int y = mn[1]; y <= mx[1]; ++y)
        for (int z = mn[2]; z <= mx[2]; ++z) {
          int v[3] = {x, y, z};
          for (int d = 0; d < 3; ++d) {
            if (v[d] == nb[d]) v[d] = 0;
            if (v[d] == -1) v[d] = nb[d] - 1;
          }
          int64_t c = v[0] + (int64_t)nb[0] * v[1] + (int64_t)nb[0] * nb[1] * v[2];
          int dup = 0;
          for (int q = 0; q < ns; ++q)
            if (sten[q] == c) dup = 1;
          if (!dup) sten[ns++] = c;
        }
    /* update_verlet_list_from_cell_list :512-546 */
    for (int q = 0; q < ns; ++q) {
      const int64_t c = sten[q];
      for (int64_t t = start[c]; t < start[c + 1]; ++t) {
        const int64_t j = items[t];
        const double *pj;
        if (j >= n) {
          pj = gpos + 3 * (j - n);
        } else {
          if (j <= i) continue;
          pj = pos + 3 * j;
        }
        if (dot3(pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]) < ce_sq) {
          if (k < cap) nbr[k] = (uint32_t)j;
          ++k;
        }
      }
    }
  }
  offs[n] = k;
  free(cell);
  free(cnt);
  free(start);
  free(items);
  return k;
}

/* ------------------------------------------------------------------ */
/* LJ: lj.cpp:393-517                                                  */
/* ------------------------------------------------------------------ */
double orc_lj_pair_energy(double r2, double eps, double sigma, double rc2) {
  /* U(r) = 4e[(s/r)^12-(s/r)^6] + 4e[6(s/rc)^12-3(s/rc)^6](r/rc)^2
   *        - 4e[7(s/rc)^12-4(s/rc)^6],  the unique U with U'(r)/r equal to
   * the reference forceCoef and U(rc) = 0. */
  const double s2 = sigma * sigma;
  const double rho2 = s2 / r2, rho6 = rho2 * rho2 * rho2, rho12 = rho6 * rho6;
  const double rc2i = 1.0 / rc2;
  const double c2 = s2 * rc2i, c6 = c2 * c2 * c2, c12 = c6 * c6;
  return 4.0 * eps * ((rho12 - rho6) + (6.0 * c12 - 3.0 * c6) * (r2 * rc2i) - (7.0 * c12 - 4.0 * c6));
}

void orc_lj(int64_t n, const double *pos, const int32_t *type, const int64_t *id,
            int64_t ng, const double *gpos, const int32_t *gtype, const int64_t *gid,
            const int64_t *offs, const uint32_t *nbr, const orc_lj_params *p, double *acc,
            double *virial, double *pe) {
  (void)ng;
  const int nt = p->ntypes;
  const double cutoff_sq = p->cutoff * p->cutoff; /* :397 */
  double c = cutoff_sq;
  double virial_local = 0.0; /* :401 */
  double pe_local = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double *pi = pos + 3 * i;
    const int ti = type[i];
    const double mass_inv_i = p->mass_inv[ti];
    const int id_i = (int)id[i]; /* :419 int id_i */
    for (int64_t e = offs[i]; e < offs[i + 1]; ++e) {
      int64_t j = nbr[e];
      const int is_ghost = j >= n; /* :424 */
      const double *pj;
      int tj, id_j;
      if (is_ghost) {
        j -= n;
        pj = gpos + 3 * j;
        tj = gtype[j];
        id_j = (int)gid[j];
      } else {
        pj = pos + 3 * j;
        tj = type[j];
        id_j = (int)id[j];
      }
      /* :446: equal molecule indices (-1 == -1 included) add nothing */
      if (p->ignore_intra_molecule && p->mol[i] == (is_ghost ? p->gmol[j] : p->mol[j])) continue;
      const double drx = pj[0] - pi[0], dry = pj[1] - pi[1], drz = pj[2] - pi[2]; /* :449 */
      double dr_sq = dot3(drx, dry, drz);                                       /* :451 */
      if (p->cutoff_list) {                                                      /* :453-454 */
        const double cl = p->cutoff_list[ti * nt + tj];
        c = cl * cl;
      }
      if (dr_sq > c) continue; /* :456 */
      if (p->jump_tol > 0.0 && dr_sq < p->jump_tol) dr_sq = p->jump_tol; /* :459-468 */
      const double eps_ij = p->eps[ti * nt + tj];
      if (eps_ij == 0) continue; /* :470-472 */
      const double sigma_ij = p->sigma[ti * nt + tj];
      /* :475-484, same association order */
      const double r_c_sq_inv = 1 / c;
      const double rho_c_sq_inv = sigma_ij * sigma_ij * r_c_sq_inv;
      const double rho_c_6_inv = rho_c_sq_inv * rho_c_sq_inv * rho_c_sq_inv;
      const double rho_c_12_inv = rho_c_6_inv * rho_c_6_inv;
      const double dr_sq_inv = 1 / dr_sq;
      const double rho_sq_inv = sigma_ij * sigma_ij * dr_sq_inv;
      const double rho_6_inv = rho_sq_inv * rho_sq_inv * rho_sq_inv;
      const double rho_12_inv = rho_6_inv * rho_6_inv;
      const double force_coef =
          4 * eps_ij *
          (-12 * rho_12_inv * dr_sq_inv + 6 * rho_6_inv * dr_sq_inv +
           12 * rho_c_12_inv * r_c_sq_inv - 6 * rho_c_6_inv * r_c_sq_inv);
      if (id_i < id_j) { /* :487-490 */
        virial_local += -force_coef * dr_sq;
        pe_local += orc_lj_pair_energy(dr_sq, eps_ij, sigma_ij, c);
      }
      const double fx = drx * force_coef, fy = dry * force_coef, fz = drz * force_coef; /* :491 */
      acc[3 * i + 0] += fx * mass_inv_i; /* :493 */
      acc[3 * i + 1] += fy * mass_inv_i;
      acc[3 * i + 2] += fz * mass_inv_i;
      if (!is_ghost) { /* :498-511 */
        const double mass_inv_j = p->mass_inv[tj];
        acc[3 * j + 0] -= fx * mass_inv_j;
        acc[3 * j + 1] -= fy * mass_inv_j;
        acc[3 * j + 2] -= fz * mass_inv_j;
      }
    }
  }
  if (virial) *virial += virial_local; /* :516 */
  if (pe) *pe += pe_local;
}

/* ------------------------------------------------------------------ */
/* kinetic energy: kinetic_energy.cpp:26-70, :409-453                  */
/* ------------------------------------------------------------------ */
double orc_kinetic_energy(int64_t n, const double *vel, const int32_t *type,
                          const double *mass, int n_r_df) {
  double e = 0.0;
  if (n_r_df == 0) {
    for (int64_t i = 0; i < n; ++i) {
      const double *v = vel + 3 * i;
      e += mass[type[i]] * dot3(v[0], v[1], v[2]);
    }
  } else {
    double vcm[3] = {0.0, 0.0, 0.0};
    double mass_sum = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double m = mass[type[i]];
      mass_sum += m;
      vcm[0] += vel[3 * i + 0] * m;
      vcm[1] += vel[3 * i + 1] * m;
      vcm[2] += vel[3 * i + 2] * m;
    }
    vcm[0] = vcm[0] / mass_sum;
    vcm[1] = vcm[1] / mass_sum;
    vcm[2] = vcm[2] / mass_sum;
    for (int64_t i = 0; i < n; ++i) {
      const double *v = vel + 3 * i;
      e += mass[type[i]] * dot3(v[0] - vcm[0], v[1] - vcm[1], v[2] - vcm[2]);
    }
  }
  return e * 0.5;
}

/* ------------------------------------------------------------------ */
/* whole-step driver                                                   */
/* ------------------------------------------------------------------ */
struct orc_sim {
  orc_sim_params p;
  int64_t n;
  double *pos, *vel, *acc, *acc_old;
  int32_t *type;
  int64_t *id;
  double *eps, *sigma, *cutoff_list, *mass, *mass_inv;
  /* ghosts */
  int64_t ng, gcap;
  double *gpos;
  int32_t *gtype;
  int64_t *gid, *gowner;
  /* neighbour list */
  int64_t *offs;
  uint32_t *nbr;
  int64_t ncap, npairs;
  double *pos_old_nl, *gpos_old_nl; /* Neighborlist::pos_old / ghost_pos_old */
  int64_t ng_old;
  double threshold_sq;
  /* thermo */
  double virial, pe, ke;
  /* Langevin coefficients, md_simulator.cpp:398-405 */
  double la, lb, lc;
  /* Lj::ignore_intra_molecule with the owned molecule indices; the ghosts
   * carry their owners' (exchange_ghost_single_md_domain.cpp:54-60) */
  int ignore_mol;
  int32_t *mol, *gmol;
  int64_t gmol_cap;
};

static void sim_ghosts(orc_sim *s) {
  for (;;) {
    int64_t k = orc_ghosts(s->n, s->pos, &s->p.box, s->p.ghost_cutoff, s->gcap, s->gpos,
                           s->gowner, NULL);
    if (k <= s->gcap) {
      s->ng = k;
      break;
    }
    s->gcap = k + k / 2 + 16;
    s->gpos = (double *)realloc(s->gpos, sizeof(double) * 3 * (size_t)s->gcap);
    s->gowner = (int64_t *)realloc(s->gowner, sizeof(int64_t) * (size_t)s->gcap);
    s->gtype = (int32_t *)realloc(s->gtype, sizeof(int32_t) * (size_t)s->gcap);
    s->gid = (int64_t *)realloc(s->gid, sizeof(int64_t) * (size_t)s->gcap);
  }
  for (int64_t g = 0; g < s->ng; ++g) {
    s->gtype[g] = s->type[s->gowner[g]];
    s->gid[g] = s->id[s->gowner[g]];
  }
}

static void sim_rebuild_list(orc_sim *s) {
  for (;;) {
    int64_t k;
    if (s->p.use_cell_list)
      k = orc_nlist_cell(s->n, s->pos, s->ng, s->gpos, &s->p.box, s->p.nl_cutoff, s->p.nl_coef,
                         s->ncap, s->offs, s->nbr);
    else
      k = orc_nlist_brute(s->n, s->pos, s->ng, s->gpos, s->p.nl_cutoff, s->p.nl_coef, s->ncap,
                          s->offs, s->nbr);
    if (k <= s->ncap) {
      s->npairs = k;
      break;
    }
    s->ncap = k + k / 4 + 64;
    s->nbr = (uint32_t *)realloc(s->nbr, sizeof(uint32_t) * (size_t)s->ncap);
  }
  /* pos_old = pos; ghost_pos_old = pos_ghost (neighborlist.cpp:292-293) */
  memcpy(s->pos_old_nl, s->pos, sizeof(double) * 3 * (size_t)s->n);
  free(s->gpos_old_nl);
  s->gpos_old_nl = (double *)malloc(sizeof(double) * 3 * (size_t)(s->ng > 0 ? s->ng : 1));
  memcpy(s->gpos_old_nl, s->gpos, sizeof(double) * 3 * (size_t)s->ng);
  s->ng_old = s->ng;
}

/* Neighborlist::update_is_needed, neighborlist.cpp:100-148 */
static int sim_update_is_needed(const orc_sim *s) {
  if (s->p.rebuild_every_step) return 1;
  if (s->ng != s->ng_old) return 1;
  for (int64_t i = 0; i < s->n; ++i) {
    const double *a = s->pos + 3 * i, *b = s->pos_old_nl + 3 * i;
    if (dot3(a[0] - b[0], a[1] - b[1], a[2] - b[2]) > s->threshold_sq) return 1;
  }
  for (int64_t i = 0; i < s->ng; ++i) {
    const double *a = s->gpos + 3 * i, *b = s->gpos_old_nl + 3 * i;
    if (dot3(a[0] - b[0], a[1] - b[1], a[2] - b[2]) > s->threshold_sq) return 1;
  }
  return 0;
}

static void sim_forces(orc_sim *s) {
  memset(s->acc, 0, sizeof(double) * 3 * (size_t)s->n); /* reset_owned_acceleration */
  orc_lj_params lj = {s->p.ntypes, s->eps, s->sigma, s->cutoff_list, s->p.lj_cutoff, s->mass_inv, 0.0, 0, NULL, NULL};
  if (s->ignore_mol) {
    if (s->ng > s->gmol_cap) {
      free(s->gmol);
      s->gmol_cap = s->ng;
      s->gmol = (int32_t *)malloc(sizeof(int32_t) * (size_t)s->gmol_cap);
    }
    for (int64_t g = 0; g < s->ng; ++g) s->gmol[g] = s->mol[s->gowner[g]];
    lj.ignore_intra_molecule = 1;
    lj.mol = s->mol;
    lj.gmol = s->gmol;
  }
  s->pe = 0.0;
  orc_lj(s->n, s->pos, s->type, s->id, s->ng, s->gpos, s->gtype, s->gid, s->offs, s->nbr, &lj,
         s->acc, &s->virial, &s->pe);
}

/* Md_simulator::re_calculate_acc, md_simulator.cpp:422-437 */
static int sim_recalc(orc_sim *s, int force_rebuild) {
  int update = orc_wrap(s->n, s->pos, &s->p.box, NULL) || force_rebuild;
  sim_ghosts(s);
  if (!update) update = sim_update_is_needed(s);
  if (update) sim_rebuild_list(s);
  sim_forces(s);
  return update;
}

static double *dup_d(const double *a, size_t k) {
  if (!a) return NULL;
  double *r = (double *)malloc(sizeof(double) * k);
  memcpy(r, a, sizeof(double) * k);
  return r;
}

orc_sim *orc_sim_create(const orc_sim_params *p, int64_t n, const double *pos,
                        const double *vel, const int32_t *type, const int64_t *id) {
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  s->p = *p;
  s->n = n;
  const size_t n3 = (size_t)(n > 0 ? n : 1) * 3;
  s->pos = dup_d(pos, 3 * (size_t)n);
  s->vel = dup_d(vel, 3 * (size_t)n);
  if (!s->pos) s->pos = (double *)calloc(n3, sizeof(double));
  if (!s->vel) s->vel = (double *)calloc(n3, sizeof(double));
  s->acc = (double *)calloc(n3, sizeof(double));
  s->acc_old = (double *)calloc(n3, sizeof(double));
  s->pos_old_nl = (double *)calloc(n3, sizeof(double));
  s->type = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  s->id = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i) {
    s->type[i] = type ? type[i] : 0;
    s->id[i] = id ? id[i] : i;
  }
  const int nt = p->ntypes;
  s->eps = dup_d(p->eps, (size_t)nt * nt);
  s->sigma = dup_d(p->sigma, (size_t)nt * nt);
  s->cutoff_list = dup_d(p->cutoff_list, (size_t)nt * nt);
  s->mass = dup_d(p->mass, (size_t)nt);
  s->mass_inv = (double *)malloc(sizeof(double) * (size_t)nt);
  for (int t = 0; t < nt; ++t) s->mass_inv[t] = 1.0 / s->mass[t]; /* atom_data.cpp:700-703 */
  s->p.eps = s->eps;
  s->p.sigma = s->sigma;
  s->p.cutoff_list = s->cutoff_list;
  s->p.mass = s->mass;
  s->offs = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  s->ncap = 64 * (n > 0 ? n : 1);
  s->nbr = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)s->ncap);
  s->gcap = n / 2 + 64;
  s->gpos = (double *)malloc(sizeof(double) * 3 * (size_t)s->gcap);
  s->gowner = (int64_t *)malloc(sizeof(int64_t) * (size_t)s->gcap);
  s->gtype = (int32_t *)malloc(sizeof(int32_t) * (size_t)s->gcap);
  s->gid = (int64_t *)malloc(sizeof(int64_t) * (size_t)s->gcap);
  /* Neighborlist::init, neighborlist.cpp:317-318 */
  const double ce = p->nl_cutoff * p->nl_coef;
  s->threshold_sq = 0.25 * (ce - p->nl_cutoff) * (ce - p->nl_cutoff);
  /* Md_simulator::initialize, md_simulator.cpp:398-405 */
  if (p->integrator == 1) {
    const double g = p->langevin_gamma, dt = p->dt;
    s->la = (2.0 - g * dt) / (2.0 + g * dt);
    s->lb = sqrt(p->langevin_kbt * g * 0.5 * dt);
    s->lc = 2.0 * dt / (2.0 + g * dt);
  }
  return s;
}

void orc_sim_destroy(orc_sim *s) {
  if (!s) return;
  free(s->pos); free(s->vel); free(s->acc); free(s->acc_old); free(s->pos_old_nl);
  free(s->gpos_old_nl); free(s->type); free(s->id); free(s->eps); free(s->sigma);
  free(s->cutoff_list); free(s->mass); free(s->mass_inv); free(s->offs); free(s->nbr);
  free(s->gpos); free(s->gowner); free(s->gtype); free(s->gid); free(s->mol); free(s->gmol);
  free(s);
}

void orc_sim_set_molecules(orc_sim *s, const int32_t *mol, int ignore) {
  free(s->mol);
  s->mol = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->n > 0 ? s->n : 1));
  for (int64_t i = 0; i < s->n; ++i) s->mol[i] = mol ? mol[i] : -1;
  s->ignore_mol = ignore;
}

static void sim_ke(orc_sim *s) {
  s->ke = orc_kinetic_energy(s->n, s->vel, s->type, s->mass, s->p.n_r_df);
}

/* Md_simulator::setup, md_simulator.cpp:565-611 */
void orc_sim_setup(orc_sim *s) {
  s->virial = 0.0; /* reset_virial */
  orc_wrap(s->n, s->pos, &s->p.box, NULL);
  sim_ghosts(s);
  sim_rebuild_list(s);
  sim_forces(s);
  sim_ke(s);
}

int orc_sim_step(orc_sim *s, const double *eta) {
  const double dt = s->p.dt;
  const int64_t n = s->n;
  int rebuilt;
  s->virial = 0.0; /* md_simulator.cpp:469 reset_virial */
  if (s->p.integrator == 0) {
    /* integrate_velocity_verlet, md_simulator.cpp:617-729 */
    for (int64_t i = 0; i < 3 * n; ++i) {
      const double a = s->acc[i];
      s->pos[i] += s->vel[i] * dt + a * 0.5 * dt * dt; /* :654 */
      s->acc_old[i] = a;                               /* :655 */
    }
    rebuilt = sim_recalc(s, 0);
    const double hdt = 0.5 * dt;
    for (int64_t i = 0; i < 3 * n; ++i) s->vel[i] += (s->acc[i] + s->acc_old[i]) * hdt; /* :678 */
  } else {
    /* integrate_velocity_verlet_langevin, md_simulator.cpp:731-869 */
    for (int64_t i = 0; i < 3 * n; ++i)
      s->vel[i] += s->acc[i] * 0.5 * dt + eta[i] * s->lb; /* :778 */
    for (int64_t i = 0; i < 3 * n; ++i) s->pos[i] += s->vel[i] * s->lc; /* :835 */
    rebuilt = sim_recalc(s, 0);
    for (int64_t i = 0; i < 3 * n; ++i)
      s->vel[i] = s->vel[i] * s->la + eta[i] * s->lb + s->acc[i] * 0.5 * dt; /* :860 */
  }
  sim_ke(s);
  return rebuilt;
}

void orc_sim_get(const orc_sim *s, double *pos, double *vel, double *acc) {
  const size_t b = sizeof(double) * 3 * (size_t)s->n;
  if (pos) memcpy(pos, s->pos, b);
  if (vel) memcpy(vel, s->vel, b);
  if (acc) memcpy(acc, s->acc, b);
}

double orc_nve_lambda(double ke_now, int dof, double energy_per_dof, double energy_tot,
                      double kbt, double kb, double temperature) {
  if (ke_now == 0.0) return 1.0;                                           /* :147-150 */
  if (energy_per_dof > 0.0) return sqrt(energy_per_dof * dof / ke_now);    /* :152-159 */
  if (energy_tot > 0.0) return sqrt(energy_tot / ke_now);                  /* :160-164 */
  if (kbt > 0.0) return sqrt(dof * kbt * 0.5 / ke_now);                    /* :165-172 */
  if (temperature > 0.0) return sqrt(dof * kb * temperature * 0.5 / ke_now); /* :173-179 */
  return 1.0;
}

void orc_sim_scale_velocities(orc_sim *s, double lambda) {
  for (int64_t i = 0; i < 3 * s->n; ++i) s->vel[i] *= lambda;
  sim_ke(s);
}

int64_t orc_sim_nghost(const orc_sim *s) { return s->ng; }
int64_t orc_sim_npairs(const orc_sim *s) { return s->npairs; }

void orc_sim_thermo(const orc_sim *s, double *ke, double *temperature, double *virial,
                    double *pe) {
  if (ke) *ke = s->ke;
  if (temperature) {
    /* finalize_temperature_mpi_domain, kinetic_energy.cpp:190-198 */
    const int dof = (int)(3 * s->n) - s->p.n_r_df;
    *temperature = 2.0 * s->ke / (s->p.k_b * dof);
  }
  if (virial) *virial = s->virial;
  if (pe) *pe = s->pe;
}
