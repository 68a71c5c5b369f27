/*
 * FP32 CPU twin of the scoring + minimization path, with the operation order of
 * the transform, the voxelisation and the pair distance pinned (explicit fmaf,
 * built with -ffp-contract=off) so that voxel indices and pair membership can be
 * compared bit for bit with the GPU.  OpenMP over poses; it also serves as the
 * CPU throughput baseline.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): used by tests/, smoke() and bench.py's
 * cpu_baseline leg; never by the product path.
 *
 * Semantics restated (sjdv1982/attract-jax):
 *   transform          test-grid-energrad.py:299-301
 *   voxel / inside     test-grid-energrad.py:216-222, 246-253
 *   trilinear          test-grid-energrad.py:183-194 (integer coordinate -> zero weight)
 *   grid gradient      test-grid-energrad.py:232-239 (minus interpolated stored gradient)
 *   nonbon             test-grid-energrad.py:95-111
 *   plateau force-shift test-grid-energrad.py:124-147
 *   quaternion pose    test4.py:396-401; ADAM loop test4.py:445-455 (optax defaults)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define T_GRID 0x1u
#define T_NB 0x2u
#define T_PLATEAU 0x4u
#define T_EXACT 0x8u
#define T_QT 0x10u

typedef struct TwinGrid {
  float origin[3];
  float spacing;
  int dim[3];
  int dim2[3];
  float gridextension;
  float plateaudis2;
  int ntypes;
  const float* inner; /* plain [t][x][y][z][4] or NULL */
  const float* outer; /* plain or NULL */
  const int32_t* counts;  /* [V] */
  const int64_t* starts;  /* [V] 0-based */
  const int32_t* lists;   /* flat receptor indices */
} TwinGrid;

typedef struct TwinFF {
  const float *rc, *ac, *ivor, *emin, *rmin2;
  int nrt, nlt;
} TwinFF;

static void transform(const float M[16], float x, float y, float z, float out[3]) {
  for (int l = 0; l < 3; ++l) out[l] = fmaf(z, M[8 + l], fmaf(y, M[4 + l], x * M[l])) + M[12 + l];
}

static void quat_to_mat(const float q[4], float R[9]) {
  const float x = q[0], y = q[1], z = q[2], w = q[3];
  const float x2 = x * x, y2 = y * y, z2 = z * z, w2 = w * w;
  const float xy = x * y, zw = z * w, xz = x * z, yw = y * w, yz = y * z, xw = x * w;
  R[0] = ((x2 - y2) - z2) + w2;
  R[1] = 2.f * (xy - zw);
  R[2] = 2.f * (xz + yw);
  R[3] = 2.f * (xy + zw);
  R[4] = ((-x2 + y2) - z2) + w2;
  R[5] = 2.f * (yz - xw);
  R[6] = 2.f * (xz - yw);
  R[7] = 2.f * (yz + xw);
  R[8] = ((-x2 - y2) + z2) + w2;
}

static void pose_matrix(const float* p, int qt, float M[16]) {
  if (qt) {
    float R[9];
    quat_to_mat(p, R);
    const float m[16] = {R[0], R[1], R[2], 1.f, R[3], R[4], R[5], 1.f, R[6], R[7], R[8], 1.f, p[4], p[5], p[6], 1.f};
    memcpy(M, m, sizeof m);
  } else {
    memcpy(M, p, 16 * sizeof(float));
  }
}

static void trilinear(const float* g, int t, const int d[3], const float v[3], int exact, float raw[4],
                      float dEdv[3]) {
  float f[3], c[3], w[3][2];
  int i0[3];
  for (int a = 0; a < 3; ++a) {
    f[a] = floorf(v[a]);
    c[a] = ceilf(v[a]);
    i0[a] = (int)f[a];
    w[a][0] = c[a] - v[a];
    w[a][1] = v[a] - f[a];
  }
  for (int k = 0; k < 4; ++k) raw[k] = 0.f;
  for (int k = 0; k < 3; ++k) dEdv[k] = 0.f;
  for (int cx = 0; cx < 2; ++cx)
    for (int cy = 0; cy < 2; ++cy)
      for (int cz = 0; cz < 2; ++cz) {
        const size_t idx = (((size_t)t * d[0] + i0[0] + cx) * d[1] + i0[1] + cy) * (size_t)d[2] + i0[2] + cz;
        const float* val = g + idx * 4;
        const float wt = (w[0][cx] * w[1][cy]) * w[2][cz];
        for (int k = 0; k < 4; ++k) raw[k] = fmaf(wt, val[k], raw[k]);
        if (exact) {
          dEdv[0] += (cx ? 1.f : -1.f) * (w[1][cy] * w[2][cz]) * val[0];
          dEdv[1] += (cy ? 1.f : -1.f) * (w[0][cx] * w[2][cz]) * val[0];
          dEdv[2] += (cz ? 1.f : -1.f) * (w[0][cx] * w[1][cy]) * val[0];
        }
      }
}

/* grid term of one bead; returns inner voxel (-1 outside) */
static int bead_grid(const TwinGrid* G, int gt, const float xp[3], unsigned flags, float* E, float g[3],
                     float raw4[4]) {
  float v[3];
  v[0] = (xp[0] - G->origin[0]) / G->spacing;
  v[1] = (xp[1] - G->origin[1]) / G->spacing;
  v[2] = (xp[2] - G->origin[2]) / G->spacing;
  int inside = 1;
  for (int a = 0; a < 3; ++a) inside &= (v[a] >= 0.f) && (v[a] < (float)(G->dim[a] - 1));
  int voxel = -1;
  if (inside) {
    const int ix = (int)floorf(v[0] + 0.5f), iy = (int)floorf(v[1] + 0.5f), iz = (int)floorf(v[2] + 0.5f);
    voxel = (ix * G->dim[1] + iy) * G->dim[2] + iz;
  }
  *E = 0.f;
  g[0] = g[1] = g[2] = 0.f;
  for (int k = 0; k < 4; ++k) raw4[k] = 0.f;
  if (!(flags & T_GRID) || !G->inner) return voxel;
  float raw[4], dEdv[3], inv;
  const int exact = (flags & T_EXACT) != 0;
  if (inside) {
    trilinear(G->inner, gt, G->dim, v, exact, raw, dEdv);
    inv = 1.f / G->spacing;
  } else if (G->outer) {
    float vo[3];
    int in_outer = 1;
    for (int a = 0; a < 3; ++a) {
      vo[a] = (v[a] + G->gridextension) / 2.f;
      in_outer &= (vo[a] >= 0.f) && (vo[a] < (float)(G->dim2[a] - 1));
    }
    if (!in_outer) return voxel;
    trilinear(G->outer, gt, G->dim2, vo, exact, raw, dEdv);
    inv = 0.5f / G->spacing;
  } else {
    return voxel;
  }
  *E = raw[0];
  for (int a = 0; a < 3; ++a) g[a] = exact ? dEdv[a] * inv : -raw[1 + a];
  for (int k = 0; k < 4; ++k) raw4[k] = raw[k];
  return voxel;
}

/* nonbon and its dsq derivative */
static void nonbon(float dsq, float rc, float ac, float iv, float em, float r2, float* E, float* dE) {
  const float u = 1.f / dsq;
  const float u3 = (u * u) * u;
  const float vlj = (rc * u - ac) * u3;
  const float dv = -((4.f * rc * u - 3.f * ac) * u3) * u;
  const int lo = dsq < r2;
  *E = lo ? vlj + (iv - 1.f) * em : iv * vlj;
  *dE = lo ? dv : iv * dv;
}

/* neighbour term of one bead; returns number of evaluated pairs */
static int bead_nb(const TwinGrid* G, const TwinFF* ff, const float* rec4, int lt, int voxel, const float xp[3],
                   unsigned flags, float* E, float g[3]) {
  *E = 0.f;
  g[0] = g[1] = g[2] = 0.f;
  if (voxel < 0 || !(flags & T_NB)) return 0;
  const int plateau = (flags & T_PLATEAU) != 0, exact = (flags & T_EXACT) != 0;
  const float P2 = G->plateaudis2;
  const float P = sqrtf(P2);
  const int n = G->counts[voxel];
  const int64_t s = G->starts[voxel];
  int np = 0;
  for (int c = 0; c < n; ++c) {
    const int r = G->lists[s + c];
    const float* rb = rec4 + 4 * (size_t)r;
    const float dx = xp[0] - rb[0], dy = xp[1] - rb[1], dz = xp[2] - rb[2];
    const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    if (!(dsq < P2)) continue;
    int rt;
    memcpy(&rt, rb + 3, 4);
    const int pi = rt * ff->nlt + lt;
    float e, de;
    nonbon(dsq, ff->rc[pi], ff->ac[pi], ff->ivor[pi], ff->emin[pi], ff->rmin2[pi], &e, &de);
    float gs = 2.f * de;
    if (plateau) {
      float pe, pde;
      nonbon(P2, ff->rc[pi], ff->ac[pi], ff->ivor[pi], ff->emin[pi], ff->rmin2[pi], &pe, &pde);
      e -= pe;
      if (!exact) gs -= 2.f * pde * P / sqrtf(dsq);
    }
    *E += e;
    g[0] += gs * dx;
    g[1] += gs * dy;
    g[2] += gs * dz;
    ++np;
  }
  return np;
}

static void pose_energrad(const TwinGrid* G, const TwinFF* ff, const float* rec4, const float* lig4, int nb,
                          const float M[16], unsigned flags, float acc[13]) {
  for (int i = 0; i < 13; ++i) acc[i] = 0.f;
  for (int j = 0; j < nb; ++j) {
    const float* b = lig4 + 4 * (size_t)j;
    int tb;
    memcpy(&tb, b + 3, 4);
    float xp[3], Eg, gg[3], raw[4], En, gn[3];
    transform(M, b[0], b[1], b[2], xp);
    const int voxel = bead_grid(G, tb & 0xffff, xp, flags, &Eg, gg, raw);
    bead_nb(G, ff, rec4, tb >> 16, voxel, xp, flags, &En, gn);
    const float g[3] = {gg[0] + gn[0], gg[1] + gn[1], gg[2] + gn[2]};
    acc[0] += Eg + En;
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) acc[1 + 3 * k + l] += b[k] * g[l];
    for (int l = 0; l < 3; ++l) acc[10 + l] += g[l];
  }
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

int twin_energrad(const TwinGrid* G, const TwinFF* ff, const float* rec4, const float* lig4, int nb,
                  const float* poses, int N, unsigned flags, float* energy, float* dEdM, int nthreads) {
  set_threads(nthreads);
  const int qt = (flags & T_QT) != 0;
#pragma omp parallel for schedule(dynamic, 16)
  for (int p = 0; p < N; ++p) {
    float M[16], acc[13];
    pose_matrix(poses + (size_t)p * (qt ? 7 : 16), qt, M);
    pose_energrad(G, ff, rec4, lig4, nb, M, flags, acc);
    energy[p] = acc[0];
    if (dEdM) {
      float* o = dEdM + (size_t)p * 16;
      for (int k = 0; k < 3; ++k) {
        for (int l = 0; l < 3; ++l) o[4 * k + l] = acc[1 + 3 * k + l];
        o[4 * k + 3] = 0.f;
      }
      o[12] = acc[10];
      o[13] = acc[11];
      o[14] = acc[12];
      o[15] = 0.f;
    }
  }
  return 0;
}

int twin_bead_terms(const TwinGrid* G, const TwinFF* ff, const float* rec4, const float* lig4, int nb,
                    const float* poses, int N, unsigned flags, float* grid4, int32_t* voxel, float* nbene,
                    float* nbgrad, int32_t* npairs, float* xp_out, int nthreads) {
  set_threads(nthreads);
  const int qt = (flags & T_QT) != 0;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < N; ++p) {
    float M[16];
    pose_matrix(poses + (size_t)p * (qt ? 7 : 16), qt, M);
    for (int j = 0; j < nb; ++j) {
      const size_t i = (size_t)p * nb + j;
      const float* b = lig4 + 4 * (size_t)j;
      int tb;
      memcpy(&tb, b + 3, 4);
      float xp[3], Eg, gg[3], raw[4], En, gn[3];
      transform(M, b[0], b[1], b[2], xp);
      const int vx = bead_grid(G, tb & 0xffff, xp, flags, &Eg, gg, raw);
      const int np = bead_nb(G, ff, rec4, tb >> 16, vx, xp, flags, &En, gn);
      if (grid4) memcpy(grid4 + 4 * i, raw, sizeof raw);
      if (voxel) voxel[i] = vx;
      if (nbene) nbene[i] = En;
      if (nbgrad) memcpy(nbgrad + 3 * i, gn, sizeof gn);
      if (npairs) npairs[i] = np;
      if (xp_out) memcpy(xp_out + 3 * i, xp, sizeof xp);
    }
  }
  return 0;
}

/* Rigid-body minimization; pose7 [N][7] updated in place, energy = last evaluated energy
 * (gravity term included).  kind 0: optax.adam (test4.py:447-453); kind 1: optax.adabelief
 * (test5.py:451-457; second moment of g - m_prev, plus eps_root); gravity adds
 * gravity |t|^2 to the energy (test5.py:418-420).  Bias corrections by powf as optax
 * computes 1 - b**count. */
/* bias_pow = 0: bias corrections 1 - b^k by the recurrence c_k = b c_{k-1} + (1 - b), as the
 * GPU minimizer computes them; 1: by powf (the two differ by rounding only). */
int twin_minimize(const TwinGrid* G, const TwinFF* ff, const float* rec4, const float* lig4, int nb, float* pose7,
                  int N, int nsteps, float lr_rot, float lr_trans, float b1, float b2, float eps, int kind,
                  float eps_root, float gravity, unsigned flags, float* energy, int nthreads, int bias_pow) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
  for (int p = 0; p < N; ++p) {
    float* P = pose7 + (size_t)p * 7;
    float m[7] = {0}, v[7] = {0};
    float e = 0.f, c1 = 0.f, c2 = 0.f;
    for (int k = 1; k <= nsteps; ++k) {
      float R[9], acc[13];
      quat_to_mat(P, R);
      const float M[16] = {R[0], R[1], R[2], 1.f, R[3], R[4], R[5], 1.f, R[6], R[7], R[8], 1.f, P[4], P[5], P[6], 1.f};
      pose_energrad(G, ff, rec4, lig4, nb, M, flags, acc);
      e = acc[0];
      const float* A = acc + 1;
      const float x = P[0], y = P[1], z = P[2], w = P[3];
      float gr[7];
      gr[0] = 2.f * (x * ((A[0] - A[4]) - A[8]) + y * (A[1] + A[3]) + z * (A[2] + A[6]) + w * (A[7] - A[5]));
      gr[1] = 2.f * (y * ((A[4] - A[0]) - A[8]) + x * (A[1] + A[3]) + z * (A[5] + A[7]) + w * (A[2] - A[6]));
      gr[2] = 2.f * (z * ((A[8] - A[0]) - A[4]) + x * (A[2] + A[6]) + y * (A[5] + A[7]) + w * (A[3] - A[1]));
      gr[3] = 2.f * (w * ((A[0] + A[4]) + A[8]) + x * (A[7] - A[5]) + y * (A[2] - A[6]) + z * (A[3] - A[1]));
      gr[4] = acc[10];
      gr[5] = acc[11];
      gr[6] = acc[12];
      if (gravity != 0.f) {
        const float n = sqrtf(fmaf(P[6], P[6], fmaf(P[5], P[5], P[4] * P[4])));
        e = e + (n * n) * gravity;
        for (int i = 4; i < 7; ++i) gr[i] = gr[i] + (2.f * gravity) * P[i];
      }
      c1 = fmaf(b1, c1, 1.f - b1);
      c2 = fmaf(b2, c2, 1.f - b2);
      const float bc1 = bias_pow ? 1.f - powf(b1, (float)k) : c1, bc2 = bias_pow ? 1.f - powf(b2, (float)k) : c2;
      for (int i = 0; i < 7; ++i) {
        const float pe = kind == 1 ? gr[i] - m[i] : gr[i];
        m[i] = (1.f - b1) * gr[i] + b1 * m[i];
        v[i] = (1.f - b2) * (pe * pe) + b2 * v[i];
        if (kind == 1) v[i] = v[i] + eps_root;
        const float upd = (m[i] / bc1) / (sqrtf(v[i] / bc2) + eps);
        P[i] = P[i] + (-(i < 4 ? lr_rot : lr_trans)) * upd;
      }
      const float nrm = sqrtf(fmaf(P[3], P[3], fmaf(P[2], P[2], fmaf(P[1], P[1], P[0] * P[0]))));
      for (int i = 0; i < 4; ++i) P[i] = P[i] / nrm;
    }
    energy[p] = e;
  }
  return 0;
}
