// libattract_amd: ATTRACT rigid-body docking scoring and minimization for gfx950.
//
// One wavefront (64 lanes) evaluates one pose at a time and takes claims of up to 8
// poses -- its first by its index, later ones from a device work counter -- until the
// list is exhausted (persistent kernel).  For
// each chunk of up to 64 ligand beads:
//   * bead phase  -- lane = bead: rigid transform in registers, grid trilinear
//     term (one 128-byte record per bead in the corner-packed layout), voxel of
//     the neighbour lookup, (count, start) of its receptor list;
//   * pair phase  -- the wave walks the concatenation of all the chunk's lists
//     in segments of 8 receptor indices (4 in a short last window; lane =
//     segment), so lanes stay
//     busy whatever the per-bead list lengths; receptor beads and the pair table
//     are read from LDS.
// Energy and dE/dM (13 floats) are accumulated per lane and reduced together by a
// transposed butterfly of lane swaps, in a fixed order.  Two consecutive poses of at
// most 32 beads share one bead phase (lanes 0-31 and 32-63), each pose's sums still
// formed exactly as when it is evaluated alone.  The minimizer gives each wavefront a
// batch of 8 poses: per optimizer step it evaluates them in order, then updates them
// together, one lane per (pose, parameter), with the state in LDS.
//
// The file is compiled with -ffp-contract=off: every fused multiply-add below
// is written explicitly, so the CPU twin can pin the operation order of the
// transform and the voxelisation bit for bit.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>

#include "attract_amd.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess)                                                      \
      return fail(ATJ_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kWave = 64;
constexpr int kListPad = 16;       // lists are padded with 0xFFFF to a multiple of this
constexpr int kHalf = 32;          // poses of at most this many beads share a bead phase two by two
constexpr int kMaxShell = 6;       // most list shells per grid point (16-bit counts in one int4)
constexpr double kShellMargin = 1e-3;  // A added to every shell radius: covers FP32 rounding of distances
constexpr int kMaxWavesPerBlock = 16;
constexpr int kAcc = 13;           // E, dE/dM[0..2][0..2], dE/dM[3][0..2]
constexpr int kPairCost = 30;      // approximate instructions per neighbour slot
constexpr int kItemCost = 90;      // approximate instructions per list segment

// kAdjoint sums the pair-table gradient in one fixed-point table per block; kAdjointWave in
// float tables of each wavefront's own (the form for pair tables too large for the first).
enum Mode { kEnergrad = 0, kMinimize = 1, kAdjoint = 2, kAdjointWave = 3 };
__host__ __device__ constexpr bool is_adjoint(int mode) { return mode >= kAdjoint; }
// Single-shot launches (energrad, adjoint) take their claims in a fixed order, with no work
// counter: wavefront g of G takes claims g, g + G, g + 2G, ...  Atomics on one counter
// serialise (about 12 ns each), and the ~12,000 claims of a 131,072-pose energrad took
// 17% of its time.  The minimizer claims through the counter, largest ligands first, which
// evens out its long launches over ligands of different sizes.
__host__ __device__ constexpr bool static_claims(int mode) { return mode != kMinimize; }

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value known to be identical on every lane: move it to a scalar register.
__device__ __forceinline__ float uni(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Value of lane `l` (uniform result in a scalar register).
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// DPP move with zero for lanes outside row_mask.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dppi(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, false);
}

// Sum over the 64 lanes in a fixed order (quad, half-row and row mirrors, then the
// row broadcasts of lanes 15 and 31); the total is returned uniformly from lane 63.
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp<0xB1>(v);        // quad_perm [1,0,3,2]
  v += dpp<0x4E>(v);        // quad_perm [2,3,0,1]
  v += dpp<0x141>(v);       // row_half_mirror
  v += dpp<0x140>(v);       // row_mirror
  v += dpp<0x142, 0xA>(v);  // row_bcast:15 -> rows 1, 3
  v += dpp<0x143, 0xC>(v);  // row_bcast:31 -> rows 2, 3
  return lane_value(v, 63);
}

__device__ __forceinline__ unsigned wave_sum_u(unsigned v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

// Inclusive prefix sum over the 64 lanes (row shifts, then row broadcasts).
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += dppi<0x111>(v);       // row_shr:1
  v += dppi<0x112>(v);       // row_shr:2
  v += dppi<0x114>(v);       // row_shr:4
  v += dppi<0x118>(v);       // row_shr:8
  v += dppi<0x142, 0xA>(v);  // row_bcast:15
  v += dppi<0x143, 0xC>(v);  // row_bcast:31
  return v;
}

// Inclusive prefix maximum over the 64 lanes (row shifts, then row broadcasts).  Lanes
// without a source lane take the maximum with 0, which callers tolerate (values >= -1,
// combined with a non-negative carry).
__device__ __forceinline__ int wave_max_scan(int v) {
  v = max(v, dppi<0x111>(v));
  v = max(v, dppi<0x112>(v));
  v = max(v, dppi<0x114>(v));
  v = max(v, dppi<0x118>(v));
  v = max(v, dppi<0x142, 0xA>(v));
  v = max(v, dppi<0x143, 0xC>(v));
  return v;
}

// Sums over the wave of all kAcc accumulators at once (a transposed butterfly): each
// halving stage pairs two accumulators, one lane half keeps the pair sum of the first and
// the other half that of the second, so a pair costs one lane swap and one add.  Stages:
// lanes l <-> l^32, rows 0 <-> 1 and 2 <-> 3, then in-row mirrors l <-> 15-l and
// l <-> 7-l, then quad swaps l <-> l^2, l <-> l^1.  The returned value of lane l is the
// sum of accumulator acc_index(l) (every lane of a quad holds it); the order of the
// additions is fixed, so results are reproducible.
__device__ __forceinline__ void swap_add(float a, float b, float& out, bool rows32) {
  const auto r = rows32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false)
                        : __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  out = __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

template <int CTRL>
__device__ __forceinline__ float mirror_add(float a, float b, bool upper) {
  return (upper ? b : a) + dpp<CTRL>(upper ? a : b);
}

__device__ __forceinline__ float wave_sum_all(const float acc[kAcc], int lane) {
  static_assert(kAcc == 13, "reduction schedule is laid out for 13 accumulators");
  float s1[7], s2[4], s3[2];
#pragma unroll
  for (int k = 0; k < 7; ++k) swap_add(acc[2 * k], 2 * k + 1 < kAcc ? acc[2 * k + 1] : 0.f, s1[k], true);
#pragma unroll
  for (int k = 0; k < 4; ++k) swap_add(s1[2 * k], 2 * k + 1 < 7 ? s1[2 * k + 1] : 0.f, s2[k], false);
#pragma unroll
  for (int k = 0; k < 2; ++k) s3[k] = mirror_add<0x140>(s2[2 * k], s2[2 * k + 1], lane & 8);  // row_mirror
  float u = mirror_add<0x141>(s3[0], s3[1], lane & 4);                                        // row_half_mirror
  u += dpp<0x4E>(u);  // quad_perm [2,3,0,1]
  u += dpp<0xB1>(u);  // quad_perm [1,0,3,2]
  return u;
}

// Accumulator whose sum wave_sum_all leaves in lane l: lane bits (5, 4, 3, 2) are its
// bits (0, 1, 2, 3).  Values 13..15 are the zero padding.
__device__ __forceinline__ int acc_index(int l) {
  return ((l >> 5) & 1) | (((l >> 4) & 1) << 1) | (((l >> 3) & 1) << 2) | (((l >> 2) & 1) << 3);
}

// Rotation matrix of a (not necessarily normalised) quaternion (x, y, z, w):
// the polynomial of scipy / jax.scipy Rotation.as_matrix, used by test4.py:399.
__device__ __forceinline__ void quat_to_mat(const float q[4], float R[9]) {
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

// dE/dq from A = dE/dR through the as_matrix polynomial.
__device__ __forceinline__ void quat_grad(const float q[4], const float A[9], float g[4]) {
  const float x = q[0], y = q[1], z = q[2], w = q[3];
  g[0] = 2.f * (x * ((A[0] - A[4]) - A[8]) + y * (A[1] + A[3]) + z * (A[2] + A[6]) + w * (A[7] - A[5]));
  g[1] = 2.f * (y * ((A[4] - A[0]) - A[8]) + x * (A[1] + A[3]) + z * (A[5] + A[7]) + w * (A[2] - A[6]));
  g[2] = 2.f * (z * ((A[8] - A[0]) - A[4]) + x * (A[2] + A[6]) + y * (A[5] + A[7]) + w * (A[3] - A[1]));
  g[3] = 2.f * (w * ((A[0] + A[4]) + A[8]) + x * (A[7] - A[5]) + y * (A[2] - A[6]) + z * (A[3] - A[1]));
}

// x'_l = ((x M0l + y M1l) + z M2l) + M3l with the first three products fused:
// fma(z, M2l, fma(y, M1l, x*M0l)) + M3l.  The CPU twin uses the same order.
__device__ __forceinline__ void transform(const float M[16], float x, float y, float z, float out[3]) {
#pragma unroll
  for (int l = 0; l < 3; ++l) out[l] = fmaf(z, M[8 + l], fmaf(y, M[4 + l], x * M[l])) + M[12 + l];
}

struct GridView {
  float ox, oy, oz, spacing, ge, p2;
  float rspacing;                   // RN(1 / spacing), for div_rn
  float lx, ly, lz, lx2, ly2, lz2;  // dim - 1 as floats: the exclusive upper bounds of the inside tests
  int dx, dy, dz, dx2, dy2, dz2, ntypes, layout;
  int nshell;            // 0: lists from nbgrid; > 0: shell-ordered lists, starts and lengths from nbshell
  float shinv;           // 1 / shell width in cell units: 2 nshell / sqrt(3)
  unsigned slab, slab2;  // records (corner) or points (plain) per atom type, inner / outer
  const float* inner;
  const float* outer;
  const int2* nbgrid;
  const uint16_t* nblist;
  const int4* nbshell;  // per grid point: the start, then 16-bit counts c0..c5 (uint16 words 2..7)
};

// a / b rounded to nearest, as IEEE division gives it, from y = RN(1/b) in three operations
// (Markstein's correction step: with y within half an ulp of 1/b and q = RN(a y) within
// one ulp of a/b, r = a - b q is exact in one fma and RN(q + r y) = RN(a/b)).  It needs r
// not to underflow: callers take a true division for |a| < 2^-99 (checked bit for bit
// against IEEE division for every float a in [2^-99, 4096) and several divisors).  Used for
// the voxel coordinates (x' - origin) / spacing of test-grid-energrad.py:216-218, whose
// voxel and trilinear weights must match a true division bit for bit.
__device__ __forceinline__ float div_rn(float a, float b, float y) {
  const float q = a * y;
  const float r = fmaf(-q, b, a);
  return fmaf(r, y, q);
}

// Trilinear interpolation over one cell (test-grid-energrad.py:183-194):
// x0 = floor(v), x1 = ceil(v), w0 = x1 - v, w1 = v - x0; corners x -> y -> z.
// raw = sum_c w_c G_c (4 channels); dEdv = d(sum_c w_c E_c)/dv (exact mode).
// Cell and point indices within one type's slab are 32-bit; one 64-bit multiply-add
// per bead places the slab.  Requires 0 <= v < dim - 1 on every axis.
template <int LAYOUT, bool ENERGY = true>
__device__ __forceinline__ void trilinear(const float* g, unsigned slab, int t, int dx, int dy, int dz, float vx,
                                          float vy, float vz, bool exact, float raw[4], float dEdv[3]) {
  const float fx = floorf(vx), fy = floorf(vy), fz = floorf(vz);
  const float cx = ceilf(vx), cy = ceilf(vy), cz = ceilf(vz);
  const unsigned x0 = (unsigned)fx, y0 = (unsigned)fy, z0 = (unsigned)fz;
  const float wx[2] = {cx - vx, vx - fx};
  const float wy[2] = {cy - vy, vy - fy};
  const float wz[2] = {cz - vz, vz - fz};
  float4 c[8];
  if (LAYOUT == ATJ_LAYOUT_CORNER) {
    const unsigned cell = (x0 * (unsigned)(dy - 1) + y0) * (unsigned)(dz - 1) + z0;
    const float4* p = reinterpret_cast<const float4*>(g) + ((size_t)(unsigned)t * slab + cell) * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) c[i] = p[i];
  } else {
    const unsigned pt = (x0 * (unsigned)dy + y0) * (unsigned)dz + z0;
    const float4* p = reinterpret_cast<const float4*>(g) + ((size_t)(unsigned)t * slab + pt);
    const unsigned sy = (unsigned)dz, sx = (unsigned)(dy * dz);
#pragma unroll
    for (int i = 0; i < 8; ++i) c[i] = p[(i >> 2) * sx + ((i >> 1) & 1) * sy + (i & 1)];
  }
  raw[0] = raw[1] = raw[2] = raw[3] = 0.f;
  dEdv[0] = dEdv[1] = dEdv[2] = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int ax = i >> 2, ay = (i >> 1) & 1, az = i & 1;
    const float w = (wx[ax] * wy[ay]) * wz[az];
    if (ENERGY || exact) raw[0] = fmaf(w, c[i].x, raw[0]);
    raw[1] = fmaf(w, c[i].y, raw[1]);
    raw[2] = fmaf(w, c[i].z, raw[2]);
    raw[3] = fmaf(w, c[i].w, raw[3]);
    if (exact) {
      const float sx = ax ? 1.f : -1.f, sy = ay ? 1.f : -1.f, sz = az ? 1.f : -1.f;
      dEdv[0] = fmaf(sx * (wy[ay] * wz[az]), c[i].x, dEdv[0]);
      dEdv[1] = fmaf(sy * (wx[ax] * wz[az]), c[i].x, dEdv[1]);
      dEdv[2] = fmaf(sz * (wx[ax] * wy[ay]), c[i].x, dEdv[2]);
    }
  }
}

// Grid term of one bead at x'.  Returns E and dE/dx', the inner voxel of the neighbour
// lookup (-1 outside the inner grid, test-grid-energrad.py:216-218,246-253) and, with
// do_lists, the start and length of the list the bead reads (0 outside): the voxel's whole
// list, or with shell-ordered lists the leading entries of shell
// min(floor(distance from the voxel's grid point / shell width), nshell - 1).  The list
// entry is loaded before the grid corners so that both loads are in flight together.
// Without ENERGY (minimizer steps whose energy is not kept) E is left 0 in non-exact mode:
// the energy channel is not interpolated, the gradient is the same.
template <int LAYOUT, bool ENERGY = true>
__device__ __forceinline__ void bead_grid(const GridView& G, int gtype, const float xp[3], bool do_grid,
                                          bool do_lists, bool exact, float& E, float g[3], int& voxel, int& cnt,
                                          int& start, unsigned& in_any, float raw_out[4]) {
  const float ax = xp[0] - G.ox, ay = xp[1] - G.oy, az = xp[2] - G.oz;
  float vx = div_rn(ax, G.spacing, G.rspacing);
  float vy = div_rn(ay, G.spacing, G.rspacing);
  float vz = div_rn(az, G.spacing, G.rspacing);
  if (fminf(fminf(fabsf(ax), fabsf(ay)), fabsf(az)) < 0x1p-99f) {  // zero or tiny: r may underflow
    vx = ax / G.spacing;
    vy = ay / G.spacing;
    vz = az / G.spacing;
  }
  const bool inside = (vx >= 0.f) && (vx < G.lx) && (vy >= 0.f) && (vy < G.ly) && (vz >= 0.f) && (vz < G.lz);
  E = 0.f;
  g[0] = g[1] = g[2] = 0.f;
  raw_out[0] = raw_out[1] = raw_out[2] = raw_out[3] = 0.f;
  in_any = 0;
  voxel = -1;
  cnt = start = 0;
  if (inside) {
    const float wx = vx + 0.5f, wy = vy + 0.5f, wz = vz + 0.5f;
    const float fx = floorf(wx), fy = floorf(wy), fz = floorf(wz);
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    voxel = (ix * G.dy + iy) * G.dz + iz;
    if (do_lists) {
      if (G.nshell > 0) {
        // offset from the grid point in cell units (w - floor(w) is exact), |o| <= 1/2 per axis
        const float ox = (wx - fx) - 0.5f, oy = (wy - fy) - 0.5f, oz = (wz - fz) - 0.5f;
        const float d = __builtin_amdgcn_sqrtf(fmaf(oz, oz, fmaf(oy, oy, ox * ox)));
        const int shell = min((int)(d * G.shinv), G.nshell - 1);
        // two loads from the entry's 16 bytes: the start (word 0) and count `shell` (uint16 2 + shell)
        const int4* e = G.nbshell + voxel;
        start = *reinterpret_cast<const int*>(e);
        cnt = reinterpret_cast<const uint16_t*>(e)[2 + shell];
      } else {
        const int2 e = G.nbgrid[voxel];
        cnt = e.x;
        start = e.y;
      }
    }
  }
  if (!do_grid || G.inner == nullptr) return;
  float raw[4], dEdv[3];
  float inv = 0.f;
  if (inside) {
    trilinear<LAYOUT, ENERGY>(G.inner, G.slab, gtype, G.dx, G.dy, G.dz, vx, vy, vz, exact, raw, dEdv);
    if (exact) inv = 1.f / G.spacing;
    in_any = 1;
  } else if (G.outer != nullptr) {
    const float ox = (vx + G.ge) / 2.f, oy = (vy + G.ge) / 2.f, oz = (vz + G.ge) / 2.f;
    const bool in_outer =
        (ox >= 0.f) && (ox < G.lx2) && (oy >= 0.f) && (oy < G.ly2) && (oz >= 0.f) && (oz < G.lz2);
    if (!in_outer) return;
    trilinear<LAYOUT, ENERGY>(G.outer, G.slab2, gtype, G.dx2, G.dy2, G.dz2, ox, oy, oz, exact, raw, dEdv);
    if (exact) inv = 0.5f / G.spacing;
    in_any = 1;
  } else {
    return;
  }
  E = raw[0];
  if (exact) {
    g[0] = dEdv[0] * inv;
    g[1] = dEdv[1] * inv;
    g[2] = dEdv[2] * inv;
  } else {
    g[0] = -raw[1];
    g[1] = -raw[2];
    g[2] = -raw[3];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) raw_out[i] = raw[i];
}

// Launch-uniform plateau constants: P2 = plateaudis^2 and powers of 1/P2 used to
// evaluate the plateau energy and its force-shift factor of any pair record.
struct PlateauConst {
  float P2, uP3, uP4, k1, k2;  // k1 = -8 P uP^5, k2 = 6 P uP^4 (2 P d vlj/d dsq at P2 = rc k1 + ac k2)
};

__device__ __forceinline__ PlateauConst plateau_const(float P2) {
  const float uP = 1.f / P2;
  const float uP2 = uP * uP;
  const float uP4 = uP2 * uP2;
  const float P = sqrtf(P2);
  PlateauConst c;
  c.P2 = P2;
  c.uP3 = uP2 * uP;
  c.uP4 = uP4;
  c.k1 = -8.f * P * (uP4 * uP);
  c.k2 = 6.f * P * uP4;
  return c;
}

// Pair-type records: a 16-byte record (rc, -3/4 ac, r2, gh / 4), an 8-byte side record
// (ac, kl) and a 4-byte kh.  ATTRACT's pair energy (test-grid-energrad.py:95-111) is
// vlj + (ivor - 1) emin for dsq < rmin2 and ivor vlj beyond, ivor = +-1 (ipon,
// attract-original/convert-attract-par.py:14-15); an ivor of another magnitude poisons the
// record (NaN energies) rather than being misread.  With w = (ivor - 1) emin,
// lo = dsq < rmin2 and s = lo ? 1 : ivor, both branches are s vlj + k with k = kl = w on the
// low side and k = kh = 0 beyond (each minus E(plateaudis) with the plateau); vlj = t1 u^3
// with t1 = rc u - ac, so E = t1 (s u^3) + k, and s folds into u^3 by one select.  For
// ivor = +1, w = 0 and s = 1 on both sides of rmin2, so the record stores r2 = FLT_MAX and
// lo always holds; for ivor = -1 it stores r2 = rmin2 (infinity for rmin2 >= FLT_MAX), and
// s = lo ? 1 : -1 either way.  The gradient needs the 16-byte record only:
// -s d vlj / d dsq = (4 rc u - 3 ac) s u^4 = 4 (rc u - 3/4 ac) s u^4, and the force shift
// adds gh / sqrt(dsq), gh = P dE/ddsq (P) being the force-shift factor -2 P dE/ddsq (P)
// (test-grid-energrad.py:124-147) divided by -2 (zero for the exact derivative and without
// the plateau); the record carries both terms divided by 4.  The side record and kh hold the
// energy constants (kh = 0 without the plateau, where it is not read).
struct PairType {
  float4 rec;
  float2 side;
  float kh;
};

__device__ __forceinline__ PairType pair_type(float rc, float ac, float iv, float em, float r2, bool plateau,
                                              bool exact, const PlateauConst& pc) {
  if (fabsf(iv) != 1.f) rc = __int_as_float(0x7fc00000);
  const float w = (iv - 1.f) * em;
  PairType t;
  t.rec = make_float4(rc, -0.75f * ac, iv < 0.f ? (r2 >= FLT_MAX ? INFINITY : r2) : FLT_MAX, 0.f);
  t.side = make_float2(ac, w);
  t.kh = 0.f;
  if (plateau) {
    const float vljP = fmaf(rc, pc.uP4, -(ac * pc.uP3));
    const bool loP = pc.P2 < r2;
    const float sP = loP ? 1.f : iv;
    const float EP = fmaf(sP, vljP, loP ? w : 0.f);
    t.side.y = w - EP;
    t.kh = -EP;
    if (!exact) t.rec.w = 0.125f * (sP * fmaf(rc, pc.k1, ac * pc.k2));
  }
  return t;
}

// ATTRACT 8-6 pair potential (test-grid-energrad.py:95-111) and its dsq-derivative, with
// the pair-type terms of the records (plateau subtraction and force shift when they carry
// them).  Returns E and h such that dE/dx'_lig = -8 h (x'_lig - x_rec):
// h = -s d vlj / d dsq / 4 (+ gh / (4 sqrt(dsq))); the factor -8 is applied by the caller once
// per list segment.  The all-zero record gives E = h = 0 for any finite dsq > 0 (padding and
// pairs beyond the plateau).  Without ENERGY only the 16-byte record is read and E is not
// formed; without the plateau kh is not read.
template <bool PLATEAU, bool EXACT, bool ENERGY = true>
__device__ __forceinline__ void pair_eval(float dsq, const float4 pa, const float2 ps, float kh, float& E,
                                          float& h) {
  float u, rs = 0.f;
  if (PLATEAU && !EXACT) {
    rs = __builtin_amdgcn_rsqf(dsq);
    u = rs * rs;
  } else {
    u = __builtin_amdgcn_rcpf(dsq);
  }
  const float u2 = u * u;
  const bool lo = dsq < pa.z;  // always for ivor = +1 (r2 = FLT_MAX)
  const float u4 = u2 * u2;
  const float su4 = lo ? u4 : -u4;  // s u^4
  h = fmaf(pa.x, u, pa.y) * su4;    // (rc u - 3/4 ac) s u^4
  if (PLATEAU && !EXACT) h = fmaf(pa.w, rs, h);
  if (ENERGY) {
    const float t1 = fmaf(pa.x, u, -ps.x);  // rc u - ac
    const float u3 = u2 * u;
    E = fmaf(t1, lo ? u3 : -u3, lo ? ps.y : (PLATEAU ? kh : 0.f));
  }
}

// pair_eval with run-time flags, returning the gradient factor gs = -8 h:
// dE/dx'_lig = gs (x'_lig - x_rec).
__device__ __forceinline__ void pair_eval_rt(bool plateau, bool exact, float dsq, const PairType& t, float& E,
                                             float& gs) {
  float h;
  if (plateau) {
    if (exact)
      pair_eval<true, true>(dsq, t.rec, t.side, t.kh, E, h);
    else
      pair_eval<true, false>(dsq, t.rec, t.side, t.kh, E, h);
  } else {
    pair_eval<false, false>(dsq, t.rec, t.side, t.kh, E, h);
  }
  gs = -8.f * h;
}

// Partial derivatives of nonbon(dsq) w.r.t. (rc, ac, emin) as independent inputs; u = 1/dsq.
__device__ __forceinline__ void pair_param_grad(float dsq, float u, const float4 pa, float d[3]) {
  const float u3 = (u * u) * u;
  const bool neg = pa.z != FLT_MAX;  // ivor = -1
  const bool lo = dsq < pa.z;
  const float s = lo ? 1.f : -1.f;
  d[0] = s * (u3 * u);
  d[1] = -s * u3;
  d[2] = lo && neg ? -2.f : 0.f;  // ivor - 1
}

// Fixed-point sums of the pair-table gradient.  A term v is rounded to a multiple of 2^-43:
// with a = v 2^12 (exact), B = rne(a) and r = a - B (exact, |r| <= 1/2), the term is
// B 2^-12 + L 2^-43 with L = rne(r 2^31), |L| <= 2^30.  A sum is two signed 64-bit words,
// w1 = sum of B and w0 = sum of L, worth w1 2^-12 + w0 2^-43.  Both take sign-extended 32-bit
// addends, so 2^31 adds cannot overflow them.  Integer adds give the same sum in any order,
// and LDS integer atomics are much faster than float ones (ds_add_u64 ~17x the rate of
// ds_add_f32 on gfx950).  Terms of 2^19 and above in magnitude (B beyond 32 bits), infinities
// and NaNs go to a float sum instead, the one part whose order is not fixed.
constexpr int kFixWords = 2;

// Adds v to the fixed-point sum whose words are t[0] (w0) and t[stride] (w1), or to the float
// *big for terms beyond the fixed-point range.
__device__ __forceinline__ void fixed_add(unsigned long long* t, unsigned stride, float* big, float v) {
  const float a = v * 0x1p12f;
  if (!(fabsf(a) < 0x1p31f)) {  // also NaN
    atomicAdd(big, v);
    return;
  }
  const float b = __builtin_rintf(a);
  const int lo = (int)__builtin_rintf((a - b) * 0x1p31f);
  const int hi = (int)b;
  atomicAdd(t, (unsigned long long)(long long)lo);
  atomicAdd(t + stride, (unsigned long long)(long long)hi);
}

// The value of a fixed-point sum, rounded to float: w0's carries go into w1, leaving w0 in
// [0, 2^31), then the two parts are added in double.
__device__ __forceinline__ float fixed_value(const unsigned long long* t, unsigned stride) {
  const long long w0 = (long long)t[0];
  const long long w1 = (long long)t[stride] + (w0 >> 31);
  return (float)fma((double)w1, 0x1p-12, (double)(w0 & 0x7fffffffll) * 0x1p-43);
}

// LDS pair-table rows (one per receptor type) have an odd stride of nlt | 1 records, so
// the records of one ligand type under different receptor types fall in different LDS
// banks (lanes working on one ligand bead read one column).  Slot nslots holds the
// all-zero record.
__host__ __device__ __forceinline__ int pair_stride(int nlt) { return nlt | 1; }
__host__ __device__ __forceinline__ int pair_slots(int nrt, int nlt) { return nrt * pair_stride(nlt); }
// Side-record slots: nslots + 1 rounded up to even, and kh slots: nslots + 1 rounded up to a
// multiple of 4, keeping what follows 16-byte aligned.
__host__ __device__ __forceinline__ int side_slots(int nslots) { return (nslots + 2) & ~1; }
__host__ __device__ __forceinline__ int kh_slots(int nslots) { return (nslots + 4) & ~3; }
// Floats of kAdjoint's block table: kFixWords 64-bit words per sum, then a float per sum for
// large terms, over [3][nslots] sums.
__host__ __device__ __forceinline__ int adjoint_table_floats(int nslots) { return (2 * kFixWords + 1) * 3 * nslots; }
// LDS bytes of the staged pair-type records.
__host__ __device__ __forceinline__ size_t pair_lds_bytes(int nslots) {
  return (size_t)(nslots + 1) * 16 + (size_t)side_slots(nslots) * 8 + (size_t)kh_slots(nslots) * 4;
}

struct WaveScratch {
  float4 xp[kWave];          // x', y', z', 16 x ligand ff type (the pair-record column in bytes, int bits)
  float4 xl[kWave];          // -8 (x, y, z) (local), list start - S * first item (int bits)
  signed char mark[kWave];   // bead whose first list segment is item base + i of the current window
  float4 gridB[kHalf];       // grid energy and gradient of the beads of the second pose of a shared bead phase
};

// Whether two consecutive poses of na and nb beads share one bead phase: both of at most kHalf
// beads.  pose_energrad takes any pair whose smaller pose has at most kHalf beads and whose
// beads fit the wave, but pairing large poses with small ones, although it saved 6% of the
// minimizer's VALU instructions, made its launches 3% slower (more time waiting on results).
__device__ __forceinline__ bool share_bead_phase(int na, int nb) { return na <= kHalf && nb <= kHalf; }

// Minimizer batch of one wavefront, in LDS: the pose slots whose energies the wave
// evaluates in order (two small ones at a time where they share a bead phase), and whose
// optimizer updates it then runs together (lane 8 slot + parameter).
constexpr int kBatch = 8;
struct BatchScratch {
  const float4* lig[kBatch];  // ligand beads of the slot's job
  int nb[kBatch];             // and their number
  float M[kBatch][12];    // rotation rows and translation for the next evaluation
  float acc[kBatch][16];  // reduced sums of the last evaluation (kAcc), then c1, c2 and the last energy
  float prm[kBatch][7];   // quaternion (x, y, z, w), translation
  float m1[kBatch][7];    // ADAM moments
  float m2[kBatch][7];
};

struct KArgs {
  GridView G;
  const float4* rec;
  int nrec;
  const float* rc;
  const float* ac;
  const float* ivor;
  const float* emin;
  const float* rmin2;
  int nrt, nlt;
  const AtjJob* jobs;
  const int* pose_job;
  const int* pose_order;
  int N;
  // energrad / adjoint
  const float* poses;
  float* energy;
  float* dEdM;
  const float* weights;
  float* partial;
  // minimize
  float* pose7;
  float* mom1;
  float* mom2;
  int* count;
  float* elast;
  int nsteps_launch;
  unsigned flags;
  int batch;          // poses per work-counter claim, 1..kBatch
  int tail_batch;     // poses per claim from tail_start on (<= batch)
  int tail_start;     // first pose of the tail (a multiple of batch; N: no tail)
  int static_blocks;  // blocks whose wavefronts take their first claim by index (the resident ones)
  int* counter;
  unsigned long long* stats;
};

// Per-wavefront launch statistics: the ATJ_NSTATS counters and, in a build with
// ATJ_PHASE_CLOCKS defined, the shader-clock cycles the wavefront spent in each phase of the
// evaluation (claims and set-up, bead phase, segment-size scans, pair phase, reductions,
// optimizer update), accumulated into stats[ATJ_NSTATS + phase].
enum Phase { kPhSetup = 0, kPhBead, kPhScan, kPhPair, kPhReduce, kPhUpdate, kNumPhases };
static_assert(kNumPhases == ATJ_NPHASES, "phase count of the header");
struct WaveStats {
  unsigned n[ATJ_NSTATS];
#ifdef ATJ_PHASE_CLOCKS
  unsigned long long clk[kNumPhases];
  unsigned long long last;
#endif
};

__device__ __forceinline__ void phase_end(WaveStats& st, int phase) {
#ifdef ATJ_PHASE_CLOCKS
  const unsigned long long now = __builtin_amdgcn_s_memtime();
  st.clk[phase] += now - st.last;
  st.last = now;
#else
  (void)st;
  (void)phase;
#endif
}

// Far-away receptor bead standing for list padding (0xFFFF entries): every pair with it
// lies beyond the plateau and uses the all-zero pair record.
__device__ __forceinline__ float4 sentinel_bead() { return make_float4(1e18f, 1e18f, 1e18f, 0.f); }

// Receptor bead with its type (pair-table row) limited to the table: a type beyond it is
// read as the last row, so no input can address memory outside the table.
__device__ __forceinline__ float4 clamp_rec_type(float4 b, int nrt) {
  b.w = __int_as_float((int)min((unsigned)__float_as_int(b.w), (unsigned)(nrt - 1)));
  return b;
}

// Pair loop over the flattened list segments (S entries each) of one chunk of beads.
// Bead lane j brings its segment count `segs` and first item `excl`; each window of 64
// items finds its beads by marking the segment starts that fall in the window and taking
// a prefix maximum (the bead of the window's first item is carried over).  Without ENERGY
// (minimizer steps whose energy is not kept) the pair energies are not formed.  A pair's
// record is addressed in bytes: the receptor bead's row (its .w in LDS; type x row bytes for
// beads read through L1) plus the ligand bead's column (ws->xp .w).
template <int S, int MODE, bool REC_LDS, bool PLATEAU, bool EXACT, bool ENERGY>
__device__ __forceinline__ void pair_items(const KArgs& a, const float4* __restrict__ recp,
                                           const float4* __restrict__ s_par, const float2* __restrict__ s_side,
                                           const float* __restrict__ s_kh, float* s_pgrad, WaveScratch* ws,
                                           int total, int segs, int excl, float wgt, float acc[kAcc],
                                           unsigned& npair, int lane) {
  const GridView& G = a.G;
  const float P2 = G.p2;
  const float uP2 = 1.f / P2;  // (adjoint) the plateau terms' 1/dsq
  const unsigned nrec = (unsigned)a.nrec;
  const unsigned row_bytes = 16u * (unsigned)pair_stride(a.nlt);
  const unsigned nslots = (unsigned)pair_slots(a.nrt, a.nlt);
  const unsigned zero_rec = 16u * nslots;  // byte offset of the all-zero record
  int carry = 0;
  static_assert(S == 4 || S == 8, "segments of 4 or 8 list entries");
  // bead of this lane's item in the window at `base`: the lanes whose first segment falls in
  // the window mark it, and a prefix maximum hands every lane its bead (the bead of the
  // window's first item is carried over from the previous window)
  auto window_bead = [&](int base) {
    ws->mark[lane] = -1;
    wave_sync();
    const int d = excl - base;
    if (segs > 0 && d >= 0 && d < kWave) ws->mark[d] = lane;
    wave_sync();
    const int jb = max(wave_max_scan(ws->mark[lane]), carry);
    carry = __builtin_amdgcn_readlane(jb, kWave - 1);
    return jb;
  };
  // the list entries of this lane's segment in the window at `base`; inactive lanes read the
  // all-padding segment at the head of the list array
  auto load_words = [&](int base, int jb, unsigned w[S / 2]) {
    const int item = base + lane;
    const int off = item < total ? __float_as_int(ws->xl[jb].w) + item * S : 0;
    if (S == 8) {
      const uint4 ent = *reinterpret_cast<const uint4*>(G.nblist + off);
      w[0] = ent.x;
      w[1] = ent.y;
      w[S / 2 > 2 ? 2 : 0] = ent.z;
      w[S / 2 > 3 ? 3 : 0] = ent.w;
    } else {
      const uint2 ent = *reinterpret_cast<const uint2*>(G.nblist + off);
      w[0] = ent.x;
      w[1] = ent.y;
    }
  };
  // the next window's list entries are loaded while this window's pairs are evaluated
  unsigned next[S / 2];
  int jb_next = 0;
  if (total > 0) {
    jb_next = window_bead(0);
    load_words(0, jb_next, next);
  }
  for (int base = 0; base < total; base += kWave) {
    const int jb = jb_next;
    unsigned words[S / 2];
#pragma unroll
    for (int i = 0; i < S / 2; ++i) words[i] = next[i];
    if (base + kWave < total) {
      jb_next = window_bead(base + kWave);
      load_words(base + kWave, jb_next, next);
    }
    const float4 bx = ws->xp[jb];
    const float4 bl = ws->xl[jb];
    const unsigned col = (unsigned)__float_as_int(bx.w);  // 16 x ligand ff type
    const unsigned zrow = zero_rec - col;                 // the "row" whose column is the zero record
    // -0 is the identity of float addition, so the first pair needs no add
    float e = -0.f, gx = -0.f, gy = -0.f, gz = -0.f;
#pragma unroll
    for (int h = 0; h < S / 4; ++h) {
      float4 rb[4];
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const unsigned wd = words[2 * h + (cc >> 1)];
        const unsigned r = (cc & 1) ? (wd >> 16) : (wd & 0xffffu);
        if (REC_LDS) {
          rb[cc] = recp[min(r, nrec)];
        } else {
          const float4 v = clamp_rec_type(recp[min(r, nrec - 1u)], a.nrt);
          rb[cc] = r < nrec ? v : sentinel_bead();
        }
      }
      float dd[4][3], dsq[4];
      float4 pa[4];
      float2 ps[4];
      float kh[4];
      bool ok[4];
      unsigned pidx[4];
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        dd[cc][0] = bx.x - rb[cc].x;
        dd[cc][1] = bx.y - rb[cc].y;
        dd[cc][2] = bx.z - rb[cc].z;
        dsq[cc] = fmaf(dd[cc][2], dd[cc][2], fmaf(dd[cc][1], dd[cc][1], dd[cc][0] * dd[cc][0]));
        ok[cc] = dsq[cc] < P2;
        // the receptor's row is read whatever ok is: an empty asm keeps the compiler from
        // turning the select below into a branch around that read
        unsigned row = (unsigned)__float_as_int(rb[cc].w);
        asm("" : "+v"(row));
        const unsigned pb = (ok[cc] ? (REC_LDS ? row : __umul24(row, row_bytes)) : zrow) + col;
        pidx[cc] = pb;
        pa[cc] = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s_par) + pb);
        ps[cc] = ENERGY && !is_adjoint(MODE) ? s_side[pb >> 4] : make_float2(0.f, 0.f);
        kh[cc] = PLATEAU && ENERGY && !is_adjoint(MODE) ? s_kh[pb >> 4] : 0.f;
      }
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        npair += (unsigned)__builtin_popcountll(__builtin_amdgcn_ballot_w64(ok[cc]));
        if (is_adjoint(MODE)) {
          // only the parameter derivatives.  kAdjoint: into the block's fixed-point table,
          // whose sums do not depend on the order of the adds.  kAdjointWave: into this
          // wavefront's own float table, which no other wavefront writes and whose adds come
          // in a fixed order.  Either way reruns give the same sums.
          if (ok[cc]) {
            float dp[3];
            pair_param_grad(dsq[cc], __builtin_amdgcn_rcpf(dsq[cc]), pa[cc], dp);
            if (PLATEAU) {
              float dq[3];
              pair_param_grad(P2, uP2, pa[cc], dq);
              dp[0] -= dq[0];
              dp[1] -= dq[1];
              dp[2] -= dq[2];
            }
            const unsigned pg = pidx[cc] >> 4;
            if (MODE == kAdjoint) {
              // words [kFixWords][3][nslots], then the float table of large terms [3][nslots]
              unsigned long long* t = reinterpret_cast<unsigned long long*>(s_pgrad);
              float* big = s_pgrad + 2 * kFixWords * 3 * nslots;
              const unsigned stride = 3 * nslots;
              fixed_add(t + pg, stride, big + pg, wgt * dp[0]);
              fixed_add(t + nslots + pg, stride, big + nslots + pg, wgt * dp[1]);
              // the emin derivative is zero except for repulsive pairs inside rmin
              if (dp[2] != 0.f) fixed_add(t + 2 * nslots + pg, stride, big + 2 * nslots + pg, wgt * dp[2]);
            } else {
              atomicAdd(&s_pgrad[pg], wgt * dp[0]);
              atomicAdd(&s_pgrad[nslots + pg], wgt * dp[1]);
              if (dp[2] != 0.f) atomicAdd(&s_pgrad[2 * nslots + pg], wgt * dp[2]);
            }
          }
        } else {
          float E = 0.f, h;
          pair_eval<PLATEAU, EXACT, ENERGY>(dsq[cc], pa[cc], ps[cc], kh[cc], E, h);
          if (ENERGY) e += E;
          gx = fmaf(h, dd[cc][0], gx);
          gy = fmaf(h, dd[cc][1], gy);
          gz = fmaf(h, dd[cc][2], gz);
        }
      }
    }
    if (is_adjoint(MODE)) continue;
    // the gradient is -8 (gx, gy, gz); bl holds -8 x_local, and scaling by -8 is exact, so
    // fma(-8 x, g', acc) and fma(-8, g', acc) equal the products with the scaled gradient
    if (ENERGY) acc[0] += e;
    const float g[3] = {gx, gy, gz};
    const float xb[3] = {bl.x, bl.y, bl.z};
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int l = 0; l < 3; ++l) acc[1 + 3 * k + l] = fmaf(xb[k], g[l], acc[1 + 3 * k + l]);
#pragma unroll
    for (int l = 0; l < 3; ++l) acc[10 + l] = fmaf(-8.f, g[l], acc[10 + l]);
  }
}

// Adds one bead's grid energy E and gradient g (at local coordinates xl) to the lane
// accumulators: energy, xl (x) g, g.
__device__ __forceinline__ void bead_acc(float acc[kAcc], float E, const float g[3], const float xl[3]) {
  acc[0] += E;
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int l = 0; l < 3; ++l) acc[1 + 3 * k + l] = fmaf(xl[k], g[l], acc[1 + 3 * k + l]);
#pragma unroll
  for (int l = 0; l < 3; ++l) acc[10 + l] += g[l];
}

// Estimated cost of t 8-entry segments: full windows, and a last window at most half full
// taken as one window of 4-entry segments (the re-cut of the pair phase).
__device__ __forceinline__ int seg8_cost(int t) {
  const int full = t / kWave, rest = t - full * kWave;
  return full * (8 * kPairCost + kItemCost) +
         (rest == 0 ? 0 : (full > 0 && rest <= kWave / 2 ? 4 : 8) * kPairCost + kItemCost);
}

// One pose of an evaluation: ligand beads, pose matrix in LDS (rows 0-2 of the rotation,
// then the translation: 12 floats) and adjoint weight.
struct PoseIn {
  const float4* lig;
  const float* M;
  int nb;
  float wgt;
};

// Energy and dE/dM of pose A, or of poses A and B.  Returns the wave-reduced sums in the
// transposed layout of wave_sum_all: lane l holds accumulator acc_index(l) (0: energy,
// 1 + 3k + l: dE/dM[k][l], 10 + l: dE/dM[3][l]) of A in uA and of B in uB.
// A pose alone puts bead c0 + j on lane j, in chunks of 64 beads.  With a second pose B
// (B.nb <= kHalf and A.nb + B.nb <= 64; the callers pair by share_bead_phase) both run in one bead
// phase, A's bead j on lane j and B's on lane sp + j (sp = 64 - B.nb in the minimizer, 32 in the
// single-shot modes); B's grid sums wait in LDS while A's pairs are evaluated, then move to lanes
// 0 .. B.nb - 1.  Either way a pose's sums are
// formed on the same lanes in the same order (grid terms of bead j on lane j, then its pairs
// in windows of its own, one reduction), so its result is the same alone or beside another.  With energy
// false (minimizer steps whose energy is not kept) the pair energies are left out of the
// energy sum; every gradient sum is the same.
template <int MODE, bool REC_LDS, bool PLATEAU, bool EXACT, int LAYOUT>
__device__ void pose_energrad(const KArgs& a, const float4* __restrict__ s_rec, const float4* __restrict__ s_par,
                              const float2* __restrict__ s_side, const float* __restrict__ s_kh, float* s_pgrad,
                              WaveScratch* ws, const PoseIn& A, const PoseIn& B, WaveStats& st, int lane, float& uA,
                              float& uB, bool energy = true) {
  const float4* __restrict__ recp = REC_LDS ? s_rec : a.rec;
  const bool do_grid = (a.flags & ATJ_TERM_GRID) != 0;
  const bool do_nb = (a.flags & ATJ_TERM_NB) != 0;
  const GridView& G = a.G;
  // B's first lane sp: the minimizer puts B on the last B.nb lanes, the single-shot modes on
  // lanes 32 + j (same sums either way; each form measured faster in its mode, whose kernels
  // use every one of their 128 registers)
  const bool halves = A.nb <= kHalf;
  const int nposes = B.nb > 0 ? 2 : 1;  // the callers pair poses of at most kHalf beads only
  const bool min_form = MODE == kMinimize;
  const int sp = min_form ? (nposes == 2 ? kWave - B.nb : kWave) : (halves ? kHalf : kWave);
  const int h = min_form ? (lane >= sp ? 1 : 0) : (halves ? lane >> 5 : 0);  // pose of this lane
  float acc[kAcc];
#pragma unroll
  for (int i = 0; i < kAcc; ++i) acc[i] = 0.f;
  uA = uB = 0.f;
  phase_end(st, kPhSetup);

  // a pose without beads has energy and gradient 0; paired with B it still runs B's chunk
  const int nloop = max(A.nb, nposes == 2 ? B.nb : 0);
  for (int c0 = 0; c0 < nloop; c0 += kWave) {
    const int j = min_form ? (h ? lane - sp : c0 + lane) : (halves ? lane & (kHalf - 1) : c0 + lane);
    const int nbh = h ? B.nb : A.nb;
    int cnt = 0, start = 0;
    float xl[3] = {0.f, 0.f, 0.f}, xp[3] = {0.f, 0.f, 0.f};
    float Eg = 0.f, g[3] = {0.f, 0.f, 0.f}, raw[4];
    int lt = 0;
    if (j < nbh) {
      const float4 b = (h ? B.lig : A.lig)[j];
      xl[0] = b.x;
      xl[1] = b.y;
      xl[2] = b.z;
      // types beyond the potential grids or the pair table are read as the last one, so no
      // input can address memory outside them (the Python layer rejects such ligands)
      const unsigned tb = (unsigned)__float_as_int(b.w);
      const int gt = min((int)(tb & 0xffffu), G.ntypes - 1);
      lt = min((int)(tb >> 16), a.nlt - 1);
      const float* mp = h ? B.M : A.M;
      float M[16];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int l = 0; l < 3; ++l) M[4 * k + l] = mp[3 * k + l];
        M[4 * k + 3] = 1.f;
      }
      transform(M, xl[0], xl[1], xl[2], xp);
      int voxel;
      unsigned in_any;
      if (MODE == kMinimize && !energy)  // the grid energy channel is not interpolated
        bead_grid<LAYOUT, false>(G, gt, xp, do_grid, do_nb, EXACT, Eg, g, voxel, cnt, start, in_any, raw);
      else
        bead_grid<LAYOUT>(G, gt, xp, do_grid, do_nb, EXACT, Eg, g, voxel, cnt, start, in_any, raw);
      if (!is_adjoint(MODE)) bead_acc(acc, Eg, g, xl);
      st.n[0] += in_any;
      if (voxel >= 0) {
        st.n[1] += 1;
        st.n[2] += (unsigned)cnt;
      }
    }
    phase_end(st, kPhBead);
    if (nposes == 2) {  // B's bead terms wait in LDS; its lanes start A's pair sums from zero
      if (h) ws->gridB[lane - sp] = make_float4(Eg, g[0], g[1], g[2]);
#pragma unroll
      for (int i = 0; i < kAcc; ++i) acc[i] = h ? 0.f : acc[i];
    }
    // segment size per pose and chunk: 8 list entries (fewer windows) or 4 (fewer padding slots)
    int t8A = 0, t4A = 0, t8B = 0, t4B = 0, segs = 0, excl = 0;
    bool u8A = true, u8B = true;
    if (do_nb) {
      const int seg8 = (cnt + 7) >> 3, seg4 = (cnt + 3) >> 2;
      const int incl8 = wave_incl_scan(seg8);
      const int incl4 = wave_incl_scan(seg4);
      t8A = __builtin_amdgcn_readlane(incl8, sp - 1);
      t4A = __builtin_amdgcn_readlane(incl4, sp - 1);
      const bool two = min_form ? nposes == 2 : halves;
      t8B = two ? __builtin_amdgcn_readlane(incl8, kWave - 1) - t8A : 0;
      t4B = two ? __builtin_amdgcn_readlane(incl4, kWave - 1) - t4A : 0;
      u8A = seg8_cost(t8A) <= ((t4A + kWave - 1) / kWave) * (4 * kPairCost + kItemCost);
      u8B = seg8_cost(t8B) <= ((t4B + kWave - 1) / kWave) * (4 * kPairCost + kItemCost);
      const bool use8 = h ? u8B : u8A;
      segs = use8 ? seg8 : seg4;
      excl = (use8 ? incl8 : incl4) - segs - (h ? (use8 ? t8A : t4A) : 0);
    }
    const int lstart = start;  // the bead's list start
    start -= excl * (h ? (u8B ? 8 : 4) : (u8A ? 8 : 4));
    if (nposes == 2 || t8A + t8B > 0) {
      ws->xp[lane] = make_float4(xp[0], xp[1], xp[2], __int_as_float(16 * lt));
      ws->xl[lane] = make_float4(-8.f * xl[0], -8.f * xl[1], -8.f * xl[2], __int_as_float(start));
    }
    wave_sync();
    phase_end(st, kPhScan);
    unsigned npair = 0;  // wave-wide count, credited to lane 0
    for (int q = 0; q < nposes; ++q) {
      if (q == 1 && !is_adjoint(MODE)) {  // A is complete; B's bead terms to lanes 0 .. B.nb - 1, summed as in B's own bead phase
        uA = wave_sum_all(acc, lane);
#pragma unroll
        for (int i = 0; i < kAcc; ++i) acc[i] = 0.f;
        if (min_form ? lane < B.nb : !h) {
          const float4 e = ws->gridB[lane];
          const float4 x = ws->xl[sp + lane];  // -8 x_local
          const float gb[3] = {e.y, e.z, e.w}, xb[3] = {-0.125f * x.x, -0.125f * x.y, -0.125f * x.z};
          bead_acc(acc, e.x, gb, xb);
        }
      }
      const bool u8 = q ? u8B : u8A;
      const int tot = q ? (u8 ? t8B : t4B) : (u8 ? t8A : t4A);
      const int sq = h == q ? segs : 0;  // the lanes of pose q mark its windows
      const float wgt = q ? B.wgt : A.wgt;
      // 8-entry segments: full windows, and a last window at most half full re-cut when the
      // entries it would hold (the rest of each bead's list) fit one window of 4-entry segments
      bool tail4 = false;
      int seg4t = 0, excl4t = 0, tot4t = 0, done = 0;
      const int nfull = tot & ~(kWave - 1);
      if (u8 && nfull > 0 && tot - nfull <= kWave / 2) {
        done = 8 * min(max(nfull - excl, 0), sq);  // entries in full windows (sq = 0: other pose)
        seg4t = h == q ? (max(cnt - done, 0) + 3) >> 2 : 0;
        const int incl = wave_incl_scan(seg4t);
        tot4t = __builtin_amdgcn_readlane(incl, kWave - 1);
        excl4t = incl - seg4t;
        tail4 = tot4t <= kWave;
      }
      const int n8 = u8 ? (tail4 ? nfull : tot) : 0;
      const int n4 = u8 ? (tail4 ? tot4t : 0) : tot;
      const bool no_energy = MODE == kMinimize && !energy;
      if (n8 > 0) {
        if (no_energy)
          pair_items<8, MODE, REC_LDS, PLATEAU, EXACT, false>(a, recp, s_par, s_side, s_kh, s_pgrad, ws, n8, sq, excl, wgt,
                                                              acc, npair, lane);
        else
          pair_items<8, MODE, REC_LDS, PLATEAU, EXACT, true>(a, recp, s_par, s_side, s_kh, s_pgrad, ws, n8, sq, excl, wgt,
                                                             acc, npair, lane);
      }
      if (tail4) {
        if (h == q) ws->xl[lane].w = __int_as_float(lstart + done - 4 * excl4t);
        wave_sync();
      }
      if (n4 > 0) {
        const int sq4 = tail4 ? seg4t : sq, ex4 = tail4 ? excl4t : excl;
        if (no_energy)
          pair_items<4, MODE, REC_LDS, PLATEAU, EXACT, false>(a, recp, s_par, s_side, s_kh, s_pgrad, ws, n4, sq4, ex4, wgt,
                                                              acc, npair, lane);
        else
          pair_items<4, MODE, REC_LDS, PLATEAU, EXACT, true>(a, recp, s_par, s_side, s_kh, s_pgrad, ws, n4, sq4, ex4, wgt,
                                                             acc, npair, lane);
      }
    }
    phase_end(st, kPhPair);
    st.n[3] += lane == 0 ? npair : 0u;
    wave_sync();
  }
  if (is_adjoint(MODE)) return;  // the parameter derivatives went to the wavefront's table
  const float u = wave_sum_all(acc, lane);
  if (nposes == 2)
    uB = u;
  else
    uA = u;
  phase_end(st, kPhReduce);
}

// Energy and dE/dM of pose p from the transposed sums: one lane per output word.
__device__ __forceinline__ void write_energrad(const KArgs& a, int p, float u, int lane) {
  if (lane & 3) return;
  const int i = acc_index(lane);
  if (i == 0) a.energy[p] = u;
  if (!a.dEdM) return;
  float* o = a.dEdM + (size_t)p * 16;
  if (i == 0) o[15] = 0.f;
  else if (i <= 9) o[4 * ((i - 1) / 3) + (i - 1) % 3] = u;
  else if (i <= 12) o[12 + (i - 10)] = u;
  else o[4 * (i - 13) + 3] = 0.f;  // i = 13..15: the zero column of rows 0..2
}

// Lane whose wave_sum_all result is accumulator i (the inverse of acc_index).
__host__ __device__ constexpr int acc_lane(int i) {
  return ((i & 1) << 5) | (((i >> 1) & 1) << 4) | (((i >> 2) & 1) << 3) | (((i >> 3) & 1) << 2);
}

// Energy and dE/d(q, t) of quaternion pose p (ATJ_GRAD_QT): the sums are gathered from
// their lanes, the chain rule through as_matrix applied as in the optimizer update, and
// lanes 0-15 write the 16 output words.
__device__ void write_energrad_qt(const KArgs& a, int p, float u, int lane) {
  float A[9], t[3];
#pragma unroll
  for (int i = 0; i < 9; ++i) A[i] = lane_value(u, acc_lane(1 + i));
#pragma unroll
  for (int i = 0; i < 3; ++i) t[i] = lane_value(u, acc_lane(10 + i));
  const float E = lane_value(u, acc_lane(0));
  const float* pp = a.poses + (size_t)p * 7;
  const float q[4] = {uni(pp[0]), uni(pp[1]), uni(pp[2]), uni(pp[3])};
  float g[4];
  quat_grad(q, A, g);
  if (lane == 0) a.energy[p] = E;
  if (!a.dEdM || lane >= 16) return;
  float v = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) v = lane == i ? g[i] : v;
#pragma unroll
  for (int i = 0; i < 3; ++i) v = lane == 4 + i ? t[i] : v;
  a.dEdM[(size_t)p * 16 + lane] = v;
}

// One optimizer step of a pose from its evaluation (acc: energy, dE/dM) with the same
// operations in the same order in the fused minimizer and in optimizer_step_kernel:
// quaternion chain rule through as_matrix, optional gravity restraint w |t|^2
// (test5.py:418-420; energy += (|t| |t|) w, translation gradient += 2 w t), then the
// optax update (adam, test4.py:447-453, or adabelief, test5.py:451-457: the belief is
// the moment of g - m_prev, plus eps_root), bias-corrected by c1 = 1 - b1^t and
// c2 = 1 - b2^t, -lr, and Rotation.from_quat(as_quat()) renormalisation.  Returns the
// pose's energy including the restraint.
__device__ __forceinline__ float optimizer_update(const AtjOptimizer& o, float prm[7], float m1[7], float m2[7],
                                                  const float acc[kAcc], float c1, float c2) {
  const float q[4] = {prm[0], prm[1], prm[2], prm[3]};
  const float A[9] = {acc[1], acc[2], acc[3], acc[4], acc[5], acc[6], acc[7], acc[8], acc[9]};
  float gr[7];
  quat_grad(q, A, gr);
  gr[4] = acc[10];
  gr[5] = acc[11];
  gr[6] = acc[12];
  float E = acc[0];
  if (o.gravity != 0.f) {
    const float n = sqrtf(fmaf(prm[6], prm[6], fmaf(prm[5], prm[5], prm[4] * prm[4])));
    E = E + (n * n) * o.gravity;
#pragma unroll
    for (int i = 4; i < 7; ++i) gr[i] = gr[i] + (2.f * o.gravity) * prm[i];
  }
  const float b1 = o.beta1, b2 = o.beta2;
  const bool belief = o.kind == ATJ_OPT_ADABELIEF;
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    const float pe = belief ? gr[i] - m1[i] : gr[i];
    m1[i] = (1.f - b1) * gr[i] + b1 * m1[i];
    m2[i] = (1.f - b2) * (pe * pe) + b2 * m2[i];
    if (belief) m2[i] = m2[i] + o.eps_root;
    const float upd = (m1[i] / c1) / (sqrtf(m2[i] / c2) + o.eps);
    prm[i] = prm[i] + (-(i < 4 ? o.lr_rot : o.lr_trans)) * upd;
  }
  const float nrm = sqrtf(fmaf(prm[3], prm[3], fmaf(prm[2], prm[2], fmaf(prm[1], prm[1], prm[0] * prm[0]))));
#pragma unroll
  for (int i = 0; i < 4; ++i) prm[i] = prm[i] / nrm;
  return E;
}

// Rotation rows and translation (12 floats) of a (quaternion, translation) pose.
// Parameter i of one pose's optimizer update: the operations optimizer_update performs for
// it, before the quaternion renormalisation (quat_renorm).  prm and acc are the pose's
// parameters and reduced sums; returns the new parameter and writes its moments and the
// energy including the restraint.
__device__ __forceinline__ float optimizer_update_param(const AtjOptimizer& o, const float* prm, const float* acc,
                                                        int i, float m1_old, float m2_old, float c1, float c2,
                                                        float& m1, float& m2, float& E) {
  const float q[4] = {prm[0], prm[1], prm[2], prm[3]};
  const float A[9] = {acc[1], acc[2], acc[3], acc[4], acc[5], acc[6], acc[7], acc[8], acc[9]};
  float gq[4];
  quat_grad(q, A, gq);
  float g = i == 0 ? gq[0] : i == 1 ? gq[1] : i == 2 ? gq[2] : i == 3 ? gq[3] : acc[10 + (i > 3 ? i - 4 : 0)];
  const float x = prm[i];
  E = acc[0];
  if (o.gravity != 0.f) {
    const float n = sqrtf(fmaf(prm[6], prm[6], fmaf(prm[5], prm[5], prm[4] * prm[4])));
    E = E + (n * n) * o.gravity;
    if (i >= 4) g = g + (2.f * o.gravity) * x;
  }
  const float b1 = o.beta1, b2 = o.beta2;
  const bool belief = o.kind == ATJ_OPT_ADABELIEF;
  const float pe = belief ? g - m1_old : g;
  m1 = (1.f - b1) * g + b1 * m1_old;
  m2 = (1.f - b2) * (pe * pe) + b2 * m2_old;
  if (belief) m2 = m2 + o.eps_root;
  const float upd = (m1 / c1) / (sqrtf(m2 / c2) + o.eps);
  return x + (-(i < 4 ? o.lr_rot : o.lr_trans)) * upd;
}

// Component i (< 4) of the renormalised quaternion prm[0..3], as optimizer_update computes it.
__device__ __forceinline__ float quat_renorm(const float* prm, int i) {
  const float nrm = sqrtf(fmaf(prm[3], prm[3], fmaf(prm[2], prm[2], fmaf(prm[1], prm[1], prm[0] * prm[0]))));
  return prm[i] / nrm;
}

__device__ __forceinline__ void pose12(const float prm[7], float M[12]) {
  float R[9];
  quat_to_mat(prm, R);
#pragma unroll
  for (int i = 0; i < 9; ++i) M[i] = R[i];
  M[9] = prm[4];
  M[10] = prm[5];
  M[11] = prm[6];
}

// Elements of pose12(prm) written by lane i (< 7) of a pose's update: M[i], and M[7 + i]
// for i < 5 (rotation rows, then the translation).
__device__ __forceinline__ void pose12_elements(const float* prm, float* M, int i) {
  const float q[4] = {prm[0], prm[1], prm[2], prm[3]};
  float R[9];
  quat_to_mat(q, R);
  float lo = R[0];
#pragma unroll
  for (int k = 1; k < 7; ++k) lo = i == k ? R[k] : lo;
  M[i] = lo;
  if (i < 5) M[7 + i] = i == 0 ? R[7] : i == 1 ? R[8] : prm[2 + i];
}

__device__ __forceinline__ void load_pose_matrix(const float* p, bool qt, float M[16]) {
  if (qt) {
    float q[4] = {p[0], p[1], p[2], p[3]};
    float R[9];
    quat_to_mat(q, R);
    M[0] = R[0]; M[1] = R[1]; M[2] = R[2]; M[3] = 1.f;
    M[4] = R[3]; M[5] = R[4]; M[6] = R[5]; M[7] = 1.f;
    M[8] = R[6]; M[9] = R[7]; M[10] = R[8]; M[11] = 1.f;
    M[12] = p[4]; M[13] = p[5]; M[14] = p[6]; M[15] = 1.f;
  } else {
#pragma unroll
    for (int i = 0; i < 16; ++i) M[i] = p[i];
  }
}


// Stage the receptor beads (+ the padding sentinel) and the pair table (+ the
// all-zero record) into LDS.  A staged bead's .w is the byte offset of its pair-record row.
__device__ void stage_lds(const KArgs& a, float4* s_rec, float4* s_par, float2* s_side, float* s_kh, bool plateau,
                          bool exact, float* s_pgrad, int npgrad) {
  const int tid = threadIdx.x;
  const int stride = pair_stride(a.nlt), nslots = pair_slots(a.nrt, a.nlt);
  if (s_rec) {
    for (int i = tid; i < a.nrec; i += blockDim.x) {
      float4 b = clamp_rec_type(a.rec[i], a.nrt);
      b.w = __int_as_float(16 * stride * __float_as_int(b.w));
      s_rec[i] = b;
    }
    if (tid == 0) s_rec[a.nrec] = sentinel_bead();
  }
  const PlateauConst pc = plateau_const(a.G.p2);
  for (int i = tid; i <= nslots; i += blockDim.x) {
    const int rt = i / stride, lt = i - rt * stride;
    PairType t;
    t.rec = make_float4(0.f, 0.f, 0.f, 0.f);
    t.side = make_float2(0.f, 0.f);
    t.kh = 0.f;
    if (i < nslots && lt < a.nlt) {
      const int k = rt * a.nlt + lt;
      t = pair_type(a.rc[k], a.ac[k], a.ivor[k], a.emin[k], a.rmin2[k], plateau, exact, pc);
    }
    s_par[i] = t.rec;
    s_side[i] = t.side;
    s_kh[i] = t.kh;
  }
  if (s_pgrad)
    for (int i = tid; i < npgrad; i += blockDim.x) s_pgrad[i] = 0.f;
}

// Work claims of a persistent wavefront.  Claim k covers a.batch poses from k a.batch on, up
// to a.tail_start, and a.tail_batch poses each after it.  In the blocks the device holds at
// once (a.static_blocks) a wavefront's first claim is fixed by its index; every other claim
// comes from the work counter and follows those first claims.  A launch whose first claims
// cover every pose thus takes no atomics (one per wavefront on a single address made small
// launches atomic-bound), and a larger one takes none for its first round.
struct Claim {
  long long k0;  // first pose of the claim (>= N: no work left)
  int n;         // poses in it
};
__device__ __forceinline__ long long claim_count(const KArgs& a) {
  const long long kb = a.tail_start / a.batch;
  return kb + (a.N - a.tail_start + a.tail_batch - 1) / a.tail_batch;
}
__device__ __forceinline__ Claim claim_of(const KArgs& a, long long k) {
  const long long kb = a.tail_start / a.batch;
  Claim c;
  if (k < kb) {
    c.k0 = k * a.batch;
    c.n = a.batch;
  } else {
    c.k0 = a.tail_start + (k - kb) * a.tail_batch;
    c.n = a.tail_batch;
  }
  if (c.k0 >= a.N) c.k0 = a.N;
  c.n = (int)min((long long)c.n, a.N - c.k0);
  return c;
}
__device__ __forceinline__ Claim next_claim(const KArgs& a, int lane) {
  const long long fr = (long long)a.static_blocks * (blockDim.x / kWave);
  if (fr >= claim_count(a)) return claim_of(a, claim_count(a));
  int c = 0;
  if (lane == 0) c = atomicAdd(a.counter, 1);
  return claim_of(a, fr + __builtin_amdgcn_readfirstlane(c));
}
__device__ __forceinline__ Claim first_claim(const KArgs& a, int lane) {
  if ((int)blockIdx.x >= a.static_blocks) return next_claim(a, lane);
  const int wave = uni((int)(threadIdx.x / kWave));  // uniform across the wavefront
  return claim_of(a, (long long)blockIdx.x * (blockDim.x / kWave) + wave);
}

// Minimizer loop of one wavefront.  The wave takes kBatch consecutive poses of the
// processing order at a time; per optimizer step it evaluates the slots still stepping
// in order (a slot alone, or together with the next stepping slot when the two share a bead
// phase), then lane 8 b + i updates parameter i of slot b.  Slots may belong to different
// jobs and have different step counts.
template <bool REC_LDS, bool PLATEAU, bool EXACT, int LAYOUT>
__device__ void minimize_wave(const KArgs& a, const float4* __restrict__ s_rec, const float4* __restrict__ s_par,
                              const float2* __restrict__ s_side, const float* __restrict__ s_kh, WaveScratch* ws,
                              BatchScratch* bs, WaveStats& st, int lane) {
  // Slot lanes of the optimizer update: lane 8 b + i updates parameter i (< 7) of slot b.
  const int sb = lane >> 3, pi = lane & 7;
  for (Claim claim = first_claim(a, lane); claim.k0 < a.N; claim = next_claim(a, lane)) {
    const int k0 = (int)claim.k0;
    const int nbatch = claim.n;
    const bool own = lane < nbatch;
    int p = 0, jid = 0, cnt = 0, nst = 0;
    if (own) {
      const int k = k0 + lane;
      p = a.pose_order ? a.pose_order[k] : k;
      jid = a.pose_job ? a.pose_job[p] : 0;
      const AtjJob job = a.jobs[jid];
      bs->lig[lane] = reinterpret_cast<const float4*>(job.ligand.beads);
      bs->nb[lane] = job.ligand.nbeads;
      cnt = a.count[p];
      nst = max(0, min(a.nsteps_launch, job.nsteps - cnt));
      // bias corrections 1 - b^t by the recurrence c_t = b c_{t-1} + (1 - b), c_0 = 0, and the
      // last energy, kept in the unused sums 13-15 of the slot's row
      float c1 = 0.f, c2 = 0.f;
      for (int i = 0; i < cnt; ++i) {
        c1 = fmaf(job.opt.beta1, c1, 1.f - job.opt.beta1);
        c2 = fmaf(job.opt.beta2, c2, 1.f - job.opt.beta2);
      }
      bs->acc[lane][kAcc] = c1;
      bs->acc[lane][kAcc + 1] = c2;
      bs->acc[lane][kAcc + 2] = a.elast[p];
      float prm[7], M[12];
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        prm[i] = a.pose7[7 * p + i];
        bs->prm[lane][i] = prm[i];
        bs->m1[lane][i] = a.mom1[7 * p + i];
        bs->m2[lane][i] = a.mom2[7 * p + i];
      }
      pose12(prm, M);
#pragma unroll
      for (int i = 0; i < 12; ++i) bs->M[lane][i] = M[i];
    }
    int nst_max = 0;
    for (int b = 0; b < nbatch; ++b) nst_max = max(nst_max, __builtin_amdgcn_readlane(nst, b));
    const int src = sb < nbatch ? sb : 0;  // the own lane of this lane's slot
    const bool slot = sb < nbatch && pi < 7;
    wave_sync();
    phase_end(st, kPhSetup);
    for (int s = 0; s < nst_max; ++s) {
      // active slots in order; a slot and the next stepping one share the bead phase when they
      // fit it (the smaller as pose B)
      for (int b = 0; b < nbatch; ++b) {
        if (s >= __builtin_amdgcn_readlane(nst, b)) continue;
        const int nbb = uni(bs->nb[b]);
        int c = b + 1;
        while (c < nbatch && s >= __builtin_amdgcn_readlane(nst, c)) ++c;
        const int partner = c < nbatch && share_bead_phase(nbb, uni(bs->nb[c])) ? c : -1;
        // slots of poses A and B
        int sa = b, sbb = -1;
        if (partner >= 0) {
          const bool swap = uni(bs->nb[partner]) > nbb;
          sa = swap ? partner : b;
          sbb = swap ? b : partner;
        }
        PoseIn P[2];
        P[0].lig = bs->lig[sa];
        P[0].nb = uni(bs->nb[sa]);
        P[0].M = bs->M[sa];
        P[0].wgt = 0.f;
        P[1] = P[0];
        P[1].nb = 0;
        if (sbb >= 0) {
          P[1].lig = bs->lig[sbb];
          P[1].nb = uni(bs->nb[sbb]);
          P[1].M = bs->M[sbb];
        }
        // the energy is kept only from a slot's last step of the launch
        bool energy = s + 1 >= __builtin_amdgcn_readlane(nst, b);
        if (partner >= 0) energy = energy || s + 1 >= __builtin_amdgcn_readlane(nst, partner);
        float uA, uB;
        pose_energrad<kMinimize, REC_LDS, PLATEAU, EXACT, LAYOUT>(a, s_rec, s_par, s_side, s_kh, nullptr, ws, P[0],
                                                                  P[1], st, lane, uA, uB, energy);
        const int i = acc_index(lane);
        if (!(lane & 3) && i < kAcc) {
          bs->acc[sa][i] = uA;
          if (sbb >= 0) bs->acc[sbb][i] = uB;
        }
        if (partner >= 0) b = partner;
      }
      wave_sync();
      // optimizer update, one lane per (slot, parameter): the operations of optimizer_update
      const int nst_s = __shfl(nst, src), jid_s = __shfl(jid, src);
      const bool act = slot && s < nst_s;
      float x = 0.f, m1 = 0.f, m2 = 0.f, c1 = 0.f, c2 = 0.f, E = 0.f;
      if (act) {
        const AtjOptimizer o = a.jobs[jid_s].opt;
        c1 = fmaf(o.beta1, bs->acc[sb][kAcc], 1.f - o.beta1);
        c2 = fmaf(o.beta2, bs->acc[sb][kAcc + 1], 1.f - o.beta2);
        x = optimizer_update_param(o, bs->prm[sb], bs->acc[sb], pi, bs->m1[sb][pi], bs->m2[sb][pi], c1, c2,
                                   m1, m2, E);
      }
      wave_sync();
      if (act) {
        bs->prm[sb][pi] = x;
        bs->m1[sb][pi] = m1;
        bs->m2[sb][pi] = m2;
        if (pi == 0) {
          bs->acc[sb][kAcc] = c1;
          bs->acc[sb][kAcc + 1] = c2;
          bs->acc[sb][kAcc + 2] = E;
        }
      }
      wave_sync();
      if (act && pi < 4) x = quat_renorm(bs->prm[sb], pi);
      wave_sync();
      if (act && pi < 4) bs->prm[sb][pi] = x;
      wave_sync();
      if (act) pose12_elements(bs->prm[sb], bs->M[sb], pi);
      wave_sync();
      phase_end(st, kPhUpdate);
    }
    if (own) {
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        a.pose7[7 * p + i] = bs->prm[lane][i];
        a.mom1[7 * p + i] = bs->m1[lane][i];
        a.mom2[7 * p + i] = bs->m2[lane][i];
      }
      a.count[p] = cnt + nst;
      a.elast[p] = bs->acc[lane][kAcc + 2];
    }
    wave_sync();
  }
}

template <int MODE, bool REC_LDS, bool PLATEAU, bool EXACT, int LAYOUT>
__global__ void __launch_bounds__(kWave * kMaxWavesPerBlock)
    pose_kernel(KArgs a) {
  extern __shared__ float4 smem[];
  const int nwaves = blockDim.x / kWave;
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int ntp = a.nrt * a.nlt;
  const int nslots = pair_slots(a.nrt, a.nlt);
  // LDS: pair records [nslots + 1] | side records [side_slots(nslots)] | kh [kh_slots(nslots)] |
  // receptor [nrec + 1] |
  // wave scratch [nwaves] | minimizer batches [nwaves], or pose matrices [nwaves][kBatch][12] and
  // (adjoint) the pair-table gradient [3][nslots], rows strided as the pair records
  float4* s_par = smem;
  float2* s_side = reinterpret_cast<float2*>(s_par + (nslots + 1));
  float* s_kh = reinterpret_cast<float*>(s_side + side_slots(nslots));
  float4* s_rec = reinterpret_cast<float4*>(s_kh + kh_slots(nslots));
  const int nrec_lds = REC_LDS ? a.nrec + 1 : 0;
  WaveScratch* s_ws = reinterpret_cast<WaveScratch*>(s_rec + nrec_lds);
  float* s_mat = reinterpret_cast<float*>(s_ws + nwaves);
  // kAdjoint: the block's fixed-point pair-table gradient; kAdjointWave: one float table
  // [3][nslots] per wavefront
  float* s_pgrad = is_adjoint(MODE) ? s_mat + nwaves * (kBatch * 12) : nullptr;
  stage_lds(a, REC_LDS ? s_rec : nullptr, s_par, s_side, s_kh, PLATEAU, EXACT, s_pgrad,
            MODE == kAdjoint ? adjoint_table_floats(nslots) : nwaves * 3 * nslots);
  __syncthreads();
  WaveScratch* ws = s_ws + wave;
  float* w_pgrad = MODE == kAdjointWave ? s_pgrad + wave * 3 * nslots : s_pgrad;
  WaveStats st = {};
#ifdef ATJ_PHASE_CLOCKS
  st.last = __builtin_amdgcn_s_memtime();
#endif

  if (MODE == kMinimize) {
    BatchScratch* bs = reinterpret_cast<BatchScratch*>(s_ws + nwaves) + wave;
    minimize_wave<REC_LDS, PLATEAU, EXACT, LAYOUT>(a, s_rec, s_par, s_side, s_kh, ws, bs, st, lane);
  } else {
    // a.batch poses of the processing order per claim; two consecutive ones share the bead phase
    // when share_bead_phase allows.  Lane b fetches pose b's index, weight and matrix
    // for the whole claim at once (one round of memory latency per claim, not per pose).
    const bool qt = (a.flags & ATJ_POSE_QT) != 0;
    float* Ms = s_mat + wave * (kBatch * 12);
    // claims in a fixed order (wavefront g of the grid: claims g, g + G, g + 2G, ...): no
    // atomics, and each adjoint wavefront table sums the same poses in the same order every run
    long long wave_id = (long long)blockIdx.x * nwaves + wave;
    const long long stride_claims = (long long)gridDim.x * nwaves;
    for (Claim claim = claim_of(a, wave_id); claim.k0 < a.N; claim = claim_of(a, wave_id += stride_claims)) {
      const int k0 = (int)claim.k0;
      const int nbatch = claim.n;
      int p_l = 0;
      float w_l = 0.f;
      wave_sync();  // the previous claim's readers of Ms are done
      if (lane < nbatch) {
        p_l = a.pose_order ? a.pose_order[k0 + lane] : k0 + lane;
        if (is_adjoint(MODE)) w_l = a.weights[p_l];
        float Mp[16];
        load_pose_matrix(a.poses + (size_t)p_l * (qt ? 7 : 16), qt, Mp);
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l) Ms[12 * lane + 3 * k + l] = Mp[4 * k + l];
      }
      wave_sync();
      for (int b = 0; b < nbatch; ++b) {
        int p[2];
        PoseIn P[2];
        p[0] = __builtin_amdgcn_readlane(p_l, b);
        const AtjJob job = a.jobs[a.pose_job ? a.pose_job[p[0]] : 0];
        P[0].lig = reinterpret_cast<const float4*>(job.ligand.beads);
        P[0].nb = job.ligand.nbeads;
        P[0].M = Ms + 12 * b;
        P[0].wgt = is_adjoint(MODE) ? lane_value(w_l, b) : 0.f;
        P[1] = P[0];
        P[1].nb = 0;
        p[1] = -1;
        // the next pose is read only when this one can share its bead phase
        if (P[0].nb <= kHalf && b + 1 < nbatch) {
          const int q = __builtin_amdgcn_readlane(p_l, b + 1);
          const AtjJob job2 = a.jobs[a.pose_job ? a.pose_job[q] : 0];
          if (share_bead_phase(P[0].nb, job2.ligand.nbeads)) {
            p[1] = q;
            P[1].lig = reinterpret_cast<const float4*>(job2.ligand.beads);
            P[1].nb = job2.ligand.nbeads;
            P[1].M = Ms + 12 * (b + 1);
            P[1].wgt = is_adjoint(MODE) ? lane_value(w_l, b + 1) : 0.f;
          }
        }
        float u[2];
        pose_energrad<MODE, REC_LDS, PLATEAU, EXACT, LAYOUT>(a, s_rec, s_par, s_side, s_kh, w_pgrad, ws, P[0], P[1], st,
                                                             lane, u[0], u[1]);
        if (!is_adjoint(MODE)) {
          if (a.flags & ATJ_GRAD_QT) {
            write_energrad_qt(a, p[0], u[0], lane);
            if (p[1] >= 0) write_energrad_qt(a, p[1], u[1], lane);
          } else {
            write_energrad(a, p[0], u[0], lane);
            if (p[1] >= 0) write_energrad(a, p[1], u[1], lane);
          }
        }
        if (p[1] >= 0) ++b;
      }
    }
  }
  if (a.stats) {
#pragma unroll
    for (int i = 0; i < ATJ_NSTATS; ++i) {
      const unsigned s = wave_sum_u(st.n[i]);
      if (lane == 0 && s) atomicAdd(a.stats + i, (unsigned long long)s);
    }
#ifdef ATJ_PHASE_CLOCKS
    phase_end(st, kPhSetup);
#pragma unroll
    for (int i = 0; i < kNumPhases; ++i)
      if (lane == 0) atomicAdd(a.stats + ATJ_NSTATS + i, st.clk[i]);
#endif
  }
  if (is_adjoint(MODE)) {  // the block's row of partial sums
    __syncthreads();
    float* out = a.partial + (size_t)blockIdx.x * 3 * ntp;
    const int stride = pair_stride(a.nlt);
    for (int i = threadIdx.x; i < 3 * ntp; i += blockDim.x) {
      const int c = i / ntp, k = i - c * ntp, rt = k / a.nlt;
      const int idx = c * nslots + rt * stride + (k - rt * a.nlt);
      if (MODE == kAdjoint) {  // the fixed-point sum, plus the float sum of large terms
        const unsigned long long* t = reinterpret_cast<const unsigned long long*>(s_pgrad);
        out[i] = fixed_value(t + idx, 3 * nslots) + s_pgrad[2 * kFixWords * 3 * nslots + idx];
      } else {  // the wavefront tables added in order
        float v = s_pgrad[idx];
        for (int w = 1; w < nwaves; ++w) v += s_pgrad[w * 3 * nslots + idx];
        out[i] = v;
      }
    }
  }
}

// Ligand RMSD of poses (test-rmsd.py:13-17): one wavefront per pose, lanes over beads,
// sum of squared deviations reduced across the wave in a fixed order.
__global__ void ligand_rmsd_kernel(const float* __restrict__ poses, int N, const float* __restrict__ lig,
                                   const float* __restrict__ ref, int nb, float* __restrict__ out) {
  const int lane = threadIdx.x % kWave;
  const int p = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (p >= N) return;  // whole waves exit together
  float M[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) M[i] = uni(poses[(size_t)p * 16 + i]);
  float acc = 0.f;
  for (int j = lane; j < nb; j += kWave) {
    float x[3];
    transform(M, lig[3 * j], lig[3 * j + 1], lig[3 * j + 2], x);
    const float dx = x[0] - ref[3 * j], dy = x[1] - ref[3 * j + 1], dz = x[2] - ref[3 * j + 2];
    acc += fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  }
  const float s = wave_sum(acc);
  if (lane == 0) out[p] = sqrtf(s / (float)nb);
}

__global__ void fill_int(int* p, int v) { *p = v; }

__global__ void zero_floats(float* p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// Per-(pose, bead) terms; one thread per bead-instance (parity / debugging path).
// With coords != NULL the items are explicit points (coords[idx], types[idx]) instead
// of (pose, bead) pairs.
__global__ void bead_terms_kernel(KArgs a, const float4* lig, int nb, const float* coords, const int* types,
                                  float* grid4, int* voxel_out, float* nbene, float* nbgrad) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)a.N * nb) return;
  const bool qt = (a.flags & ATJ_POSE_QT) != 0;
  const bool plateau = (a.flags & ATJ_NB_PLATEAU) != 0;
  const bool exact = (a.flags & ATJ_GRAD_EXACT) != 0;
  float xp[3];
  int tb;
  if (coords) {
    xp[0] = coords[3 * idx];
    xp[1] = coords[3 * idx + 1];
    xp[2] = coords[3 * idx + 2];
    tb = types[idx];
  } else {
    const int p = (int)(idx / nb), j = (int)(idx % nb);
    float M[16];
    load_pose_matrix(a.poses + (size_t)p * (qt ? 7 : 16), qt, M);
    const float4 b = lig[j];
    tb = __float_as_int(b.w);
    transform(M, b.x, b.y, b.z, xp);
  }
  const int gt = min((int)((unsigned)tb & 0xffffu), a.G.ntypes - 1), lt = min((int)((unsigned)tb >> 16), a.nlt - 1);
  float E, g[3], raw[4];
  const bool do_grid = (a.flags & ATJ_TERM_GRID) != 0, do_nb = (a.flags & ATJ_TERM_NB) != 0;
  int voxel, cnt, start;
  unsigned in_any;
  if (a.G.layout == ATJ_LAYOUT_CORNER)
    bead_grid<ATJ_LAYOUT_CORNER>(a.G, gt, xp, do_grid, do_nb, exact, E, g, voxel, cnt, start, in_any, raw);
  else
    bead_grid<ATJ_LAYOUT_PLAIN>(a.G, gt, xp, do_grid, do_nb, exact, E, g, voxel, cnt, start, in_any, raw);
  if (grid4) {
#pragma unroll
    for (int i = 0; i < 4; ++i) grid4[idx * 4 + i] = raw[i];
  }
  if (voxel_out) voxel_out[idx] = voxel;
  float e = 0.f, gg[3] = {0.f, 0.f, 0.f};
  if (voxel >= 0 && do_nb) {
    for (int c = 0; c < cnt; ++c) {
      const unsigned r = a.G.nblist[start + c];
      if (r >= (unsigned)a.nrec) continue;  // not a bead of this receptor
      const float4 rb = clamp_rec_type(a.rec[r], a.nrt);
      const float ddx = xp[0] - rb.x, ddy = xp[1] - rb.y, ddz = xp[2] - rb.z;
      const float dsq = fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx));
      if (dsq < a.G.p2) {
        const int rt = __float_as_int(rb.w);
        const int pi = rt * a.nlt + lt;
        const PairType t =
            pair_type(a.rc[pi], a.ac[pi], a.ivor[pi], a.emin[pi], a.rmin2[pi], plateau, exact, plateau_const(a.G.p2));
        float Ep, gs;
        pair_eval_rt(plateau, exact, dsq, t, Ep, gs);
        e += Ep;
        gg[0] = fmaf(gs, ddx, gg[0]);
        gg[1] = fmaf(gs, ddy, gg[1]);
        gg[2] = fmaf(gs, ddz, gg[2]);
      }
    }
  }
  if (nbene) nbene[idx] = e;
  if (nbgrad) {
#pragma unroll
    for (int i = 0; i < 3; ++i) nbgrad[idx * 3 + i] = gg[i];
  }
}

// Plain one-launch-per-step ADAM (the per-job comparison form).
__global__ void optimizer_step_kernel(const float* dEdM, float* energy, int N, float* pose7, float* mom1,
                                      float* mom2, int* count, AtjOptimizer o) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const float* G = dEdM + (size_t)p * 16;
  float* P = pose7 + (size_t)p * 7;
  const float acc[kAcc] = {energy ? energy[p] : 0.f, G[0], G[1], G[2], G[4], G[5], G[6], G[8], G[9], G[10],
                           G[12], G[13], G[14]};
  const int cnt = count[p] + 1;
  // bias corrections 1 - b^t by the recurrence c_t = b c_{t-1} + (1 - b), c_0 = 0
  float c1 = 0.f, c2 = 0.f;
  for (int i = 0; i < cnt; ++i) {
    c1 = fmaf(o.beta1, c1, 1.f - o.beta1);
    c2 = fmaf(o.beta2, c2, 1.f - o.beta2);
  }
  float prm[7], m1[7], m2[7];
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    prm[i] = P[i];
    m1[i] = mom1[7 * p + i];
    m2[i] = mom2[7 * p + i];
  }
  const float E = optimizer_update(o, prm, m1, m2, acc, c1, c2);
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    P[i] = prm[i];
    mom1[7 * p + i] = m1[i];
    mom2[7 * p + i] = m2[i];
  }
  if (energy) energy[p] = E;
  count[p] = cnt;
}

// Column sums of partial [blocks][n] in a fixed order (deterministic): a block of 256
// threads takes 64 columns; thread (c, r) sums rows r, r + 4, ... of column c, and the four
// row groups are added in order.
constexpr int kReduceCols = 64, kReduceRows = 4;
__global__ void reduce_partials_kernel(const float* partial, int blocks, int n, float* out) {
  __shared__ float part[kReduceRows][kReduceCols];
  const int c = threadIdx.x % kReduceCols, r = threadIdx.x / kReduceCols;
  const int i = blockIdx.x * kReduceCols + c;
  float s = 0.f;
  if (i < n) {
    // unrolled so that eight loads are in flight together; the additions keep their order
#pragma unroll 8
    for (int b = r; b < blocks; b += kReduceRows) s += partial[(size_t)b * n + i];
  }
  part[r][c] = s;
  __syncthreads();
  if (r == 0 && i < n) out[i] = ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c];
}

// ---------------------------------------------------------------------------
// Grid construction
// ---------------------------------------------------------------------------

// Potential + force (ATTRACT sign, = -dE/dp) of the ligand grid types at every grid point:
// plateau-capped inside plateaudis, zero beyond cutoff (test-calc-grid-energy.py:132-148); the
// force of a capped receptor bead is the plateau gradient projected on the plateau sphere (the
// complement of the force-shift of test-grid-energrad.py:140-146).
//
// One thread per grid point computes up to kBuildTypes ligand types (blockIdx.y selects the
// group).  Every distance it meets is at least plateaudis (capped), so a pair's energy is
// ivor (rc u^4 - ac u^3) with u = 1/max(dsq, P^2), its branch below rmin2 only mattering for
// pair types with ivor != 1 and rmin2 > P^2.  The potential of type t at the point therefore
// factorises over receptor types rt:
//   E_t = sum_rt ivor rc(rt, t) U4(rt) - ivor ac(rt, t) U3(rt),
//   F_t = sum_rt 4 ivor rc(rt, t) V5(rt) - 3 ivor ac(rt, t) V4(rt),
// with U4 = sum u^4, U3 = sum u^3, V5 = sum 2 c u^5 d, V4 = sum 2 c u^4 d over the beads of type
// rt within the cutoff (d = point - bead, c = P / |d| for capped beads, else 1).  The bead loop
// computes the 8 sums once for all types; the pair types whose lower branch can hold get their
// difference added bead by bead.  Beads are staged in LDS in type order (a stable counting sort
// per block, so sums run in the same order in every block and every run).  The contraction
// coefficients (ivor rc, -(ivor ac), -3 ivor ac) of up to kBuildRowTile receptor types are
// formed once per block into LDS, with a mask per receptor type of the ligand types whose lower
// branch applies, so a grid point spends 8 FMAs per (receptor type, ligand type).
constexpr int kBuildTypes = 32;
constexpr int kBuildChunk = 2048;  // receptor beads staged per pass
constexpr int kBuildThreads = 256;
constexpr int kBuildRowTile = 32;  // receptor types whose coefficients are staged at once

// LDS bytes of build_potentials_kernel for nrt receptor types.
__host__ __device__ __forceinline__ size_t build_lds_bytes(int nrt) {
  return (size_t)kBuildChunk * 16 + (size_t)kBuildRowTile * kBuildTypes * 16 + (size_t)kBuildRowTile * 4 +
         (size_t)(2 * nrt + 1 + kBuildThreads) * 4;
}

// Coefficients of receptor types rt0 .. rt0 + nrow - 1 against the block's ligand types
// t0 .. t0 + ng - 1: s_coef[r][k] = (ivor rc, -(ivor ac), -3 (ivor ac), 0), the products the
// contraction multiplies by U4, U3 and V4 (V5's factor 4 ivor rc is taken as ivor rc times 4 V5,
// the same product), and bit k of s_low[r] set where the lower branch applies (ivor != 1 and
// rmin2 > P^2).  Ligand types beyond the group get zeros.
__device__ void stage_build_coefs(const float* rc, const float* ac, const float* ivor, const float* rmin2, int nlt,
                                  const int* lig_fftype, int t0, int ng, int rt0, int nrow, float P2,
                                  float4* s_coef, unsigned* s_low) {
  static_assert(kBuildTypes == 32 && kBuildThreads % kWave == 0, "two coefficient rows per wavefront");
  const int lane = threadIdx.x % kWave;
  for (int e0 = threadIdx.x - lane; e0 < kBuildRowTile * kBuildTypes; e0 += blockDim.x) {
    const int e = e0 + lane, r = e / kBuildTypes, k = e % kBuildTypes;
    float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
    bool low = false;
    if (r < nrow && k < ng) {
      // types beyond the pair table are read as its last row / column (memory-safe for any input)
      const int lt = (int)min((unsigned)lig_fftype[t0 + k], (unsigned)(nlt - 1));
      const int pi = (rt0 + r) * nlt + lt;
      const float iv = ivor[pi];
      const float a = iv * rc[pi], bb = iv * ac[pi];
      c = make_float4(a, -bb, -3.f * bb, 0.f);
      low = iv != 1.f && rmin2[pi] > P2;
    }
    s_coef[e] = c;
    const unsigned long long m = __builtin_amdgcn_ballot_w64(low);
    if (lane == 0) s_low[r] = (unsigned)m;
    if (lane == kWave / 2) s_low[r] = (unsigned)(m >> 32);
  }
}

__device__ void stage_by_type(const float4* rec, int c0, int n, int nrt, float4* s_bead, int* s_off, int* s_cur,
                              int* s_type) {
  const int tid = threadIdx.x;
  for (int i = tid; i <= nrt; i += blockDim.x) s_off[i] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += blockDim.x) {
    const int t = (int)min((unsigned)__float_as_int(rec[c0 + i].w), (unsigned)(nrt - 1));
    atomicAdd(&s_off[t + 1], 1);  // integer counts: the order of the increments does not matter
  }
  __syncthreads();
  if (tid == 0)
    for (int i = 1; i <= nrt; ++i) s_off[i] += s_off[i - 1];
  __syncthreads();
  for (int i = tid; i < nrt; i += blockDim.x) s_cur[i] = s_off[i];
  __syncthreads();
  // rounds of blockDim beads: a bead's place is its type's cursor plus the beads of its type
  // earlier in the round (stable, so the bead order within a type is the receptor order)
  for (int r0 = 0; r0 < n; r0 += blockDim.x) {
    const int i = r0 + tid;
    float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
    int t = -1;
    if (i < n) {
      b = rec[c0 + i];
      t = (int)min((unsigned)__float_as_int(b.w), (unsigned)(nrt - 1));
      b.w = __int_as_float(t);
    }
    s_type[tid] = t;
    __syncthreads();
    int rank = 0, later = 0;
    if (t >= 0) {
      for (int j = 0; j < tid; ++j) rank += s_type[j] == t;
      for (int j = tid + 1; j < (int)blockDim.x; ++j) later += s_type[j] == t;
      s_bead[s_cur[t] + rank] = b;
    }
    __syncthreads();
    if (t >= 0 && later == 0) s_cur[t] += rank + 1;  // the round's last bead of the type
    __syncthreads();
  }
}

// Three blocks per CU (three wavefronts per SIMD, at most 168 VGPRs): the bead loop waits on
// LDS, and the 128 output accumulators leave no room for a fourth.
__global__ void __launch_bounds__(kBuildThreads, 3)
    build_potentials_kernel(GridView G, const float4* rec, int nrec, const float* rc, const float* ac,
                            const float* ivor, const float* emin, const float* rmin2, int nrt, int nlt,
                            const int* lig_fftype, int ntypes, float cutoff2, float* out, int dx, int dy, int dz,
                            int outer) {
  extern __shared__ float4 smem[];
  float4* s_bead = smem;
  float4* s_coef = s_bead + kBuildChunk;
  unsigned* s_low = reinterpret_cast<unsigned*>(s_coef + kBuildRowTile * kBuildTypes);
  int* s_off = reinterpret_cast<int*>(s_low + kBuildRowTile);
  int* s_cur = s_off + (nrt + 1);
  int* s_type = s_cur + nrt;
  const int t0 = blockIdx.y * kBuildTypes;
  const int ng = min(kBuildTypes, ntypes - t0);
  const long long nv = (long long)dx * dy * dz;
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long vc = v < nv ? v : nv - 1;  // threads past the end compute a copy and store nothing
  const int iz = (int)(vc % dz), iy = (int)((vc / dz) % dy), ix = (int)(vc / ((long long)dy * dz));
  float px, py, pz;
  if (outer) {
    px = G.ox + G.spacing * (2.f * ix - G.ge);
    py = G.oy + G.spacing * (2.f * iy - G.ge);
    pz = G.oz + G.spacing * (2.f * iz - G.ge);
  } else {
    px = G.ox + G.spacing * ix;
    py = G.oy + G.spacing * iy;
    pz = G.oz + G.spacing * iz;
  }
  const float P2 = G.p2;
  const float P = sqrtf(P2);
  const float twoP = 2.f * P;
  float E[kBuildTypes], Fx[kBuildTypes], Fy[kBuildTypes], Fz[kBuildTypes];
#pragma unroll
  for (int k = 0; k < kBuildTypes; ++k) E[k] = Fx[k] = Fy[k] = Fz[k] = 0.f;
  for (int c0 = 0; c0 < nrec; c0 += kBuildChunk) {
    const int n = min(kBuildChunk, nrec - c0);
    __syncthreads();  // the previous chunk's readers are done
    stage_by_type(rec, c0, n, nrt, s_bead, s_off, s_cur, s_type);
    for (int rt0 = 0; rt0 < nrt; rt0 += kBuildRowTile) {
      const int nrow = min(kBuildRowTile, nrt - rt0);
      __syncthreads();  // the previous tile's readers are done
      stage_build_coefs(rc, ac, ivor, rmin2, nlt, lig_fftype, t0, ng, rt0, nrow, P2, s_coef, s_low);
      __syncthreads();
      for (int rt = rt0; rt < rt0 + nrow; ++rt) {
        const int b0 = s_off[rt], b1 = s_off[rt + 1];
        if (b0 == b1) continue;
        float U4 = 0.f, U3 = 0.f, V5[3] = {0.f, 0.f, 0.f}, V4[3] = {0.f, 0.f, 0.f};
        // unrolled so that several LDS reads of beads are in flight at once
#pragma unroll 4
        for (int b = b0; b < b1; ++b) {
          const float4 rb = s_bead[b];
          const float ddx = px - rb.x, ddy = py - rb.y, ddz = pz - rb.z;
          const float dsq = fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx));
          const bool capped = dsq < P2;
          // max(dsq, P^2) within the cutoff; beyond it (and for NaN) +infinity, whose reciprocal
          // 0 makes u^3, u^4 and both weights of the bead exactly 0.  Hardware reciprocal and
          // reciprocal square root (1 ulp): the grid is data compared with the FP64 restatement
          // at 1e-4, and IEEE division and square root were most of this loop's instructions.
          const float dd = dsq < cutoff2 ? fmaxf(dsq, P2) : INFINITY;
          const float u = __builtin_amdgcn_rcpf(dd);
          const float u3 = (u * u) * u;
          const float u4 = u3 * u;
          // 2c: c = P / |d| for capped beads (0 at d = 0, where the force has no direction), else 1
          const float c2 = capped ? (dsq > 0.f ? twoP * __builtin_amdgcn_rsqf(dsq) : 0.f) : 2.f;
          const float w4 = c2 * u4;
          const float w5 = w4 * u;
          U4 += u4;
          U3 += u3;
          V5[0] = fmaf(w5, ddx, V5[0]);
          V5[1] = fmaf(w5, ddy, V5[1]);
          V5[2] = fmaf(w5, ddz, V5[2]);
          V4[0] = fmaf(w4, ddx, V4[0]);
          V4[1] = fmaf(w4, ddy, V4[1]);
          V4[2] = fmaf(w4, ddz, V4[2]);
        }
        // 4 ivor rc V5 as ivor rc (4 V5): scaling by 4 is exact, so the fused products agree
        const float V5s[3] = {4.f * V5[0], 4.f * V5[1], 4.f * V5[2]};
        const float4* coef = s_coef + (rt - rt0) * kBuildTypes;
        const unsigned low = __builtin_amdgcn_readfirstlane(s_low[rt - rt0]);
        // types k >= ng of the last group have zero coefficients and no low bit; their sums are
        // not stored
#pragma unroll
        for (int k = 0; k < kBuildTypes; ++k) {
          const float4 cf = coef[k];
          E[k] = fmaf(cf.x, U4, fmaf(cf.y, U3, E[k]));
          Fx[k] = fmaf(cf.x, V5s[0], fmaf(cf.z, V4[0], Fx[k]));
          Fy[k] = fmaf(cf.x, V5s[1], fmaf(cf.z, V4[1], Fy[k]));
          Fz[k] = fmaf(cf.x, V5s[2], fmaf(cf.z, V4[2], Fz[k]));
          if (low & (1u << k)) {
            // the lower branch below rmin2: energy (1 - ivor)(vlj - emin), force from (1 - ivor) dvlj
            const int lt = (int)min((unsigned)lig_fftype[t0 + k], (unsigned)(nlt - 1));
            const int pi = rt * nlt + lt;
            const float iv = ivor[pi], r2 = rmin2[pi];
            const float em = emin[pi], crc = rc[pi], cac = ac[pi];
            for (int b = b0; b < b1; ++b) {
              const float4 rb = s_bead[b];
              const float ddx = px - rb.x, ddy = py - rb.y, ddz = pz - rb.z;
              const float dsq = fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx));
              const bool capped = dsq < P2;
              const float dd = capped ? P2 : dsq;
              if (!(dsq < cutoff2) || !(dd < r2)) continue;
              const float u = __builtin_amdgcn_rcpf(dd);
              const float u3 = (u * u) * u;
              const float vlj = (crc * u - cac) * u3;
              const float dv = -(fmaf(4.f, crc * u, -3.f * cac) * u3) * u;
              const float c = capped ? (dsq > 0.f ? P * __builtin_amdgcn_rsqf(dsq) : 0.f) : 1.f;
              const float gs = 2.f * (1.f - iv) * dv * c;
              E[k] += (1.f - iv) * (vlj - em);
              Fx[k] = fmaf(-gs, ddx, Fx[k]);
              Fy[k] = fmaf(-gs, ddy, Fy[k]);
              Fz[k] = fmaf(-gs, ddz, Fz[k]);
            }
          }
        }
      }
    }
  }
  if (v >= nv) return;
#pragma unroll
  for (int k = 0; k < kBuildTypes; ++k) {
    if (k >= ng) continue;
    *reinterpret_cast<float4*>(out + ((size_t)(t0 + k) * nv + v) * 4) = make_float4(E[k], Fx[k], Fy[k], Fz[k]);
  }
}

// Neighbour lists are built over bricks of kBrickX x kBrickY x kBrickZ grid points, one point
// per thread of a block.  The block stages, in receptor order, the beads that can lie within
// neighbourdis of some point of its brick (within neighbourdis + the brick's half diagonal +
// 0.01 A of the brick's centre, a margin far above the rounding of these float distances), and
// each point tests only those with the same arithmetic as a test against every bead: the
// counts and lists are those of the whole receptor.
constexpr int kBrickX = 4, kBrickY = 8, kBrickZ = 8;
constexpr int kBrickThreads = kBrickX * kBrickY * kBrickZ;
constexpr int kListCand = 2048;  // candidate beads staged per pass

struct ListBrick {
  long long v;  // the thread's grid point (flat index), -1 beyond the grid
  float px, py, pz;
  float cx, cy, cz, cull2;
};

__host__ __device__ __forceinline__ long long brick_count(int dx, int dy, int dz) {
  return (long long)((dx + kBrickX - 1) / kBrickX) * ((dy + kBrickY - 1) / kBrickY) * ((dz + kBrickZ - 1) / kBrickZ);
}

__device__ ListBrick list_brick(const GridView& G, float nbdis) {
  const int nby = (G.dy + kBrickY - 1) / kBrickY, nbz = (G.dz + kBrickZ - 1) / kBrickZ;
  const long long b = blockIdx.x;
  const int bz = (int)(b % nbz), by = (int)((b / nbz) % nby), bx = (int)(b / ((long long)nbz * nby));
  const int t = threadIdx.x;
  const int ix = bx * kBrickX + t / (kBrickY * kBrickZ), iy = by * kBrickY + (t / kBrickZ) % kBrickY,
            iz = bz * kBrickZ + t % kBrickZ;
  ListBrick k;
  // grid point coordinates exactly as the per-point test has them
  k.px = G.ox + G.spacing * ix;
  k.py = G.oy + G.spacing * iy;
  k.pz = G.oz + G.spacing * iz;
  k.v = (ix < G.dx && iy < G.dy && iz < G.dz) ? ((long long)ix * G.dy + iy) * G.dz + iz : -1;
  const int x1 = min(bx * kBrickX + kBrickX, G.dx) - 1, y1 = min(by * kBrickY + kBrickY, G.dy) - 1,
            z1 = min(bz * kBrickZ + kBrickZ, G.dz) - 1;
  const float lx = G.ox + G.spacing * (bx * kBrickX), hx = G.ox + G.spacing * x1;
  const float ly = G.oy + G.spacing * (by * kBrickY), hy = G.oy + G.spacing * y1;
  const float lz = G.oz + G.spacing * (bz * kBrickZ), hz = G.oz + G.spacing * z1;
  k.cx = 0.5f * (lx + hx);
  k.cy = 0.5f * (ly + hy);
  k.cz = 0.5f * (lz + hz);
  const float ex = 0.5f * (hx - lx), ey = 0.5f * (hy - ly), ez = 0.5f * (hz - lz);
  // |nbdis|: the point test depends on its square only.  The margin covers the rounding of
  // both distances, relative to the coordinates' magnitude, with 0.01 A to spare.
  const float reach = fabsf(nbdis) + sqrtf(fmaf(ez, ez, fmaf(ey, ey, ex * ex)));
  const float mag = fabsf(k.cx) + fabsf(k.cy) + fabsf(k.cz) + reach;
  const float cull = fmaf(mag, 1e-5f, reach + 0.01f);
  k.cull2 = cull * cull;
  return k;
}

// Stages into s_cand, in receptor order, the beads from index pos on that lie within the brick's
// cull radius (x, y, z and the receptor index as int bits), until the next round of
// blockDim.x beads might not fit or the receptor ends; pos moves past the beads examined.
// Returns the number staged.  Every thread of the block calls it; it ends with a barrier.
__device__ int stage_candidates(const float4* rec, int nrec, int& pos, const ListBrick& k, float4* s_cand,
                                int* s_wcount) {
  const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave, nw = blockDim.x / kWave;
  int n = 0;
  while (pos < nrec && n + (int)blockDim.x <= kListCand) {
    const int r = pos + tid;
    float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
    bool keep = false;
    if (r < nrec) {
      b = rec[r];
      const float dx = b.x - k.cx, dy = b.y - k.cy, dz = b.z - k.cz;
      keep = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) < k.cull2;
    }
    const unsigned long long m = __builtin_amdgcn_ballot_w64(keep);
    if (lane == 0) s_wcount[wave] = __builtin_popcountll(m);
    __syncthreads();
    int before = n, total = n;
    for (int w = 0; w < nw; ++w) {
      const int cw = s_wcount[w];
      before += w < wave ? cw : 0;
      total += cw;
    }
    if (keep) {
      const unsigned long long below = lane ? m & (~0ull >> (kWave - lane)) : 0ull;
      s_cand[before + __builtin_popcountll(below)] = make_float4(b.x, b.y, b.z, __int_as_float(r));
    }
    __syncthreads();
    n = total;
    pos += blockDim.x;
  }
  return n;
}

__global__ void __launch_bounds__(kBrickThreads)
    count_neighbours_kernel(GridView G, const float4* rec, int nrec, float nbdis, float nbdis2, int* counts) {
  __shared__ float4 s_cand[kListCand];
  __shared__ int s_wcount[kBrickThreads / kWave];
  const ListBrick k = list_brick(G, nbdis);
  int c = 0;
  for (int pos = 0; pos < nrec;) {
    __syncthreads();  // the previous pass's readers are done
    const int n = stage_candidates(rec, nrec, pos, k, s_cand, s_wcount);
    for (int i = 0; i < n; ++i) {
      const float4 rb = s_cand[i];
      const float ddx = k.px - rb.x, ddy = k.py - rb.y, ddz = k.pz - rb.z;
      const float dsq = fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx));
      c += dsq < nbdis2;
    }
  }
  if (k.v >= 0) counts[k.v] = c;
}

// The first counts[v] beads within neighbourdis of point v, in receptor order, then 0xFFFF
// padding to a multiple of kListPad entries.
__global__ void __launch_bounds__(kBrickThreads)
    fill_neighbours_kernel(GridView G, const float4* rec, int nrec, float nbdis, float nbdis2, const int* counts,
                           const long long* starts, int2* nbgrid, uint16_t* nblist) {
  __shared__ float4 s_cand[kListCand];
  __shared__ int s_wcount[kBrickThreads / kWave];
  const ListBrick k = list_brick(G, nbdis);
  const long long s = k.v >= 0 ? starts[k.v] : 0;
  const int n = k.v >= 0 ? counts[k.v] : 0;
  int c = 0;
  for (int pos = 0; pos < nrec;) {
    __syncthreads();  // the previous pass's readers are done
    const int nc = stage_candidates(rec, nrec, pos, k, s_cand, s_wcount);
    for (int i = 0; i < nc; ++i) {
      const float4 rb = s_cand[i];
      const float ddx = k.px - rb.x, ddy = k.py - rb.y, ddz = k.pz - rb.z;
      const float dsq = fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx));
      if (dsq < nbdis2 && c < n) nblist[s + c++] = (uint16_t)__float_as_int(rb.w);
    }
  }
  if (k.v < 0) return;
  const int padded = (n + kListPad - 1) / kListPad * kListPad;
  for (int j = n; j < padded; ++j) nblist[s + j] = 0xFFFFu;
  nbgrid[k.v] = make_int2(n, (int)s);
}

// Shell order of one grid point's list (atj_grid_shell_lists).  Entry r at distance d
// from the point goes to group 0 when d < r0 (= plateaudis + margin), else to group
// min(n + 1, ceil((d - r0) / h)) (h = shell width); groups in order, receptor order kept
// within a group.  Count k of nbshell[v] = the entries in groups 0..k + 1: every entry
// closer than r0 + (k + 1) h (the quotient rounds monotonically, so no such entry lands in
// a later group).  An entry that is not a bead of this receptor (r >= nrec: lists made for
// another receptor) goes to the last group, which no shell reads; the pair kernels read it
// as the far-away sentinel bead when they walk whole lists.
__device__ __forceinline__ int shell_group(float d, float r0, float h, int n) {
  return d < r0 ? 0 : min(n + 1, (int)ceilf((d - r0) / h));
}

// A counting sort in two passes over the list: the first counts the entries of each group,
// the second writes each entry at its group's running offset.  The n + 2 <= 8 group counters
// (each below 2^16: a list holds fewer than 65535 entries) are packed four to a 64-bit word,
// so the group index selects a shift rather than a register.
__global__ void shell_lists_kernel(GridView G, const float4* rec, int nrec, int n, float r0, float h, int npts,
                                   uint16_t* out, int4* nbshell) {
  static_assert(kMaxShell + 2 <= 8, "group counters: four 16-bit fields in each of two words");
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= npts) return;
  const int2 pl = G.nbgrid[v];
  const int iz = v % G.dz, iy = (v / G.dz) % G.dy, ix = v / (G.dy * G.dz);
  const float px = fmaf((float)ix, G.spacing, G.ox);
  const float py = fmaf((float)iy, G.spacing, G.oy);
  const float pz = fmaf((float)iz, G.spacing, G.oz);
  auto group_of = [&](unsigned r) {
    if (r >= (unsigned)nrec) return n + 1;
    const float4 rb = rec[r];
    const float dx = rb.x - px, dy = rb.y - py, dz = rb.z - pz;
    return shell_group(sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx))), r0, h, n);
  };
  // adds `one` << 16 g to the field of group g
  auto bump = [](unsigned long long& lo, unsigned long long& hi, int g, unsigned long long one) {
    const unsigned long long f = one << (16 * (g & 3));
    lo += g < 4 ? f : 0ull;
    hi += g < 4 ? 0ull : f;
  };
  auto field = [](unsigned long long lo, unsigned long long hi, int g) {
    return (int)(((g < 4 ? lo : hi) >> (16 * (g & 3))) & 0xffffu);
  };
  unsigned long long nlo = 0, nhi = 0;  // entries per group
  for (int t = 0; t < pl.x; ++t) bump(nlo, nhi, group_of(G.nblist[pl.y + t]), 1ull);
  // offsets: entries of the groups before; count k of the shells = entries of groups 0..k + 1
  unsigned long long olo = 0, ohi = 0;
  unsigned c[3] = {0u, 0u, 0u};
  int w = 0;
  for (int j = 0; j <= n + 1; ++j) {
    bump(olo, ohi, j, (unsigned long long)w);
    w += field(nlo, nhi, j);
    const int k = j - 1;  // shell k reads groups 0..k + 1
    if (k >= 0 && k < n) c[k >> 1] |= (unsigned)w << (16 * (k & 1));
  }
  for (int t = 0; t < pl.x; ++t) {
    const uint16_t r = G.nblist[pl.y + t];
    const int g = group_of(r);
    out[pl.y + field(olo, ohi, g)] = r;
    bump(olo, ohi, g, 1ull);
  }
  nbshell[v] = make_int4(pl.y, (int)c[0], (int)c[1], (int)c[2]);
}

__global__ void pack_corners_kernel(const float* plain, int ntypes, int dx, int dy, int dz, float* packed) {
  const long long nrec = (long long)ntypes * (dx - 1) * (dy - 1) * (dz - 1);
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nrec * 8) return;
  const int c = (int)(idx & 7);
  long long r = idx >> 3;
  const int z = (int)(r % (dz - 1));
  r /= (dz - 1);
  const int y = (int)(r % (dy - 1));
  r /= (dy - 1);
  const int x = (int)(r % (dx - 1));
  const int t = (int)(r / (dx - 1));
  const int ix = x + (c >> 2), iy = y + ((c >> 1) & 1), iz = z + (c & 1);
  const size_t src = (((size_t)t * dx + ix) * dy + iy) * (size_t)dz + iz;
  *reinterpret_cast<float4*>(packed + idx * 4) = *reinterpret_cast<const float4*>(plain + src * 4);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

int make_view(const AtjGrid* g, GridView& G, bool need_lists) {
  if (!g) return fail(ATJ_ERR_INVALID, "grid is NULL");
  for (int i = 0; i < 3; ++i)
    if (g->dim[i] < 2) return fail(ATJ_ERR_INVALID, "grid dim must be >= 2");
  if (!(g->spacing > 0.f)) return fail(ATJ_ERR_INVALID, "gridspacing must be > 0");
  if (g->layout != ATJ_LAYOUT_PLAIN && g->layout != ATJ_LAYOUT_CORNER)
    return fail(ATJ_ERR_INVALID, "unknown grid layout");
  if (g->outer)
    for (int i = 0; i < 3; ++i)
      if (g->dim2[i] < 2) return fail(ATJ_ERR_INVALID, "outer grid dim2 must be >= 2");
  if (need_lists && (!g->nbgrid || !g->nblist)) return fail(ATJ_ERR_INVALID, "neighbour lists are NULL");
  G.ox = g->origin[0];
  G.oy = g->origin[1];
  G.oz = g->origin[2];
  G.spacing = g->spacing;
  G.rspacing = 1.f / g->spacing;  // IEEE single division: correctly rounded
  G.ge = g->gridextension;
  G.p2 = g->plateaudis2;
  G.dx = g->dim[0];
  G.dy = g->dim[1];
  G.dz = g->dim[2];
  G.dx2 = g->dim2[0];
  G.dy2 = g->dim2[1];
  G.dz2 = g->dim2[2];
  if (g->inner && (g->ntypes <= 0 || g->ntypes > 65536))
    return fail(ATJ_ERR_INVALID, "ntypes must be 1..65536 with potential grids");
  G.ntypes = g->ntypes > 0 ? g->ntypes : 1;  // bead grid types are limited to [0, ntypes - 1]
  G.layout = g->layout;
  if (g->nshell < 0 || g->nshell > kMaxShell) return fail(ATJ_ERR_INVALID, "nshell must be 0..6");
  if (need_lists && g->nshell > 0 && !g->nbshell) return fail(ATJ_ERR_INVALID, "nbshell is NULL");
  G.nshell = g->nshell;
  G.shinv = (float)(2.0 * g->nshell / std::sqrt(3.0));
  G.lx = (float)(g->dim[0] - 1);
  G.ly = (float)(g->dim[1] - 1);
  G.lz = (float)(g->dim[2] - 1);
  G.lx2 = (float)(g->dim2[0] - 1);
  G.ly2 = (float)(g->dim2[1] - 1);
  G.lz2 = (float)(g->dim2[2] - 1);
  // per-type slab sizes; voxel and cell indices within a slab are 32-bit
  const bool corner = g->layout == ATJ_LAYOUT_CORNER;
  const unsigned long long pts = (unsigned long long)g->dim[0] * g->dim[1] * g->dim[2];
  const unsigned long long cells = (unsigned long long)(g->dim[0] - 1) * (g->dim[1] - 1) * (g->dim[2] - 1);
  if (pts >= (1ull << 31)) return fail(ATJ_ERR_UNSUPPORTED, "inner grid has 2^31 or more points");
  G.slab = (unsigned)(corner ? cells : pts);
  G.slab2 = 0;
  if (g->outer) {
    const unsigned long long pts2 = (unsigned long long)g->dim2[0] * g->dim2[1] * g->dim2[2];
    const unsigned long long cells2 = (unsigned long long)(g->dim2[0] - 1) * (g->dim2[1] - 1) * (g->dim2[2] - 1);
    if (pts2 >= (1ull << 31)) return fail(ATJ_ERR_UNSUPPORTED, "outer grid has 2^31 or more points");
    G.slab2 = (unsigned)(corner ? cells2 : pts2);
  }
  G.inner = g->inner;
  G.outer = g->outer;
  G.nbgrid = reinterpret_cast<const int2*>(g->nbgrid);
  G.nblist = g->nblist;
  G.nbshell = reinterpret_cast<const int4*>(g->nbshell);
  return ATJ_OK;
}

int fill_common(KArgs& a, const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec, uint32_t flags) {
  if (!ff || !rec) return fail(ATJ_ERR_INVALID, "forcefield / receptor is NULL");
  int rc = make_view(grid, a.G, (flags & ATJ_TERM_NB) != 0);
  if (rc) return rc;
  if ((flags & ATJ_TERM_GRID) && !grid->inner)
    return fail(ATJ_ERR_INVALID, "grid term requested but the grid has no potentials");
  if (ff->nrt <= 0 || ff->nlt <= 0) return fail(ATJ_ERR_INVALID, "empty pair table");
  if ((long long)ff->nrt * ff->nlt > 4096) return fail(ATJ_ERR_UNSUPPORTED, "pair table larger than 4096 entries");
  if (!ff->rc || !ff->ac || !ff->ivor || !ff->emin || !ff->rmin2) return fail(ATJ_ERR_INVALID, "pair table array is NULL");
  if (rec->nrec < 0 || rec->nrec > 65535) return fail(ATJ_ERR_UNSUPPORTED, "receptor must have < 65536 beads");
  if (rec->nrec > 0 && !rec->beads) return fail(ATJ_ERR_INVALID, "receptor beads NULL");
  a.rec = reinterpret_cast<const float4*>(rec->beads);
  a.nrec = rec->nrec;
  a.rc = ff->rc;
  a.ac = ff->ac;
  a.ivor = ff->ivor;
  a.emin = ff->emin;
  a.rmin2 = ff->rmin2;
  a.nrt = ff->nrt;
  a.nlt = ff->nlt;
  a.flags = flags;
  return ATJ_OK;
}

struct DevInfo {
  int cus = 0;
  size_t lds_per_block = 0;
  bool ok = false;
};

int device_info(DevInfo& d) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  static thread_local int cached_dev = -1;
  static thread_local DevInfo cached;
  if (cached_dev == dev && cached.ok) {
    d = cached;
    return ATJ_OK;
  }
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, dev));
  d.cus = p.multiProcessorCount;
  d.lds_per_block = p.sharedMemPerBlock;
  d.ok = true;
  cached = d;
  cached_dev = dev;
  return ATJ_OK;
}

// Poses per work-counter claim for N poses on `waves` wavefronts: the size in 1, 2, 4, 8 with
// the fewest pose evaluations per wave on the longest path (rounds of claims x size), the
// largest on a tie (more poses of at most 32 beads then share a bead phase).  Only launches of
// fewer than 8 poses per wave take fewer than 8 poses per claim.  Each pose's result is the
// same whatever the claim size.
int claim_size(long long N, long long waves) {
  int best = kBatch;
  long long best_cost = -1;
  for (int b = kBatch; b >= 1; b >>= 1) {
    const long long rounds = (N + b * waves - 1) / (b * waves);
    const long long cost = rounds * b;
    if (best_cost < 0 || cost < best_cost) {
      best = b;
      best_cost = cost;
    }
  }
  return best;
}

template <int MODE, bool REC_LDS, bool PLATEAU, bool EXACT, int LAYOUT>
int launch_pose_kernel_t(KArgs& a, int wpb, size_t lds, int blocks, const DevInfo& d, hipStream_t stream) {
  auto kern = pose_kernel<MODE, REC_LDS, PLATEAU, EXACT, LAYOUT>;
  if (lds > 65536)
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
  int per_cu = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, wpb * kWave, lds));
  if (per_cu < 1) per_cu = 1;
  const int resident = per_cu * d.cus;
  const bool auto_blocks = blocks == 0;
  if (auto_blocks) blocks = resident;
  if (a.batch == 0) a.batch = claim_size(a.N, (long long)min(blocks, resident) * wpb);
  if (a.tail_batch == 0 || a.tail_batch >= a.batch || is_adjoint(MODE)) a.tail_batch = a.batch;
  // a launch whose poses all fall in the tail (its rounds of claims cover them) is claimed
  // tail_batch at a time from the start
  const int tail_rounds = a.tail_start;
  int first = a.batch;
  if (a.tail_batch < a.batch &&
      (long long)a.N <= (long long)tail_rounds * min(blocks, resident) * wpb * a.batch)
    first = a.tail_batch;
  if (auto_blocks) {  // no more blocks than the first round of claims needs
    const long long per_block = (long long)wpb * first;
    const long long need = ((long long)a.N + per_block - 1) / per_block;
    if (need < blocks) blocks = (int)(need > 0 ? need : 1);
  }
  a.static_blocks = min(blocks, resident);
  // the tail: the poses of the last tail_rounds rounds of claims, claimed a.tail_batch at a time
  if (a.tail_batch < a.batch) {
    const long long round = (long long)a.static_blocks * wpb * a.batch;
    const long long t = (long long)a.N - (long long)tail_rounds * round;
    a.tail_start = t <= 0 ? 0 : (int)(t / a.batch * a.batch);
  } else {
    a.tail_start = a.N;
  }
  if (!static_claims(MODE)) {  // the minimizer's claims come from the work counter
    hipLaunchKernelGGL(fill_int, dim3(1), dim3(1), 0, stream, a.counter, 0);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(wpb * kWave), lds, stream, a);
  HIPCHK(hipGetLastError());
  return ATJ_OK;
}

template <int MODE, bool REC_LDS, bool PLATEAU, bool EXACT>
int launch_layout(KArgs& a, int wpb, size_t lds, int blocks, const DevInfo& d, hipStream_t stream) {
  if (a.G.layout == ATJ_LAYOUT_CORNER)
    return launch_pose_kernel_t<MODE, REC_LDS, PLATEAU, EXACT, ATJ_LAYOUT_CORNER>(a, wpb, lds, blocks, d, stream);
  return launch_pose_kernel_t<MODE, REC_LDS, PLATEAU, EXACT, ATJ_LAYOUT_PLAIN>(a, wpb, lds, blocks, d, stream);
}

template <int MODE, bool REC_LDS>
int launch_flags(KArgs& a, int wpb, size_t lds, int blocks, const DevInfo& d, hipStream_t stream) {
  // the adjoint's parameter derivatives are those of the energy: the gradient convention
  // (exact or force-shifted) does not enter them
  const bool plateau = (a.flags & ATJ_NB_PLATEAU) != 0;
  if constexpr (!is_adjoint(MODE)) {
    if (a.flags & ATJ_GRAD_EXACT) {
      if (plateau) return launch_layout<MODE, REC_LDS, true, true>(a, wpb, lds, blocks, d, stream);
      return launch_layout<MODE, REC_LDS, false, true>(a, wpb, lds, blocks, d, stream);
    }
  }
  if (plateau) return launch_layout<MODE, REC_LDS, true, false>(a, wpb, lds, blocks, d, stream);
  return launch_layout<MODE, REC_LDS, false, false>(a, wpb, lds, blocks, d, stream);
}

template <int MODE>
int launch_pose_kernel(KArgs& a, const AtjLaunch* launch, hipStream_t stream) {
  if (!a.counter && !static_claims(MODE)) return fail(ATJ_ERR_INVALID, "work counter is NULL");
  if (!a.jobs) return fail(ATJ_ERR_INVALID, "jobs is NULL");
  int wpb = launch && launch->waves_per_block > 0 ? launch->waves_per_block : 4;
  if (wpb > kMaxWavesPerBlock) return fail(ATJ_ERR_INVALID, "waves_per_block > 16");
  const uint32_t options = launch ? launch->options : 0u;
  if (options & ~(ATJ_LAUNCH_REC_GLOBAL | ATJ_LAUNCH_CLAIM_MASK | ATJ_LAUNCH_TAIL_MASK | ATJ_LAUNCH_TAIL_ROUNDS_MASK |
                  ATJ_LAUNCH_ADJOINT_WAVE_TABLES))
    return fail(ATJ_ERR_INVALID, "unknown launch option");
  const int claim = (int)((options & ATJ_LAUNCH_CLAIM_MASK) >> 4);
  if (claim != 0 && claim != 1 && claim != 2 && claim != 4 && claim != kBatch)
    return fail(ATJ_ERR_INVALID, "poses per claim must be 1, 2, 4 or 8");
  const int tail = (int)((options & ATJ_LAUNCH_TAIL_MASK) >> 8);
  if (tail != 0 && tail != 1 && tail != 2 && tail != 4)
    return fail(ATJ_ERR_INVALID, "poses per tail claim must be 1, 2 or 4");
  const int tail_rounds = (int)((options & ATJ_LAUNCH_TAIL_ROUNDS_MASK) >> 12);
  a.batch = claim;  // 0: chosen at launch
  a.tail_batch = tail;
  a.tail_start = tail_rounds > 0 ? tail_rounds : 1;  // rounds until the launch shape is known
  DevInfo d;
  int rc = device_info(d);
  if (rc) return rc;
  const int nslots = pair_slots(a.nrt, a.nlt);
  const size_t fixed = pair_lds_bytes(nslots) + (MODE == kAdjoint ? adjoint_table_floats(nslots) * sizeof(float) : 0);
  const size_t per_wave = sizeof(WaveScratch) +
                          (MODE == kMinimize ? sizeof(BatchScratch) : kBatch * 12 * sizeof(float)) +
                          (MODE == kAdjointWave ? 3 * (size_t)nslots * sizeof(float) : 0);
  if constexpr (MODE == kAdjoint) {
    // the per-wavefront float tables when asked for, or when the block's fixed-point table
    // leaves no room for a wavefront (pair tables of more than ~1,800 slots)
    if ((options & ATJ_LAUNCH_ADJOINT_WAVE_TABLES) || fixed + per_wave > d.lds_per_block)
      return launch_pose_kernel<kAdjointWave>(a, launch, stream);
  }
  const size_t rec_bytes = (size_t)(a.nrec + 1) * sizeof(float4);
  bool rec_lds = !(options & ATJ_LAUNCH_REC_GLOBAL);
  if (is_adjoint(MODE)) {
    // as many wavefronts as fit beside the pair table and the gradient tables, and the
    // receptor in LDS only when it costs no wavefront (C4 with per-wavefront tables: 8
    // wavefronts with the receptor read through L1 take 0.45 ms, 6 with it in LDS 0.48 ms)
    int w_lds = wpb, w_glob = wpb;
    while (w_lds > 1 && fixed + w_lds * per_wave + rec_bytes > d.lds_per_block) --w_lds;
    while (w_glob > 1 && fixed + w_glob * per_wave > d.lds_per_block) --w_glob;
    if (rec_lds && w_lds == w_glob && fixed + w_lds * per_wave + rec_bytes <= d.lds_per_block) {
      wpb = w_lds;
    } else {
      rec_lds = false;
      wpb = w_glob;
    }
  }
  const size_t base = fixed + wpb * per_wave;
  if (base > d.lds_per_block) return fail(ATJ_ERR_UNSUPPORTED, "pair table does not fit in LDS");
  const int blocks = launch && launch->blocks > 0 ? launch->blocks : 0;
  const size_t with_rec = base + rec_bytes;
  if (rec_lds && with_rec <= d.lds_per_block)
    return launch_flags<MODE, true>(a, wpb, with_rec, blocks, d, stream);
  if (a.nrec == 0) return fail(ATJ_ERR_INVALID, "global receptor path needs at least one bead");
  return launch_flags<MODE, false>(a, wpb, base, blocks, d, stream);
}

// Scratch memory of the allocation-free entry point (atj_minimize_noalloc): the work
// counter of its persistent launch lives here instead of in a caller workspace.  One copy per
// GPU, held until the process ends.  Host code never uses the array's name as an address:
// each device's address comes from hipGetSymbolAddress on that device's first call.  Every
// call on a GPU shares that GPU's copy, so calls that overlap on one GPU (on different
// streams) race on the counter.
constexpr size_t kScratchBytes = 256;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) int scratch[kScratchBytes / sizeof(int)];

std::mutex g_scratch_mu;
int* g_scratch_dev[kMaxDevices] = {};

// Device address of the scratch array on the current device (looked up once per device).
int scratch_address(int** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(ATJ_ERR_UNSUPPORTED, "device index beyond 64");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(ATJ_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_dev[dev] = static_cast<int*>(p);
  }
  *out = g_scratch_dev[dev];
  return ATJ_OK;
}

}  // namespace

extern "C" {

const char* atj_last_error(void) { return g_err.c_str(); }

const char* atj_version(void) { return "attract_amd 0.1 gfx950"; }

int atj_energrad(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec, const AtjJob* jobs_dev,
                 int32_t njobs, const int32_t* pose_job, const int32_t* pose_order, const float* poses, int32_t N,
                 float* energy, float* dEdM, uint32_t flags, const AtjLaunch* launch, int32_t* counter,
                 unsigned long long* stats, void* stream) {
  KArgs a = {};
  int rc = fill_common(a, grid, ff, rec, flags);
  if (rc) return rc;
  if (N < 0 || njobs < 1) return fail(ATJ_ERR_INVALID, "N < 0 or no job");
  if (N == 0) return ATJ_OK;
  if (!poses || !energy) return fail(ATJ_ERR_INVALID, "poses / energy NULL");
  if ((flags & ATJ_GRAD_QT) && !(flags & ATJ_POSE_QT))
    return fail(ATJ_ERR_INVALID, "ATJ_GRAD_QT needs quaternion poses (ATJ_POSE_QT)");
  a.jobs = jobs_dev;
  a.pose_job = pose_job;
  a.pose_order = pose_order;
  a.N = N;
  a.poses = poses;
  a.energy = energy;
  a.dEdM = dEdM;
  a.counter = counter;
  a.stats = stats;
  return launch_pose_kernel<kEnergrad>(a, launch, (hipStream_t)stream);
}

int atj_minimize(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec, const AtjJob* jobs_dev,
                 int32_t njobs, const int32_t* pose_job, const int32_t* pose_order, int32_t N,
                 const AtjPoseState* state, int32_t nsteps_launch, uint32_t flags, const AtjLaunch* launch,
                 int32_t* counter, unsigned long long* stats, void* stream) {
  KArgs a = {};
  int rc = fill_common(a, grid, ff, rec, flags & ~ATJ_POSE_QT);
  if (rc) return rc;
  if (N < 0 || njobs < 1 || nsteps_launch < 0) return fail(ATJ_ERR_INVALID, "N < 0, no job or nsteps < 0");
  if (N == 0 || nsteps_launch == 0) return ATJ_OK;
  if (!state || !state->pose || !state->m || !state->v || !state->count || !state->energy)
    return fail(ATJ_ERR_INVALID, "pose state array NULL");
  a.jobs = jobs_dev;
  a.pose_job = pose_job;
  a.pose_order = pose_order;
  a.N = N;
  a.pose7 = state->pose;
  a.mom1 = state->m;
  a.mom2 = state->v;
  a.count = state->count;
  a.elast = state->energy;
  a.nsteps_launch = nsteps_launch;
  a.counter = counter;
  a.stats = stats;
  return launch_pose_kernel<kMinimize>(a, launch, (hipStream_t)stream);
}

int atj_minimize_noalloc(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec,
                         const AtjJob* jobs_dev, int32_t njobs, const int32_t* pose_job, const int32_t* pose_order,
                         int32_t N, const AtjPoseState* state, int32_t nsteps_launch, uint32_t flags,
                         const AtjLaunch* launch, unsigned long long* stats, void* stream) {
  int* counter = nullptr;
  int rc = scratch_address(&counter);
  if (rc) return rc;
  // the work counter takes the first int of the scratch array; the launch resets it with a
  // fill kernel (no memset node, so the call may be captured in a graph)
  return atj_minimize(grid, ff, rec, jobs_dev, njobs, pose_job, pose_order, N, state, nsteps_launch, flags, launch,
                      counter, stats, stream);
}

int atj_energrad_adjoint(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec,
                         const AtjJob* jobs_dev, int32_t njobs, const int32_t* pose_job, const float* poses,
                         int32_t N, const float* weights, float* energy, float* dEdM, float* partial,
                         uint32_t flags, const AtjLaunch* launch, int32_t* counter, void* stream) {
  KArgs a = {};
  int rc = fill_common(a, grid, ff, rec, flags);
  if (rc) return rc;
  if (N < 0 || njobs < 1) return fail(ATJ_ERR_INVALID, "N < 0 or no job");
  if (!launch || launch->blocks <= 0) return fail(ATJ_ERR_INVALID, "adjoint needs launch->blocks > 0");
  if (!poses || !weights || !partial) return fail(ATJ_ERR_INVALID, "NULL argument");
  if (dEdM && !energy) return fail(ATJ_ERR_INVALID, "dEdM needs energy");
  a.jobs = jobs_dev;
  a.pose_job = pose_job;
  a.N = N;
  a.poses = poses;
  hipStream_t st = (hipStream_t)stream;
  if (energy && N > 0) {  // energies and dE/dM: the energrad launch
    KArgs e = a;
    e.energy = energy;
    e.dEdM = dEdM;
    e.counter = counter;
    AtjLaunch le = *launch;
    le.blocks = 0;
    rc = launch_pose_kernel<kEnergrad>(e, &le, st);
    if (rc) return rc;
  }
  // the pair-table gradient: a pass of its own over the neighbour-list pairs (the potential
  // grids do not depend on the pair table), with partial sums per block
  a.flags = flags & ~(uint32_t)ATJ_TERM_GRID;
  a.weights = weights;
  a.partial = partial;
  if (!(a.flags & ATJ_TERM_NB) || N == 0) {
    const long long n = (long long)launch->blocks * 3 * ff->nrt * ff->nlt;
    hipLaunchKernelGGL(zero_floats, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, partial, n);
    HIPCHK(hipGetLastError());
    return ATJ_OK;
  }
  return launch_pose_kernel<kAdjoint>(a, launch, st);
}

int atj_reduce_partials(const float* partial, int32_t blocks, int32_t n, float* out, void* stream) {
  if (!partial || !out || blocks <= 0 || n <= 0) return fail(ATJ_ERR_INVALID, "bad reduce arguments");
  hipLaunchKernelGGL(reduce_partials_kernel, dim3((n + kReduceCols - 1) / kReduceCols),
                     dim3(kReduceCols * kReduceRows), 0, (hipStream_t)stream, partial, blocks, n, out);
  HIPCHK(hipGetLastError());
  return ATJ_OK;
}

int atj_bead_terms(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec, const AtjLigand* ligand,
                   const float* poses, int32_t N, float* grid4, int32_t* voxel, float* nbene, float* nbgrad,
                   uint32_t flags, void* stream) {
  KArgs a = {};
  int rc = fill_common(a, grid, ff, rec, flags);
  if (rc) return rc;
  if (!ligand || !poses) return fail(ATJ_ERR_INVALID, "ligand / poses NULL");
  a.N = N;
  a.poses = poses;
  const long long tot = (long long)N * ligand->nbeads;
  if (tot == 0) return ATJ_OK;
  hipLaunchKernelGGL(bead_terms_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a,
                     reinterpret_cast<const float4*>(ligand->beads), ligand->nbeads, nullptr, nullptr, grid4, voxel,
                     nbene, nbgrad);
  HIPCHK(hipGetLastError());
  return ATJ_OK;
}

int atj_point_terms(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec, const float* coords,
                    const int32_t* types, int32_t M, float* grid4, int32_t* voxel, float* nbene, float* nbgrad,
                    uint32_t flags, void* stream) {
  KArgs a = {};
  int rc = fill_common(a, grid, ff, rec, flags);
  if (rc) return rc;
  if (M < 0) return fail(ATJ_ERR_INVALID, "M < 0");
  if (M == 0) return ATJ_OK;
  if (!coords || !types) return fail(ATJ_ERR_INVALID, "coords / types NULL");
  a.N = M;
  hipLaunchKernelGGL(bead_terms_kernel, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a,
                     nullptr, 1, coords, types, grid4, voxel, nbene, nbgrad);
  HIPCHK(hipGetLastError());
  return ATJ_OK;
}

int atj_ligand_rmsd(const float* poses, int32_t N, const float* lig, const float* ref, int32_t nb, float* out,
                    void* stream) {
  if (N < 0 || nb <= 0) return fail(ATJ_ERR_INVALID, "N < 0 or no beads");
  if (N == 0) return ATJ_OK;
  if (!poses || !lig || !ref || !out) return fail(ATJ_ERR_INVALID, "NULL argument");
  constexpr int kWavesPerBlock = 4;
  hipLaunchKernelGGL(ligand_rmsd_kernel, dim3((N + kWavesPerBlock - 1) / kWavesPerBlock), dim3(kWave * kWavesPerBlock),
                     0, (hipStream_t)stream, poses, N, lig, ref, nb, out);
  HIPCHK(hipGetLastError());
  return ATJ_OK;
}

int atj_optimizer_step(const float* dEdM, float* energy, int32_t N, const AtjPoseState* state,
                       const AtjOptimizer* opt, void* stream) {
  if (N <= 0) return N == 0 ? ATJ_OK : fail(ATJ_ERR_INVALID, "N < 0");
  if (!dEdM || !state || !opt || !state->pose || !state->m || !state->v || !state->count)
    return fail(ATJ_ERR_INVALID, "NULL argument");
  if (opt->kind != ATJ_OPT_ADAM && opt->kind != ATJ_OPT_ADABELIEF) return fail(ATJ_ERR_INVALID, "unknown optimizer");
  hipLaunchKernelGGL(optimizer_step_kernel, dim3((N + 255) / 256), dim3(256), 0, (hipStream_t)stream, dEdM, energy,
                     N, state->pose, state->m, state->v, state->count, *opt);
  HIPCHK(hipGetLastError());
  return ATJ_OK;
}

int atj_grid_build_potentials(const AtjGrid* geom, const AtjForceField* ff, const AtjReceptor* rec,
                              const int32_t* lig_fftype_dev, float cutoff, float* inner_out, float* outer_out,
                              void* stream) {
  GridView G;
  int rc = make_view(geom, G, false);
  if (rc) return rc;
  if (!ff || !rec || !lig_fftype_dev || !inner_out) return fail(ATJ_ERR_INVALID, "NULL argument");
  if (geom->ntypes <= 0 || geom->ntypes > 65535) return fail(ATJ_ERR_INVALID, "bad ntypes");
  if (ff->nrt <= 0 || ff->nlt <= 0 || !ff->rc || !ff->ac || !ff->ivor || !ff->emin || !ff->rmin2)
    return fail(ATJ_ERR_INVALID, "empty pair table or NULL pair-table array");
  if (rec->nrec < 0 || (rec->nrec > 0 && !rec->beads)) return fail(ATJ_ERR_INVALID, "receptor beads NULL");
  if ((long long)ff->nrt * ff->nlt > 4096) return fail(ATJ_ERR_UNSUPPORTED, "pair table larger than 4096 entries");
  const float c2 = cutoff * cutoff;
  const float4* rb = reinterpret_cast<const float4*>(rec->beads);
  if (rec->nrec == 0) return fail(ATJ_ERR_INVALID, "receptor has no beads");
  const size_t lds = build_lds_bytes(ff->nrt);
  auto kern = build_potentials_kernel;
  if (lds > 65536)
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
  const unsigned groups = (unsigned)((geom->ntypes + kBuildTypes - 1) / kBuildTypes);
  long long nv = (long long)G.dx * G.dy * G.dz;
  hipLaunchKernelGGL(kern, dim3((unsigned)((nv + kBuildThreads - 1) / kBuildThreads), groups), dim3(kBuildThreads),
                     lds, (hipStream_t)stream, G, rb, rec->nrec, ff->rc, ff->ac, ff->ivor, ff->emin, ff->rmin2,
                     ff->nrt, ff->nlt, lig_fftype_dev, geom->ntypes, c2, inner_out, G.dx, G.dy, G.dz, 0);
  HIPCHK(hipGetLastError());
  if (outer_out) {
    nv = (long long)G.dx2 * G.dy2 * G.dz2;
    hipLaunchKernelGGL(kern, dim3((unsigned)((nv + kBuildThreads - 1) / kBuildThreads), groups),
                       dim3(kBuildThreads), lds, (hipStream_t)stream, G, rb, rec->nrec, ff->rc, ff->ac, ff->ivor,
                       ff->emin, ff->rmin2, ff->nrt, ff->nlt, lig_fftype_dev, geom->ntypes, c2, outer_out, G.dx2,
                       G.dy2, G.dz2, 1);
    HIPCHK(hipGetLastError());
  }
  return ATJ_OK;
}

int atj_grid_shell_lists(const AtjGrid* grid, const AtjReceptor* rec, int32_t nshell, uint16_t* nblist_out,
                         int32_t* nbshell_out, void* stream) {
  if (nshell < 1 || nshell > kMaxShell) return fail(ATJ_ERR_INVALID, "nshell must be 1..6");
  if (!grid || grid->nshell != 0) return fail(ATJ_ERR_INVALID, "shell order is built from lists read whole (nshell 0)");
  GridView G;
  int rc = make_view(grid, G, true);
  if (rc) return rc;
  if (!rec || (rec->nrec > 0 && !rec->beads) || !nblist_out || !nbshell_out)
    return fail(ATJ_ERR_INVALID, "NULL argument");
  const long long npts = (long long)G.dx * G.dy * G.dz;
  const double h = (double)grid->spacing * std::sqrt(3.0) / (2.0 * nshell);
  const float r0 = (float)((double)grid->plateaudis + kShellMargin);
  hipLaunchKernelGGL(shell_lists_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, (hipStream_t)stream, G,
                     reinterpret_cast<const float4*>(rec->beads), rec->nrec, (int)nshell, r0, (float)h, (int)npts,
                     nblist_out, reinterpret_cast<int4*>(nbshell_out));
  HIPCHK(hipGetLastError());
  return ATJ_OK;
}

int atj_grid_count_neighbours(const AtjGrid* geom, const AtjReceptor* rec, float neighbourdis, int32_t* counts,
                              void* stream) {
  GridView G;
  int rc = make_view(geom, G, false);
  if (rc) return rc;
  if (!rec || !counts) return fail(ATJ_ERR_INVALID, "NULL argument");
  if (rec->nrec < 0 || (rec->nrec > 0 && !rec->beads)) return fail(ATJ_ERR_INVALID, "receptor beads NULL");
  hipLaunchKernelGGL(count_neighbours_kernel, dim3((unsigned)brick_count(G.dx, G.dy, G.dz)), dim3(kBrickThreads), 0,
                     (hipStream_t)stream, G, reinterpret_cast<const float4*>(rec->beads), rec->nrec, neighbourdis,
                     neighbourdis * neighbourdis, counts);
  HIPCHK(hipGetLastError());
  return ATJ_OK;
}

int atj_grid_fill_neighbours(const AtjGrid* geom, const AtjReceptor* rec, float neighbourdis, const int32_t* counts,
                             const int64_t* starts, int32_t* nbgrid, uint16_t* nblist, void* stream) {
  GridView G;
  int rc = make_view(geom, G, false);
  if (rc) return rc;
  if (!rec || !counts || !starts || !nbgrid || !nblist) return fail(ATJ_ERR_INVALID, "NULL argument");
  if (rec->nrec < 0 || (rec->nrec > 0 && !rec->beads)) return fail(ATJ_ERR_INVALID, "receptor beads NULL");
  hipLaunchKernelGGL(fill_neighbours_kernel, dim3((unsigned)brick_count(G.dx, G.dy, G.dz)), dim3(kBrickThreads), 0,
                     (hipStream_t)stream, G, reinterpret_cast<const float4*>(rec->beads), rec->nrec, neighbourdis,
                     neighbourdis * neighbourdis, counts, reinterpret_cast<const long long*>(starts),
                     reinterpret_cast<int2*>(nbgrid), nblist);
  HIPCHK(hipGetLastError());
  return ATJ_OK;
}

int atj_grid_pack_corners(const float* plain, int32_t ntypes, const int32_t dim[3], float* packed, void* stream) {
  if (!plain || !packed || !dim || ntypes <= 0) return fail(ATJ_ERR_INVALID, "NULL argument");
  for (int i = 0; i < 3; ++i)
    if (dim[i] < 2) return fail(ATJ_ERR_INVALID, "dim must be >= 2");
  const long long n = (long long)ntypes * (dim[0] - 1) * (dim[1] - 1) * (dim[2] - 1) * 8;
  hipLaunchKernelGGL(pack_corners_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, plain,
                     ntypes, dim[0], dim[1], dim[2], packed);
  HIPCHK(hipGetLastError());
  return ATJ_OK;
}

}  // extern "C"
