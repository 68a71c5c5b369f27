/*
 * attract_amd.h -- C-ABI of the MI355X (gfx950) ATTRACT scoring and rigid-body
 * minimization library (libattract_amd.so).
 *
 * Every buffer argument is DEVICE memory owned by the caller (hipMalloc, or a
 * torch tensor's data pointer); the library never allocates or frees caller
 * buffers.  All work is stream-ordered on the hipStream_t passed in (NULL = the
 * default stream).  Every entry point returns ATJ_OK (0) or an error code; the
 * message of the last error on the calling thread is atj_last_error().  No C++
 * exception crosses this boundary.
 *
 * Conventions follow the reference scripts (sjdv1982/attract-jax):
 *  - a pose is a 4x4 row-vector matrix M: x' = [x, y, z, 1] . M, i.e.
 *    x'_l = sum_k x_k M[k][l] + M[3][l]                (test-grid-energrad.py:299-301)
 *  - or a quaternion (x, y, z, w; scipy scalar-last) + translation, with
 *    M[0:3][0:3] = as_matrix(q) (unnormalised polynomial) (test4.py:396-401)
 *  - grid gradients are stored with the ATTRACT sign (force); the energy
 *    gradient of the grid term is minus their trilinear interpolation
 *    (test-grid-energrad.py:232-239).
 */
#ifndef ATTRACT_AMD_H
#define ATTRACT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATJ_OK 0
#define ATJ_ERR_INVALID 1      /* bad argument (shape, size, null pointer) */
#define ATJ_ERR_HIP 2          /* HIP runtime error */
#define ATJ_ERR_UNSUPPORTED 3  /* valid but unsupported configuration */

/* Potential-grid memory layouts (AtjGrid.layout). */
#define ATJ_LAYOUT_PLAIN 0   /* float [ntypes][dx][dy][dz][4]  (reference layout, read_grid.py:81) */
#define ATJ_LAYOUT_CORNER 1  /* float [ntypes][dx-1][dy-1][dz-1][8][4]: the 8 trilinear corners of a
                                voxel cell packed into one 128-byte record, corner c = 4*cx + 2*cy + cz */

/* Flags (bitwise or). */
#define ATJ_TERM_GRID 0x1u    /* grid trilinear term (inner, else outer grid) */
#define ATJ_TERM_NB 0x2u      /* neighbour-list 8-6 pair term */
#define ATJ_NB_PLATEAU 0x4u   /* subtract the plateau energy / force-shift the pair term
                                 (test-grid-energrad.py:124-147); without it the pair term is
                                 plain nonbon (test-nb-energrad.py:142-145) */
#define ATJ_GRAD_EXACT 0x8u   /* gradient = exact derivative of the energy (finite-difference
                                 checkable) instead of the ATTRACT convention */
#define ATJ_POSE_QT 0x10u     /* poses are 7 floats (qx, qy, qz, qw, tx, ty, tz), not 16 */
#define ATJ_GRAD_QT 0x20u     /* atj_energrad with ATJ_POSE_QT: dEdM[p] receives the gradient
                                 w.r.t. the pose's 7 parameters (dE/dq through as_matrix,
                                 dE/dt), then 9 zeros (test4.py:415 differentiates the
                                 quaternion form) */

/* Device receptor grid.  Geometry follows read_grid.py:35-48. */
typedef struct AtjGrid {
  float origin[3];
  float spacing;          /* gridspacing */
  int32_t dim[3];         /* inner grid points */
  int32_t dim2[3];        /* outer grid points (ignored when outer == NULL) */
  float gridextension;
  float plateaudis;
  float plateaudis2;      /* plateaudis**2 rounded once to float32 (test-grid-energrad.py:113) */
  int32_t ntypes;        /* number of potential grids (ligand grid types) */
  int32_t layout;         /* ATJ_LAYOUT_* of inner and outer */
  int32_t nshell;         /* 0: a bead reads its point's whole list (nbgrid).  1..6: the lists are in
                             shell order (atj_grid_shell_lists) and a bead of distance shell k
                             from point v reads nbshell[v]'s count k from nbshell[v]'s start */
  int32_t reserved;
  const float* inner;     /* potentials + gradients, NULL for a neighbour-only grid */
  const float* outer;     /* NULL = no outer grid */
  const int32_t* nbgrid;  /* [dx*dy*dz][2] = (count, start): 0-based start into nblist,
                             start % 8 == 0 and start >= 16 (lists are 16-byte aligned; 64-byte
                             alignment, start % 32 == 0, saves memory lines) */
  const uint16_t* nblist; /* receptor bead indices; each list padded with 0xFFFF to a multiple
                             of 16 entries; entries 0..15 are 0xFFFF */
  const int32_t* nbshell; /* [dx*dy*dz][4] when nshell > 0, else unused: (start, c0 | c1 << 16,
                             c2 | c3 << 16, c4 | c5 << 16), ck the list length shell k reads */
} AtjGrid;

/* ATTRACT pair table, compacted to receptor types x ligand types
 * (test-grid-energrad.py:67-79).  All arrays are float32 [nrt][nlt], row-major.  The pose
 * kernels stage it into each block's LDS as 28 bytes per slot, nrt x (nlt | 1) slots, next
 * to per-wavefront scratch (about 4.3 KB per wavefront for the minimizer, 3 KB for energrad;
 * the adjoint adds 12 bytes per slot).  A table that does not fit in 160 KB beside them
 * returns ATJ_ERR_UNSUPPORTED: at 16 wavefronts per block the minimizer takes up to about
 * 3,400 slots (57 receptor x 57 ligand types), fewer wavefronts per block leave room for more. */
typedef struct AtjForceField {
  const float* rc;
  const float* ac;
  const float* ivor;   /* +1 or -1 (ATTRACT ipon); any other value yields NaN energies */
  const float* emin;   /* -27 ac^4 / (256 rc^3) */
  const float* rmin2;  /* 4 rc / (3 ac) */
  int32_t nrt;
  int32_t nlt;
} AtjForceField;

/* Receptor beads: float4 [nrec] = (x, y, z, ff type as int32 bit pattern).  A type of
 * nrt or more is read as row nrt - 1 of the pair table. */
typedef struct AtjReceptor {
  const float* beads;
  int32_t nrec;
} AtjReceptor;

/* One rigid ligand: float4 [nbeads] = (x, y, z, int32 bit pattern of
 * grid_type | ff_type << 16).  grid_type indexes the potential grids
 * (lig_atomtype_pos), ff_type the ligand column of the pair table; types beyond
 * them are read as the last grid / column. */
typedef struct AtjLigand {
  const float* beads;
  int32_t nbeads;
  int32_t reserved;
} AtjLigand;

/* Optimizer of the rigid-body minimization (per job).  Both update the parameter pytree
 * (quaternion (x, y, z, w), translation) like optax, then renormalise the quaternion
 * (Rotation.from_quat(as_quat())).
 *   ATJ_OPT_ADAM       optax.adam (test4.py:447-453): m = (1-b1) g + b1 m, v = (1-b2) g^2 + b2 v
 *   ATJ_OPT_ADABELIEF  optax.adabelief (test5.py:451-457): s = (1-b2) (g - m_prev)^2 + b2 s + eps_root
 * with the update -lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).  gravity > 0 adds
 * gravity * |t|^2 to each pose's energy and 2 gravity t to its translation gradient
 * (test5.py:418-420, weight 0.1 in the stages that use it). */
#define ATJ_OPT_ADAM 0
#define ATJ_OPT_ADABELIEF 1
typedef struct AtjOptimizer {
  int32_t kind;     /* ATJ_OPT_* */
  float lr_rot;     /* learning rate of the quaternion */
  float lr_trans;   /* learning rate of the translation (the reference uses one lr for both) */
  float beta1, beta2, eps;
  float eps_root;   /* adabelief only (optax default 1e-16) */
  float gravity;    /* restraint weight, 0 = off */
} AtjOptimizer;

/* A group of poses sharing one ligand (one docking job, or one ligand of a batch). */
typedef struct AtjJob {
  AtjLigand ligand;
  int32_t pose_begin;  /* first pose slot of the job */
  int32_t pose_count;
  int32_t nsteps;      /* total optimizer steps of the job (minimize only) */
  int32_t reserved;
  AtjOptimizer opt;    /* minimize only */
} AtjJob;

/* Optional per-launch counters (device uint64[4]): beads inside a grid,
 * beads inside the inner grid, listed neighbour entries (sum of K_v),
 * evaluated pairs (dsq < plateaudis^2).  Accumulated, not reset.  A library built with
 * ATJ_PHASE_CLOCKS defined also accumulates, in stats[ATJ_NSTATS + i], the shader-clock
 * cycles its wavefronts spent in phase i (ATJ_NPHASES: claims and set-up, bead phase,
 * segment-size scans, pair phase, reductions, optimizer update); stats then needs
 * ATJ_NSTATS + ATJ_NPHASES entries. */
#define ATJ_NSTATS 4
#define ATJ_NPHASES 6

/* Pose-state arrays of the minimizer (all device, indexed by pose slot). */
typedef struct AtjPoseState {
  float* pose;      /* [N][7] = quaternion (x, y, z, w), translation (x, y, z) */
  float* m;         /* [N][7] first moment (quat, trans) */
  float* v;         /* [N][7] second moment (adam) / belief (adabelief) */
  int32_t* count;   /* [N] optimizer steps done (zero them to re-initialise the optimizer) */
  float* energy;    /* [N] energy of the last evaluated pose (gravity term included) */
} AtjPoseState;

/* Launch configuration of the pose kernels (chosen by the autotuner). */
#define ATJ_LAUNCH_REC_GLOBAL 0x1u  /* read receptor beads through the vector-memory (L1) path, not LDS */
#define ATJ_LAUNCH_CLAIM(n) ((uint32_t)(n) << 4)  /* poses per claim, n = 1, 2, 4 or 8 (bits 4-7;
                                                    0 = chosen from N and the wavefronts) */
#define ATJ_LAUNCH_CLAIM_MASK 0xF0u
#define ATJ_LAUNCH_TAIL(n) ((uint32_t)(n) << 8)  /* poses per claim in the tail, n = 1, 2 or 4 below
                                                   the claim size (bits 8-11; 0 = no tail) */
#define ATJ_LAUNCH_TAIL_MASK 0xF00u
#define ATJ_LAUNCH_TAIL_ROUNDS(r) ((uint32_t)(r) << 12)  /* length of the tail: the last r rounds of
                                                           claims, r = 1..15 (bits 12-15; 0 = 1) */
#define ATJ_LAUNCH_TAIL_ROUNDS_MASK 0xF000u
#define ATJ_LAUNCH_ADJOINT_WAVE_TABLES 0x10000u  /* adjoint: per-wavefront float tables instead of the
                                                   block's fixed-point table (the form used anyway
                                                   when the fixed-point table does not fit in LDS) */
typedef struct AtjLaunch {
  int32_t waves_per_block;  /* 1..16 (0 = default 4) */
  int32_t blocks;           /* 0 = as many as the device holds, fewer for small launches */
  uint32_t options;         /* ATJ_LAUNCH_* */
} AtjLaunch;
/* Poses are claimed 1, 2, 4 or 8 at a time (chosen from N and the wavefronts launched, or set
 * by ATJ_LAUNCH_CLAIM).  With ATJ_LAUNCH_TAIL the poses of the last rounds of claims (one round:
 * a claim for every wavefront the device holds at once) are claimed in smaller claims, which
 * even out the end of a persistent launch while the rest keeps the larger batches.  In
 * atj_minimize the first claim of each wavefront in the blocks the device holds at once is
 * fixed by its index, further claims come from the work counter.  atj_energrad and
 * atj_energrad_adjoint take their claims in a fixed order (wavefront g of G: claims g, g + G,
 * ...) and use no counter.  A pose's result does not depend on any of these options. */

const char* atj_last_error(void);
const char* atj_version(void);

/* Energy and pose gradient of N poses (reference: main + jax.value_and_grad,
 * test-grid-energrad.py:298-317; quaternion form test4.py:396-401,415).
 * poses: [N][16] (row-major 4x4) or [N][7] with ATJ_POSE_QT.
 * jobs/njobs give the ligand of each pose range; pose_order (nullable) is the
 * processing order of pose slots (a permutation of [0, N)).
 * energy: [N]; dEdM: [N][16] = dE/dM (nullable; with ATJ_GRAD_QT dE/d(q, t), see the flag);
 * counter: unused (claims follow a fixed order; nullable, kept for the signature); stats:
 * nullable. */
int atj_energrad(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec,
                 const AtjJob* jobs_dev, int32_t njobs, const int32_t* pose_job,
                 const int32_t* pose_order, const float* poses, int32_t N, float* energy,
                 float* dEdM, uint32_t flags, const AtjLaunch* launch, int32_t* counter,
                 unsigned long long* stats, void* stream);

/* Per-bead terms, for parity checks and for potential_energrad
 * (test-grid-energrad.py:214-225) / generate_nb_table (:246-260).
 * Outputs per (pose, bead) (all nullable):
 *   grid4 [N][nb][4]  interpolated (E, gx, gy, gz) of the grid term (stored sign)
 *   voxel [N][nb]     flat inner voxel index of the neighbour lookup, -1 if outside
 *   nbene [N][nb]     neighbour-list energy of the bead
 *   nbgrad [N][nb][3] neighbour-list dE/dx' of the bead */
int atj_bead_terms(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec,
                   const AtjLigand* ligand, const float* poses, int32_t N, float* grid4,
                   int32_t* voxel, float* nbene, float* nbgrad, uint32_t flags, void* stream);

/* Per-point terms at explicit coordinates (the reference's per-bead functions take
 * transformed coordinates all_coors_lig).  coords [M][3], types [M] = grid_type |
 * ff_type << 16; outputs as atj_bead_terms with [M] leading dimension. */
int atj_point_terms(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec, const float* coords,
                    const int32_t* types, int32_t M, float* grid4, int32_t* voxel, float* nbene, float* nbgrad,
                    uint32_t flags, void* stream);

/* Batched rigid-body minimization: every pose of every job runs
 * min(nsteps_launch, job.nsteps - count) steps of its job's optimizer inside ONE
 * persistent launch (reference loops: test4.py:445-455 adam, test5.py:450-458 adabelief
 * + gravity; quaternion renormalised after every step).  state.energy receives the
 * energy evaluated in the last step. */
int atj_minimize(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec,
                 const AtjJob* jobs_dev, int32_t njobs, const int32_t* pose_job,
                 const int32_t* pose_order, int32_t N, const AtjPoseState* state,
                 int32_t nsteps_launch, uint32_t flags, const AtjLaunch* launch,
                 int32_t* counter, unsigned long long* stats, void* stream);

/* atj_minimize without the work-counter workspace: for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops that forbid hipMalloc).  No
 * counterpart in the reference.  The counter lives in a 256-byte scratch array the library
 * holds on each GPU until the process ends; the call allocates nothing and needs no size
 * limit beyond atj_minimize's (the scratch does not grow with N).  Every call on one GPU
 * shares that GPU's scratch: calls that overlap on one GPU (different streams) race and
 * must be serialised by the caller.  Arguments and results as atj_minimize. */
int atj_minimize_noalloc(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec,
                         const AtjJob* jobs_dev, int32_t njobs, const int32_t* pose_job,
                         const int32_t* pose_order, int32_t N, const AtjPoseState* state,
                         int32_t nsteps_launch, uint32_t flags, const AtjLaunch* launch,
                         unsigned long long* stats, void* stream);

/* One optimizer step of N poses from dE/dM (the plain one-launch-per-step form):
 * quaternion chain rule through as_matrix, the gravity term (energy[p] += gravity |t|^2
 * when energy is not NULL), the optimizer update, renormalisation.  Same arithmetic as
 * one step of atj_minimize. */
int atj_optimizer_step(const float* dEdM, float* energy, int32_t N, const AtjPoseState* state,
                       const AtjOptimizer* opt, void* stream);

/* Ligand RMSD of N poses (reference: test-rmsd.py:13-17): poses [N][16] row-vector
 * matrices applied to lig [nb][3], deviations from ref [nb][3]; out [N] =
 * sqrt(sum_j |x'_j - ref_j|^2 / nb).  All device pointers. */
int atj_ligand_rmsd(const float* poses, int32_t N, const float* lig, const float* ref, int32_t nb, float* out,
                    void* stream);

/* Adjoint of the pair term w.r.t. the pair table: for weights w[N]
 * (upstream dL/dE_pose) accumulates sum_pose w * dE_pose/d(rc, ac, emin) into
 * partial[blocks][3][nrt*nlt] (one slab per block, written, not accumulated), recomputing
 * the pairs (nothing per pair is stored).  Reduce the slabs with atj_reduce_partials.
 * Requires blocks = launch->blocks > 0.  The sums are reproducible bit for bit for a given
 * number of blocks: a block adds each pair's terms, rounded to multiples of 2^-43, into two-word
 * 64-bit fixed-point sums (exact, in any order, for terms below 2^19 in magnitude; larger terms,
 * infinities and NaNs go to float sums), and the slabs are added in a fixed order.  Pair tables
 * too large for the block's fixed-point table in LDS (or ATJ_LAUNCH_ADJOINT_WAVE_TABLES) take
 * per-wavefront float tables: each wavefront takes its poses in a fixed order, and the tables
 * are added in a fixed order (reproducible for a given launch shape).  energy (nullable) and dEdM (nullable, needs energy)
 * receive what atj_energrad writes, by an energrad launch before the gradient pass; counter is
 * unused and may be NULL. */
int atj_energrad_adjoint(const AtjGrid* grid, const AtjForceField* ff, const AtjReceptor* rec,
                         const AtjJob* jobs_dev, int32_t njobs, const int32_t* pose_job,
                         const float* poses, int32_t N, const float* weights, float* energy,
                         float* dEdM, float* partial, uint32_t flags, const AtjLaunch* launch,
                         int32_t* counter, void* stream);

/* out[i] = sum_b partial[b][i] for i < n, in block order (deterministic). */
int atj_reduce_partials(const float* partial, int32_t blocks, int32_t n, float* out, void* stream);

/* Grid construction on the device (semantics of test-calc-grid-energy.py:132-196
 * plus a force channel):
 * potentials of ligand grid type t at every inner / outer grid point, plain layout.
 * lig_fftype[ntypes] gives the pair-table column of each grid type. */
int atj_grid_build_potentials(const AtjGrid* geom, const AtjForceField* ff,
                              const AtjReceptor* rec, const int32_t* lig_fftype_dev,
                              float cutoff, float* inner_out, float* outer_out, void* stream);
/* Shell order of the neighbour lists (no counterpart in the reference: a reordering of its
 * lists, test-grid-energrad.py:246-253, that lets a bead skip entries beyond the plateau).
 * Bead x in point v's cell (the beads whose list is v's) lies within spacing * sqrt(3) / 2
 * of v; split that range into nshell shells of width h = spacing * sqrt(3) / (2 nshell), and
 * let k = min(floor(|x - v| / h), nshell - 1) (nshell 1..6).  A receptor bead r of v's list can be within
 * plateaudis of x only if |r - v| < plateaudis + (k + 1) h.  This call writes nblist_out, a
 * copy of grid->nblist (same size; the caller copies it), with each list reordered into
 * groups of increasing |r - v| (group 0: |r - v| < r0 = plateaudis + 0.001; group j:
 * ceil((|r - v| - r0) / h) = j, clamped to nshell + 1; receptor order within a group), and
 * nbshell_out [dx*dy*dz][4] = (start, c0 | c1 << 16, c2 | c3 << 16, c4 | c5 << 16), ck =
 * the entries of groups 0..k+1 of the list.  The pairs
 * within plateaudis of every bead are those of the full lists (the 0.001 A covers float
 * rounding); the evaluated pair set is unchanged, the summation order is not.  grid->nshell
 * must be 0; the result is used with nblist = nblist_out, nbshell = nbshell_out, nshell. */
int atj_grid_shell_lists(const AtjGrid* grid, const AtjReceptor* rec, int32_t nshell, uint16_t* nblist_out,
                         int32_t* nbshell_out, void* stream);

/* Neighbour lists: receptor beads with |p - r| < neighbourdis of every inner grid point p.
 * Pass 1 writes counts [dx*dy*dz]; the caller turns round-up-to-16 counts into 0-based
 * starts (and fills entries 0..15 with 0xFFFF); pass 2 writes nbgrid [V][2] and nblist
 * (ascending receptor index, 0xFFFF padding to a multiple of 16 entries). */
int atj_grid_count_neighbours(const AtjGrid* geom, const AtjReceptor* rec, float neighbourdis,
                              int32_t* counts, void* stream);
int atj_grid_fill_neighbours(const AtjGrid* geom, const AtjReceptor* rec, float neighbourdis,
                             const int32_t* counts, const int64_t* starts, int32_t* nbgrid,
                             uint16_t* nblist, void* stream);
/* Plain [t][x][y][z][4] -> corner-packed [t][x-1][y-1][z-1][8][4]. */
int atj_grid_pack_corners(const float* plain, int32_t ntypes, const int32_t dim[3],
                          float* packed, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ATTRACT_AMD_H */
