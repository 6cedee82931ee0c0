/*
 * sphx.h - C ABI of the MI355X-native SPH-EXA single-rank hydro hot path.
 *
 * Every entry point takes plain device pointers and sizes, runs asynchronously
 * on the stream held by its context, and returns an int status (SPHX_OK = 0).
 * A failing call leaves a message retrievable with sphx_last_error(ctx).
 * Nothing here calls exit(); nothing uses process-global device state, so two
 * contexts on one device are independent (calls that overlap on ONE context race).
 *
 * Citations name the reference C++ interface each entry point replaces
 * (paths relative to the SPH-EXA source tree).
 */
#ifndef SPHX_H
#define SPHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
enum
{
    SPHX_OK          = 0,
    SPHX_ERR_HIP     = 1, /* HIP runtime error (message has the HIP error string)          */
    SPHX_ERR_ARG     = 2, /* invalid argument (null pointer, bad size, bad precision)      */
    SPHX_ERR_STACK   = 3, /* traversal queue overflow (domain/.../find_neighbors.cuh:252)  */
    SPHX_ERR_NGMAX   = 4, /* ng0 > ngmax (sph/include/sph/find_neighbors.hpp:49)           */
    SPHX_ERR_STATE   = 5, /* SPH loop called without matching neighbor lists in context   */
    SPHX_ERR_NOMEM   = 6, /* device allocation failed                                      */
    SPHX_ERR_NODEV   = 7  /* no HIP device                                                 */
};

/* cstone::BoundaryType, domain/include/cstone/sfc/box.hpp:80-85 */
enum
{
    SPHX_BC_OPEN     = 0,
    SPHX_BC_PERIODIC = 1,
    SPHX_BC_FIXED    = 2
};

/* element type of the hydro fields (everything except x, y, z and keys) */
enum
{
    SPHX_FP64  = 0, /* double fields, FP64 arithmetic everywhere                         */
    SPHX_MIXED = 1  /* float fields; FP64 positions and pair offsets, FP32 weights and
                       Kahan-compensated FP32 neighbor sums                               */
};

/* cstone::Box<double>, domain/include/cstone/sfc/box.hpp:94-174 (lengths and inverse
 * lengths are derived inside the library exactly as the reference constructor does) */
typedef struct sphx_box
{
    double  xmin, xmax, ymin, ymax, zmin, zmax;
    int32_t bx, by, bz; /* SPHX_BC_* */
    int32_t reserved;
} sphx_box;

/* Internal octree + neighbor-search view; mirrors cstone::OctreeView
 * (domain/include/cstone/tree/octree.hpp:224-245) and cstone::OctreeNsView (:259-281).
 * All pointers are device pointers, allocated by the caller with the sizes given
 * by sphx_octree_sizes(). */
typedef struct sphx_octree
{
    int32_t   numLeafNodes;
    int32_t   numInternalNodes; /* (numLeafNodes - 1) / 7                                */
    int32_t   numNodes;         /* numLeafNodes + numInternalNodes                        */
    int32_t   reserved;
    uint64_t* prefixes;         /* [numNodes] Warren-Salmon placeholder-bit keys          */
    int32_t*  childOffsets;     /* [numNodes + 1], 0 = leaf                               */
    int32_t*  parents;          /* [max(1, (numNodes - 1) / 8)]                           */
    int32_t*  levelRange;       /* [23] first node index per level, device               */
    int32_t*  internalToLeaf;   /* [numNodes]                                             */
    int32_t*  leafToInternal;   /* [numNodes]                                             */
    uint64_t* leaves;           /* [numLeafNodes + 1] cornerstone leaf keys               */
    uint32_t* layout;           /* [numLeafNodes + 1] first particle of each leaf         */
    double*   centers;          /* [3 * numNodes] node centres, xyz interleaved           */
    double*   sizes;            /* [3 * numNodes] node half-sizes, xyz interleaved        */
} sphx_octree;

/* sph::GroupView, domain/include/cstone/traversal/groups.hpp:28-34.  Target groups
 * hold at most 64 particles (one wave64).  groupStart == NULL selects fixed groups of
 * 64 consecutive particles covering [firstBody, lastBody). */
typedef struct sphx_groups
{
    uint32_t        firstBody, lastBody, numGroups;
    const uint32_t* groupStart; /* device [numGroups] */
    const uint32_t* groupEnd;   /* device [numGroups] */
} sphx_groups;

/* The particle fields consumed by the hot path; mirrors the field contract of
 * sphexa::ParticlesData (sph/include/sph/particles_data.hpp:218-246, 85-143).
 * x, y, z are always double.  Every void* field is double* when precision ==
 * SPHX_FP64 and float* when precision == SPHX_MIXED.  wh is the kernel table of
 * SPHX_TABLE_SIZE entries in the same element type. */
typedef struct sphx_particles
{
    uint64_t      n;
    int32_t       precision;
    uint32_t      ng0, ngmax;
    float         Kcour;
    double        K, gamma;
    double *      x, *y, *z; /* written only by sphx_compute_positions                  */
    void *        h, *m, *vx, *vy, *vz, *u;
    void *        rho, *p, *c;
    void *        c11, *c12, *c13, *c22, *c23, *c33;
    void *        ax, *ay, *az, *du;
    uint32_t*     nc;
    const void*   wh;
    double        minDtCourant; /* output of sphx_compute_momentum_energy */
    /* integration state (particles_data.hpp:219,236), used by sphx_compute_positions only:
     * x_m1, y_m1, z_m1 = X_n - X_{n-1}, du_m1 = du of the previous step; field type */
    void *x_m1, *y_m1, *z_m1, *du_m1;
} sphx_particles;

#define SPHX_TABLE_SIZE 20000 /* sph::lt::kTableSize, sph/include/sph/table_lookup.hpp:10 */
#define SPHX_MAX_LEVEL 21     /* maxTreeLevel<uint64_t>, domain/include/cstone/tree/definitions.h:55 */

typedef struct sphx_ctx sphx_ctx;

/* ---- context ------------------------------------------------------------- */
/* Creates a context on `device` that launches on `stream` (a hipStream_t, NULL =
 * the device's null stream).  The context owns all scratch memory. */
int         sphx_ctx_create(int device, void* stream, sphx_ctx** out);
int         sphx_ctx_destroy(sphx_ctx* ctx);
int         sphx_ctx_set_stream(sphx_ctx* ctx, void* stream);
int         sphx_ctx_synchronize(sphx_ctx* ctx);
const char* sphx_last_error(const sphx_ctx* ctx);
const char* sphx_version(void);

/* context options */
enum
{
    SPHX_OPT_FP64_EXACT = 1, /* 1: FP64 particle loops compiled without FMA contraction, so
                                every operation rounds like the reference CPU loops (default 0) */
    SPHX_OPT_TIMING     = 2, /* 1: record HIP events on the context stream around every hot
                                kernel launch (read with sphx_timing_summary); 2: around
                                the neighbor-candidate and SPH particle-loop launches
                                only; 0: none */
    SPHX_OPT_LOOP_SHAPE_FP64  = 3, /* particle-loop kernel shape for FP64 fields:          */
    SPHX_OPT_LOOP_SHAPE_MIXED = 4, /* 0 = neighbor records staged in LDS per 16-target     */
                                   /* sub-group (default), 1 = one lane per target with    */
                                   /* neighbor records gathered from global memory         */
    SPHX_OPT_DEFER_MIN_DT = 5      /* 1: the momentum-energy entries do not wait for the
                                      Courant time step: d->minDtCourant is set to NaN and
                                      sphx_min_dt_courant returns the value (default 0: the
                                      entries read it back, synchronising the stream) */
};
int sphx_ctx_set_option(sphx_ctx* ctx, int option, int value);

/* No counterpart in the reference (its momentum-energy calls always copy minDtCourant to
 * the host).  For callers that keep results on the device between calls, with
 * SPHX_OPT_DEFER_MIN_DT set: the Courant time step of the context's last
 * sphx_compute_momentum_energy, sphx_compute_hydro_std or sphx_compute_momentum_energy_ve
 * call, waiting only for that value's copy (INFINITY before any call).  Switching the
 * option off or the context's stream reads a pending value first. */
int sphx_min_dt_courant(sphx_ctx* ctx, double* out);

/* statistics of the last call that produced them */
enum
{
    SPHX_STAT_CANDIDATE_TESTS = 1, /* (target, candidate) pairs tested by the last neighbor search */
    SPHX_STAT_SCRATCH_BYTES   = 2  /* device bytes the context's scratch arena holds              */
};
int sphx_get_stat(sphx_ctx* ctx, int stat, double* value);

/* Per-kernel totals of the launches recorded since the last call (synchronises the
 * stream, then clears the record): up to `cap` entries of kernel name (static
 * string), number of launches and summed milliseconds; *numEntries receives the
 * number of distinct kernels.  Timing is a library service with no counterpart in
 * the reference (which times whole stages, main/src/util/timer.hpp:34-80). */
int sphx_timing_summary(sphx_ctx* ctx, int cap, const char** names, int* launches, double* totalMs,
                        int* numEntries);


/* ---- kernel tables ------------------------------------------------------- */
/* Host-side table and normalisation: replaces ParticlesData::createTables
 * (sph/include/sph/particles_data.hpp:377-387): K = kernel_3D_k(sinc^n, 2)
 * (sph_kernel_tables.hpp:78-85) and wh = tabulateFunction (:88-102) in double and
 * in float.  Any output pointer may be NULL. */
int sphx_kernel_tables(double sincIndex, double* wh64, float* wh32, double* K);

/* ---- coordinate box ------------------------------------------------------ */
/* cstone::makeGlobalBox (domain/include/cstone/sfc/box_mpi.hpp:68-107) on one rank:
 * periodic and fixed dimensions keep the limits of *previous; open dimensions get
 * the min and max of the n coordinates (unchanged when n == 0).  Synchronises the
 * stream once (the box is returned on the host). */
int sphx_make_global_box(sphx_ctx* ctx, const double* x, const double* y, const double* z, uint64_t n,
                         const sphx_box* previous, sphx_box* out);

/* ---- SFC keys ------------------------------------------------------------ */
/* cstone::computeSfcKeys<true> (domain/include/cstone/sfc/sfc_gpu.h:26-31,
 * CPU semantics sfc/sfc.hpp:267-275): 64-bit Hilbert keys; entries already equal
 * to removeKey (2^63) are left untouched. */
int sphx_compute_sfc_keys(sphx_ctx* ctx, const double* x, const double* y, const double* z, uint64_t* keys,
                          uint64_t n, const sphx_box* box);

/* ---- sort / permute ------------------------------------------------------ */
/* cstone::sortByKey<true> (domain/include/cstone/primitives/primitives_acc.hpp:115-121,
 * primitives_gpu.cu:308-339): stable ascending radix sort of (key, value) pairs, in
 * place.  From 2^18 pairs on, four passes sort by the top 32 key bits and runs of equal
 * top halves up to 64 long are then sorted by the low halves (a longer run makes the
 * call run all eight passes); when the context's previous such sort found every run of
 * equal top 24 bits at most 64 long, three passes sort by the top 24 bits and those runs
 * are fixed up instead.  The result is the same.  Synchronises the stream once (the
 * digit check of the eight-pass sort, or the run-length check).  Scratch comes from the
 * context. */
int sphx_sort_by_key(sphx_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t n);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * stable sort as sphx_sort_by_key, bit for bit, with every work buffer in one fixed device
 * array of the library (about 25.3 MiB, one copy per GPU, held until the process ends).
 * n must be <= 2097152 (2^21) keys, else SPHX_ERR_ARG.  It allocates nothing, never
 * synchronises the stream and sets no memory with memset (every digit pass runs).  All
 * calls on one GPU share the array: calls that overlap on one GPU race. */
int sphx_sort_by_key_noalloc(sphx_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t n);
/* cstone::sequence<true> (primitives_acc.hpp:108-113): values[i] = first + i */
int sphx_sequence(sphx_ctx* ctx, uint32_t* values, uint64_t n, uint32_t first);
/* cstone::gather (domain/include/cstone/primitives/gather.hpp:90-98): out[i] = in[order[i]] */
int sphx_gather_f64(sphx_ctx* ctx, const uint32_t* order, const double* in, double* out, uint64_t n);
int sphx_gather_f32(sphx_ctx* ctx, const uint32_t* order, const float* in, float* out, uint64_t n);
int sphx_gather_u32(sphx_ctx* ctx, const uint32_t* order, const uint32_t* in, uint32_t* out, uint64_t n);

/* cstone::gatherArrays (domain/include/cstone/domain/layout.hpp:231-260) with the output
 * offset 0: out[k][i] = in[k][order[i]] for count <= SPHX_GATHER_MAX_ARRAYS arrays of
 * 4- or 8-byte elements (elemBytes[k]) in one launch.  in, out and elemBytes are host
 * arrays of device pointers; outputs may overlap neither inputs nor each other (the
 * reference gathers into scratch buffers and swaps them in). */
#define SPHX_GATHER_MAX_ARRAYS 16
int sphx_gather_arrays(sphx_ctx* ctx, const uint32_t* order, uint64_t n, int32_t count, const void* const* in,
                       void* const* out, const uint32_t* elemBytes);

/* ---- cornerstone octree --------------------------------------------------- */
/* cstone::computeOctree (domain/include/cstone/tree/csarray.hpp:430-440; GPU loop
 * updateOctreeGpu, tree/update_gpu.cuh:41-64): the converged cornerstone leaf array
 * of sorted `keys` for `bucketSize`.  Results are written into context-owned device
 * memory, valid until the next sphx_compute_octree on this context:
 * *leaves -> [*numLeaves + 1] keys, *counts -> [*numLeaves] particle counts
 * (capped at maxCount), *layout -> [*numLeaves + 1] exclusive scan of the counts.
 * Synchronises the stream once (to return numLeaves). */
int sphx_compute_octree(sphx_ctx* ctx, const uint64_t* keys, uint64_t n, uint32_t bucketSize, uint32_t maxCount,
                        int32_t* numLeaves, uint64_t** leaves, uint32_t** counts, uint32_t** layout);

/* Array sizes of an internal octree with numLeafNodes leaves (OctreeData::resize,
 * tree/octree.hpp:291-306). */
int sphx_octree_sizes(int32_t numLeafNodes, int32_t* numInternalNodes, int32_t* numNodes, int32_t* numParents);

/* cstone::buildOctreeGpu / updateInternalTree (tree/octree_gpu.h:34,
 * tree/octree_gpu.cu:138-188, tree/octree.hpp:168-196,361-367): link the internal
 * octree of `leaves` into the caller-allocated arrays of `tree` (prefixes,
 * childOffsets, parents, levelRange, internalToLeaf, leafToInternal; numLeafNodes
 * must be set; the other counts are filled in). */
int sphx_build_octree(sphx_ctx* ctx, const uint64_t* leaves, sphx_octree* tree);

/* cstone::nodeFpCenters / computeGeoCentersGpu (focus/source_center.hpp:146-158,
 * focus/source_center_gpu.cu:222-235): geometric centre and half-size of every node. */
int sphx_node_centers(sphx_ctx* ctx, const uint64_t* prefixes, int32_t numNodes, const sphx_box* box,
                      double* centers, double* sizes);

/* ---- target groups -------------------------------------------------------- */
/* sph::computeGroups -> cstone::computeGroupSplits (sph/include/sph/groups.hpp:41-74,
 * traversal/groups_gpu.cu:108-137): SFC-consecutive groups of at most 64
 * particles: each fixed 64-particle group is split after particle p when
 * |X(p+1) - X(p)| > min(tolFactor * cbrt(smallest leaf volume in the group),
 * 2 h_p / min box length), X = position / box lengths (unit-box coordinates; h may be
 * NULL: 2h term = 1; h is double for SPHX_FP64, float for SPHX_MIXED).  Writes
 * group boundaries into context memory; *groups is filled with device pointers
 * valid until the next call (groupEnd = groupStart + 1).  Synchronises the stream
 * once (to return numGroups).  tree needs numLeafNodes, leaves and layout. */
int sphx_compute_groups(sphx_ctx* ctx, uint32_t first, uint32_t last, const double* x, const double* y,
                        const double* z, const void* h, int32_t precision, const sphx_octree* tree,
                        const sphx_box* box, float tolFactor, sphx_groups* groups);

/* ---- neighbor search ------------------------------------------------------ */
/* cstone::findNeighbors (domain/include/cstone/findneighbors.hpp:149-170): for
 * i in [first, last), the neighbors j != i with d^2(i,j) < 4 h_i^2 (minimum image
 * for periodic dimensions when the search sphere leaves the box).  nbrs
 * (optional, may be NULL) uses the reference CPU layout nbrs[(i - first) * ngmax + k],
 * in ascending j, truncated at ngmax; nc[i - first] is the full count. */
int sphx_find_neighbors(sphx_ctx* ctx, const double* x, const double* y, const double* z, const double* h,
                        uint32_t first, uint32_t last, const sphx_box* box, const sphx_octree* tree,
                        uint32_t ngmax, uint32_t* nbrs, uint32_t* nc);

/* sph::findNeighborsSfc -> findNeighborsSph (sph/include/sph/find_neighbors.hpp:10-53):
 * neighbor search with up to 10 smoothing-length updates (updateH, kernels.hpp:27-33)
 * per particle outside [ng0/4, ngmax + 1]; d->nc[i] counts self.  The lists are kept
 * in the context for the SPH loops below, grouped by `groups`. */
int sphx_find_neighbors_sph(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, const sphx_box* box,
                            const sphx_octree* tree);

/* Copy the context's lists (from sphx_find_neighbors_sph) into the reference CPU
 * layout nbrs[(i - firstBody) * ngmax + k]. */
int sphx_export_neighbors(sphx_ctx* ctx, uint32_t* nbrs);

/* ---- std-SPH particle loops ------------------------------------------------ */
/* sph::computeDensity (sph/include/sph/hydro_std/density.hpp:41-60, xmassJLoop
 * hydro_ve/xmass_kern.hpp:50-79): rho_i = K/h_i^3 (m_i + sum_j m_j W(r_ij/h_i)). */
int sphx_compute_density(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, const sphx_box* box);

/* sph::computeEOS_HydroStd, ideal gas (hydro_std/eos.hpp:52-78,144-150; sph/eos.hpp:34-44):
 * p = rho (gamma - 1) u, c = sqrt(gamma (gamma - 1) u) for i in [first, last). */
int sphx_compute_eos(sphx_ctx* ctx, uint32_t first, uint32_t last, sphx_particles* d);

/* sph::computeIAD (hydro_std/iad.hpp:41-79, IADJLoopSTD hydro_std/iad_kern.hpp:12-78):
 * the inverse of the symmetric tau tensor -> c11, c12, c13, c22, c23, c33. */
int sphx_compute_iad(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, const sphx_box* box);

/* sph::computeMomentumEnergySTD (hydro_std/momentum_energy.hpp:41-101,
 * momentumAndEnergyJLoop hydro_std/momentum_energy_kern.hpp:12-134): ax, ay, az,
 * du and d->minDtCourant (Courant time step, kernels.hpp:10-17).  Synchronises
 * the stream once to return minDtCourant (with SPHX_OPT_DEFER_MIN_DT: not at all). */
int sphx_compute_momentum_energy(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d,
                                 const sphx_box* box);

/* The four calls above in the order HydroProp::computeForces makes them after the
 * neighbor search (main/src/propagator/std_hydro.hpp:139-153): density, EOS over the
 * groups' [first, last), IAD, momentum-energy, on one rank (no halo exchange between
 * them).  Same results as the four calls.  When the groups cover every particle, the
 * density loop's epilogue writes the IAD loop's neighbor records (x, y, z, m/rho) and,
 * for MIXED fields, the IAD loop's epilogue the momentum-energy records, in place of
 * packing launches.  Synchronises the stream once to return minDtCourant (with
 * SPHX_OPT_DEFER_MIN_DT: not at all). */
int sphx_compute_hydro_std(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, const sphx_box* box);

/* ---- integration step ------------------------------------------------------ */
/* The dataset's global time-step scalars (sph/include/sph/particles_data.hpp:93-111,138). */
typedef struct sphx_timestep
{
    double minDt, minDt_m1, ttot; /* state: updated by sphx_compute_timestep               */
    double minDtCourant;          /* candidate limits: Courant (sphx_compute_momentum_energy */
    double minDtRho;              /* output) and any other limit (INFINITY when unused)     */
    double g, etaAcc;             /* gravitational constant; the acceleration limit is used  */
                                  /* when g != 0 (accelerationTimestep, ts_global.hpp:48-71) */
    double minDtAcc;              /* output: that limit (INFINITY when g == 0)               */
} sphx_timestep;

/* sph::computeTimestep (sph/include/sph/ts_global.hpp:100-130) on one rank:
 * minDt_new = min(minDtAcc, minDtCourant, minDtRho, 1.1 minDt), where minDtAcc =
 * etaAcc * (min_i h_i^2 / |a_i|^2)^(1/4) over [first, last) when g != 0; then
 * ttot += minDt_new, minDt_m1 = minDt, minDt = minDt_new.  Synchronises the stream
 * once when g != 0 (the minimum is returned on the host). */
int sphx_compute_timestep(sphx_ctx* ctx, uint32_t first, uint32_t last, const sphx_particles* d,
                          sphx_timestep* ts);

/* sph::computePositions, CPU semantics (sph/include/sph/positions.hpp:113-137,162-179,
 * 207-230): for the particles of `groups`, the 2nd-order Press position update
 * positionUpdate(minDt, minDt_m1, X, A, X_m1, box) (:100-111, with putInBox for periodic
 * dimensions, cstone/sfc/box.hpp:193-214) writes x, y, z, x_m1, y_m1, z_m1 and vx, vy, vz; near fixed
 * boundaries A and X_m1 are scaled by fbcAdjustFactors (:47-74).  Then the Adams-Bashforth
 * energy update u = energyUpdate(u, minDt f, minDt_m1 f, du, du_m1) (:77-84), f the smallest
 * factor, and du_m1 = du.  Uses ts->minDt and ts->minDt_m1 as sphx_compute_timestep left them. */
int sphx_compute_positions(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, const sphx_box* box,
                           const sphx_timestep* ts);

/* sph::updateSmoothingLength (sph/include/sph/update_h.hpp:13-49): for the particles of
 * `groups`, h = updateH(ng0, nc, h) (kernels.hpp:27-33); particles with nc <= 1 get
 * keys[i] = removeKey (2^63; keys may be NULL) and set *keysRemoved = 1 (else 0).
 * Synchronises the stream once (to return *keysRemoved). */
int sphx_update_smoothing_length(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, uint64_t* keys,
                                 int* keysRemoved);

/* Host-side factors of updateH (sph/include/sph/kernels.hpp:27-33), which every h update
 * of the library applies as h' = (h * 0.5) * f[nc]: f64[nc] = pow(1 + 1023 ng0 / nc, 0.1)
 * in double and f32[nc] the same in float (powf), for nc in [0, count), computed with the
 * host libm as the reference CPU path computes them.  The library tabulates nc < 131072
 * per context; sphx_find_neighbors_sph and sphx_update_smoothing_length return an error
 * for a count beyond that.  Either output may be NULL. */
int sphx_smoothing_length_factors(uint32_t ng0, uint32_t count, double* f64, float* f32);

/* ---- VE hydro (the reference's default propagator) ------------------------- */
/* The fields of the volume-element (VE) loops that sphx_particles does not hold
 * (sph/include/sph/particles_data.hpp:218-246); every void* is double* for SPHX_FP64
 * and float* for SPHX_MIXED, as in sphx_particles.  Outputs of one loop are inputs of
 * the next, in the order of HydroVeProp::computeForces (main/src/propagator/
 * ve_hydro.hpp:126-205): xmass -> veDefGradh -> EOS -> iadDivvCurlv -> AVswitches ->
 * momentum-energy.  dV11..dV33 (the symmetrised velocity gradient) and curlv are
 * optional (NULL: not computed). */
typedef struct sphx_ve_fields
{
    void *      xm, *kx, *gradh, *prho, *divv, *curlv, *alpha;
    void *      dV11, *dV12, *dV13, *dV22, *dV23, *dV33;
    const void* whd;                          /* kernel-derivative table, SPHX_TABLE_SIZE entries */
    double      alphamin, alphamax, decay;    /* AV switches (particles_data.hpp:127-129)         */
    double      Atmin, Atmax, ramp;           /* Atwood ramp (particles_data.hpp:134-136)         */
    int32_t     avClean;                      /* 1: AV cleaning in momentum-energy (needs dV)     */
    int32_t     reserved;
} sphx_ve_fields;

/* Host-side derivative table: tabulateFunction of getSphKernelDerivative(sinc_n)
 * (sph/include/sph/sph_kernel_tables.hpp:88-102,161-172), the whd of createTables
 * (particles_data.hpp:377-387), in double and float; either pointer may be NULL. */
int sphx_kernel_derivative_tables(double sincIndex, double* whd64, float* whd32);

/* sph::computeXMass (sph/include/sph/hydro_ve/xmass.hpp:39-75, xmassJLoop
 * xmass_kern.hpp:50-79): xm_i = m_i / (K/h_i^3 (m_i + sum_j m_j W(r_ij/h_i))). */
int sphx_compute_xmass(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, sphx_ve_fields* ve,
                       const sphx_box* box);

/* sph::computeVeDefGradh (hydro_ve/ve_def_gradh.hpp:40-87, veDefGradhJLoop
 * ve_def_gradh_kern.hpp:43-89): the VE normalisation kx and the grad-h term gradh. */
int sphx_compute_ve_def_gradh(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, sphx_ve_fields* ve,
                              const sphx_box* box);

/* sph::computeEOS, ideal gas with u (hydro_ve/eos.hpp:52-91,192-197): rho = kx m / xm,
 * p = rho (gamma - 1) u, c = sqrt(gamma (gamma - 1) u), prho = p / (kx m^2 gradh);
 * writes ve->prho and d->c, and d->rho, d->p when they are not NULL. */
int sphx_compute_eos_ve(sphx_ctx* ctx, uint32_t first, uint32_t last, sphx_particles* d, sphx_ve_fields* ve);

/* sph::computeIadDivvCurlv (hydro_ve/iad_divv_curlv.hpp:40-98): the IAD tensor with the
 * VE volume xm/kx (IADJLoop, hydro_ve/iad_kern.hpp:43-110) into d->c11..c33, then
 * divv, curlv and the symmetrised velocity gradient (divV_curlVJLoop,
 * divv_curlv_kern.hpp:43-122) with those tensors. */
int sphx_compute_iad_divv_curlv(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, sphx_ve_fields* ve,
                                const sphx_box* box);

/* sph::computeAVswitches (hydro_ve/av_switches.hpp:40-88, AVswitchesJLoop
 * av_switches_kern.hpp:43-136): updates ve->alpha in place over the time step dt. */
int sphx_compute_av_switches(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, sphx_ve_fields* ve,
                             const sphx_box* box, double dt);

/* sph::computeMomentumEnergy<avClean> (hydro_ve/momentum_energy.hpp:40-117,
 * momentumAndEnergyJLoop momentum_energy_kern.hpp:65-221, tdpdTrho absent), avClean =
 * ve->avClean (then the AV correction avRvCorrection :43-63 reads dV11..dV33): ax, ay, az,
 * du and d->minDtCourant.  Synchronises the stream once to return minDtCourant (with
 * SPHX_OPT_DEFER_MIN_DT: not at all). */
int sphx_compute_momentum_energy_ve(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d,
                                    sphx_ve_fields* ve, const sphx_box* box);

/* sph::rhoTimestep (sph/include/sph/ts_global.hpp:73-98): *minDtRho = Krho / |max divv|
 * over [first, last) (ve->divv, field type of d->precision).  Synchronises the stream once. */
int sphx_rho_timestep(sphx_ctx* ctx, uint32_t first, uint32_t last, const sphx_particles* d,
                      const sphx_ve_fields* ve, double Krho, double* minDtRho);

/* ---- adjoints -------------------------------------------------------------- */
/* No counterpart in the reference (its loops are forward-only).  Vector-Jacobian
 * product of sphx_compute_density with respect to the positions and h, on the lists
 * of the last sphx_find_neighbors_sph (same groups, covering [0, n)): given
 * gRho[i] = dL/drho_i (field type), writes gx[i] = dL/dx_i, gy, gz (double, n each,
 * overwritten) and, if gh != NULL, gh[i] = dL/dh_i (field type).  The derivative is
 * that of the forward's own arithmetic: the table-linear kernel (FP64) or the FP32
 * polynomial kernel (MIXED); neighbor sets are held fixed.  j-side terms are summed
 * with atomic adds: reproducible to rounding, not bit for bit. */
int sphx_density_adjoint(sphx_ctx* ctx, const sphx_groups* groups, const sphx_particles* d, const sphx_box* box,
                         const void* gRho, double* gx, double* gy, double* gz, void* gh);

/* No counterpart in the reference.  Vector-Jacobian product of
 * sphx_compute_momentum_energy with respect to the positions, holding v, h, m, rho,
 * p, c and c11..c33 fixed: given dL/dax, dL/day, dL/daz, dL/ddu (field type; any may
 * be NULL = zero), writes gx, gy, gz (double, n each, overwritten).  Lists and groups
 * as for sphx_density_adjoint. */
int sphx_momentum_energy_adjoint(sphx_ctx* ctx, const sphx_groups* groups, const sphx_particles* d,
                                 const sphx_box* box, const void* gax, const void* gay, const void* gaz,
                                 const void* gdu, double* gx, double* gy, double* gz);

#ifdef __cplusplus
}
#endif

#endif /* SPHX_H */
