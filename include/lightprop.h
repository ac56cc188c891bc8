/*
 * lightprop.h -- C ABI of the MI355X photon-transport library (liblp.so).
 *
 * Plain C types only: pointers, sizes, int status codes.  Inputs are copied
 * in (the caller keeps ownership); outputs are copied into caller buffers.
 * Nothing allocated by the library crosses the boundary.  A context is
 * single-threaded; separate contexts may run concurrently.  No call throws.
 *
 * The reference crosses its host/device boundary through the OpenCL kernel
 * functor of src/Simulation/CalculatePropagation.hpp:70-88 (19 arguments,
 * invoked at CalculatePropagation.cpp:1516-1535) driven by the
 * CalculatePropagation setters / Init() / Propagation() / Get*Image()
 * (CalculatePropagation.hpp:212-389; call order main.cpp:174-222).  Each
 * entry point below names the reference interface it replaces.  The C++
 * host layer lightpropagation_amd/host/lp_host.hpp keeps the reference's
 * class interfaces (CalculatePropagation and its neighbours) on top of this
 * ABI; include/lightprop_host.h is that layer's own C ABI.
 */
#ifndef LIGHTPROP_H
#define LIGHTPROP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  SUCCESS / FAILURE / EXPECTED_FAILURE keep the values of
 * CalculatePropagation.cpp:23-25; the rest refine FAILURE. */
#define LP_OK 0
#define LP_FAILURE 1
#define LP_EXPECTED_FAILURE 2
#define LP_EINVAL 3      /* bad argument                         */
#define LP_ENOMEM 4      /* device allocation failed             */
#define LP_EDEVICE 5     /* HIP runtime / launch error           */
#define LP_ESTATE 6      /* call out of order (e.g. no mesh yet) */

/* Random-number modes (lp_set_params). */
#define LP_RNG_PHILOX 0  /* Philox4x32-10 keyed by (seed, photon index)        */
#define LP_RNG_RAN3 1    /* per-photon stream of the reference's float ran3     */

/* Tally selectors (lp_set_tallies bit mask, lp_get_tally `which`). */
#define LP_TALLY_SFT 0    /* exit image (R, F, T) over x,y: GetSFTImage        */
#define LP_TALLY_RTD 1    /* exit (radius, angle) image (R, F, T): GetRTdImage */
#define LP_TALLY_ARZ 2    /* absorbed weight over (signed r, z): GetArzImage   */
#define LP_TALLY_PHIRZ 3  /* absorbed weight / mua over (r, z): GetPhirzImage  */
#define LP_TALLY_REGION 4 /* absorbed weight per tissue                        */
#define LP_TALLY_VOX 5    /* absorbed weight on a voxel grid                   */
#define LP_TALLY_RAD 6    /* exit weight (R, F, T) over radial bins             */
#define LP_TALLY_DET 7    /* reflected weight (R, F) per annular detector       */
#define LP_TALLY_MISC 8   /* counters: photons, steps, interactions, boundary
                             events, escaped, capped, off-grid Arz weight,
                             off-grid voxel weight, BVH node visits, triangle
                             tests, skipped searches, NaN photons, free-radius
                             queries, their node visits, list face tests,
                             wavefront iterations, near-list search batches,
                             BVH search batches, replay mismatches, marches,
                             edge-crawl crossings replayed (32 slots)         */
#define LP_NTALLY 9

/* Tally values are int64 fixed point: 2^32 per unit photon weight, except
 * LP_TALLY_PHIRZ (2^24 per unit of weight / mua) and LP_TALLY_MISC (plain
 * counts and 2^32 weights).  Sums are exact and independent of the order in
 * which lanes add them. */
#define LP_FIXED_SCALE 4294967296.0
#define LP_FIXED_SCALE_PHI 16777216.0

typedef struct lp_context lp_context;

/* Sizes of the tally images, derived as the reference host derives them
 * (InitLogVectors, CalculatePropagation.cpp:120-243). */
typedef struct lp_layout {
    int32_t dimx, dimy, dimz, dimR, dimAlpha;
    int32_t sft_nx, sft_ny; /* SFT: x in [-dimx/2, dimx/2], y likewise, 3 comps */
    int32_t arz_nr, arz_nz; /* Arz: r in [-dimR, dimR], z in [0, dimz]          */
    float dist_min;         /* R/T classification distance (MCLightPropagation.cl:1351) */
    float rsp;              /* specular reflectance (MCLightPropagation.cl:273-299)     */
    int64_t n[LP_NTALLY];   /* element count of each tally                      */
} lp_layout;

typedef struct lp_stats {
    double kernel_ms;       /* device time of the last lp_run (HIP events)      */
    int64_t photons;        /* photons of the last lp_run                       */
    int64_t bvh_nodes;      /* nodes of the face-search BVH                     */
    int32_t bvh_depth;
    int32_t grid_blocks;    /* persistent work-groups launched                  */
    int64_t device_bytes;   /* device memory held by the context               */
    double replay_ms;       /* lp_adjoint: device time of the replay launch     */
    int64_t replayed;       /* lp_adjoint: (photon, detector) pairs replayed    */
    int64_t replay_mismatch; /* lp_adjoint: replays that ended with another R
                                weight than the forward photon (0 when exact)   */
} lp_stats;

/* Replaces the CalculatePropagation constructor + createContext /
 * CreateCommandQueue (CalculatePropagation.cpp:27-80, 1225-1339). */
int lp_create(lp_context **ctx, int32_t device);
void lp_destroy(lp_context *ctx);
/* Text of the last error of this context (never NULL). */
const char *lp_last_error(const lp_context *ctx);

/* Mesh: triangle t has vertices tri[9t..9t+8], normal nrm[3t..3t+2] and
 * (inner, outer) tissue ids med[2t], med[2t+1].  Replaces SetTriangles /
 * SetNormals / SetInnerAndOuterTissueIndices (CalculatePropagation.cpp:
 * 541-667) and the triangles/normals/mediumIDs kernel buffers.  Builds the
 * BVH and the safety grid on the host and uploads them. */
int lp_set_mesh(lp_context *ctx, const float *tri, const float *nrm, const int32_t *med, int64_t ntri);

/* Several geometries in one mesh (no reference counterpart: the reference
 * kernel searches one triangle list, MCLightPropagation.cl:1262-1270, and
 * a simulation loads one mesh, CalculatePropagation.cpp:541-667): the mesh
 * of lp_set_mesh is ngeom geometries, geometry g being the next
 * tri_count[g] triangles (the counts add up to the mesh's).  Each geometry
 * gets a BVH and near-face grid of its own, so a photon of geometry g meets
 * only g's faces and takes exactly the path it takes in a mesh of g alone;
 * geometries may overlap in space and share the tissue ids, source and
 * detection grid.  lp_run_batch_geom picks each instance's geometry; lp_run,
 * lp_trace, lp_adjoint and lp_autotune transport in geometry 0.  lp_set_mesh
 * resets to one geometry.  The safety grid (lp_set_launch) needs a single
 * geometry.  At most 64 geometries. */
int lp_set_geometries(lp_context *ctx, const int64_t *tri_count, int32_t ngeom);

/* nmuasg[4i..4i+3] = (n, mua, mus, g) of tissue i; tissue 0 is where photons
 * start.  Replaces SetTissueProperties (cpp:427-451) + SetFetalTissueID. */
int lp_set_materials(lp_context *ctx, const float *nmuasg, int32_t ntissue, int32_t fetal_id);

/* Source: start point, nstart directions vstart[3*nstart] with photon counts
 * vocc[nstart], and (POI, eX, eY, eZ) = poi_basis[12].  Replaces
 * SetStartTrajectories / SetStartPoint / SetPOI / SetBasisVectors
 * (cpp:316-350, 713-773) and the vStart, vStartOcc, pStart, e_P0XYZ args. */
int lp_set_source(lp_context *ctx, const float pstart[3], const float *vstart, const int32_t *vocc,
                  int32_t nstart, const float poi_basis[12]);

/* Model extent (x, y, z) and detection grid (x, y, z, r, phi, rho).
 * Replaces SetDimensions / SetDetectionGridDimensions (cpp:539-540, 706-712). */
int lp_set_grid(lp_context *ctx, const float model_dims[3], const float det_grid[6]);

/* Roulette chance, weight threshold, PixelPerUnit factor, seed, RNG mode and
 * a per-photon step cap (<= 0: the default of 10^6; at most 2^31 - 1).  Replaces
 * SetChanceForRoulette / SetWeightThreshold / SetPixelPerUnitFactor
 * (cpp:405-425) and the PSNR seeding of setup() (cpp:1893-1899). */
int lp_set_params(lp_context *ctx, float chance, float w_th, float factor, uint64_t seed, int32_t rng_mode,
                  int64_t max_steps);

/* Tallies without a reference counterpart: voxel grid in the emitter frame
 * (origin, voxel size, counts), radial exit bins, annular detectors on the
 * exit radius [rlo, rhi). */
int lp_set_voxel_grid(lp_context *ctx, const float origin[3], const float size[3], const int32_t n[3]);
int lp_set_radial(lp_context *ctx, float dr, int32_t nbins);
int lp_set_detectors(lp_context *ctx, const float *rlo, const float *rhi, int32_t ndet);

/* Bit mask of (1 << LP_TALLY_*) to keep (MISC and REGION are always kept).
 * record_exit != 0 keeps the per-photon float8 exit records of the next
 * lp_run (the SurfFetalTransRxyG buffer of the reference kernel). */
int lp_set_tallies(lp_context *ctx, uint32_t mask, int32_t record_exit);

int lp_get_layout(lp_context *ctx, lp_layout *out);
int lp_reset_tallies(lp_context *ctx);

/* Transport photons [first, first + count) and add to the tallies; blocks
 * until done.  Replaces Init() + Propagation() / ExecKernel
 * (cpp:1389-1759, 2003-2028): one persistent launch instead of a launch,
 * wait, copy and host binning per 6400 photons. */
int lp_run(lp_context *ctx, int64_t first, int64_t count);

/* Copy tally `which` (n elements) to out. */
int lp_get_tally(lp_context *ctx, int32_t which, int64_t *out, int64_t n);

/* Per-photon exit records of the last lp_run (8 floats each: coords(3),
 * SFT(3), exit angle, exit radius; MCLightPropagation.cl:1586-1594). */
int lp_get_exit_records(lp_context *ctx, float *out, int64_t nphotons);

/* Runs photons [first, first + count) with a per-step event trace (12 floats
 * per step: kind, face, tissue, w, p(3), v(3), step, draws) of up to
 * max_rec steps each; nrec[i] receives the number of records of photon i.
 * Tallies are accumulated as by lp_run.  No reference counterpart: the
 * step-by-step view the parity tests compare with the oracle's. */
int lp_trace(lp_context *ctx, int64_t first, int32_t count, int32_t max_rec, float *out, int32_t *nrec);

/* Kernel time and launch figures of the last lp_run / lp_adjoint /
 * lp_run_batch (the reference's PrintTimeSpent, CalculatePropagation.cpp:
 * 2063, prints wall time only). */
int lp_get_stats(lp_context *ctx, lp_stats *out);

/* ---- adjoint (no reference counterpart) ----
 * Gradients of the annular detector signals S_d (sum of the R exit weights
 * W of the photons whose exit radius lies in detector d, as LP_TALLY_DET
 * counts them) with respect to each tissue's mua and mus, and to mua in each
 * voxel of a grid in the emitter frame, by the path-length (perturbation
 * Monte Carlo) estimator on every detected path:
 *   dW/dmua_t = -L_t W,   dW/dmus_t = (n_t / mus_t - L_t) W,   dW/dmua_v = -L_v W
 * (L: path length, n: interactions in tissue t / voxel v).  lp_adjoint runs
 * photons [first, first + count) forward (adding to the tallies as lp_run
 * does) while recording which photons reach which detector, then replays
 * only those photons from their (seed, index) random streams in a second
 * launch that accumulates the gradients.  Outputs are 2^32 fixed point:
 * signals[ndet], grad[ndet][ntissue][2] (d/dmua, d/dmus), jac[ndet][nvox]
 * (x fastest; NULL skips the voxel Jacobian); npairs = (photon, detector)
 * pairs replayed.  Needs detectors and at most 16 tissues. */
int lp_set_jacobian_grid(lp_context *ctx, const float origin[3], const float size[3], const int32_t n[3]);
int lp_adjoint(lp_context *ctx, int64_t first, int64_t count, int64_t *signals, int64_t *grad, int64_t *jac,
               int64_t *npairs);
/* lp_adjoint with the per-voxel mus Jacobian as well (no reference
 * counterpart): jac_mus[ndet][nvox] = dS_d / dmus_v, the mus of the tissue
 * in voxel v perturbed there only, by the same estimator:
 *   dW/dmus_v = (sum over the interactions in v of 1 / mus) W - L_v W,
 * 2^32 fixed point, accumulated as the mua term (jac) plus W / mus of each
 * interaction in the voxel, so jac_mus - jac is the interaction term alone.
 * jac or jac_mus may be NULL. */
int lp_adjoint_voxels(lp_context *ctx, int64_t first, int64_t count, int64_t *signals, int64_t *grad, int64_t *jac,
                      int64_t *jac_mus, int64_t *npairs);
/* Per-pair path records of the adjoint (no reference counterpart): with
 * enable != 0, every later lp_adjoint / lp_adjoint_voxels keeps, for each
 * (photon, detector) pair it replays, the photon's path length and
 * interaction count in each tissue (4 + 4 bytes per tissue and pair on the
 * device until the next call); enable = 0 frees them.  These are the
 * perturbation Monte Carlo records: the pair's weight at other optics is
 *   W' = W prod_t (mus'_t / mus_t)^n_t exp(-(mua'_t + mus'_t - mua_t - mus_t) L_t). */
int lp_set_adjoint_paths(lp_context *ctx, int32_t enable);
/* Copies the records of the last lp_adjoint out, sorted by (photon,
 * detector): photon[npairs], det[npairs], w[npairs] (R exit weight),
 * length[npairs][ntissue], interactions[npairs][ntissue].  npairs must be
 * the count that call returned. */
int lp_get_adjoint_paths(lp_context *ctx, int64_t npairs, int64_t *photon, int32_t *det, float *w, float *length,
                         uint32_t *interactions);

/* Evaluates the kernel's math core element-wise on the device (op 0 sqrt,
 * 1 divide, 2 log, 3 cos, 4 acos, 5 multiply, 6 add, 7 the square root of
 * the scattering path, exact for 0 and arguments of at least 2^-96, 8 the
 * reciprocal of its normalize, exact on [2^-60, 2^60], 9 the step-length
 * log, for draws in [2^-126, 1]) over n (x, y) pairs.
 * No reference counterpart (the reference calls OpenCL built-ins): it lets
 * the tests compare the device math with the oracle's bit for bit. */
int lp_debug_math(lp_context *ctx, int32_t op, const float *xy, float *out, int64_t n);

/* Runs a plain device-to-device copy of `bytes` (multiple of 16) through a
 * streaming kernel and reports its time: the reference for converting the
 * profiler's FETCH_SIZE / WRITE_SIZE counters into bytes. */
int lp_debug_copy(lp_context *ctx, int64_t bytes, double *ms);

/* Host only, no device: builds the merged arrays of a mesh of ngeom
 * geometries as lp_set_geometries does (tri[9 * ntri], med[2 * ntri],
 * tri_count[ngeom]; BVH tops of lds_per_part nodes each; near grid of radius
 * r0 and cell size `cell`, r0 <= 0: none) and returns them for inspection.
 * sizes[6] = (nodes, triangles, cell records incl. one sentinel per
 * geometry, list entries, cells per geometry incl. its sentinel, 0); each
 * output pointer may be NULL, else receives its whole array: nodes[8 * nodes]
 * (lo.xyz, hit, hi.xyz, miss as raw 32-bit words), tri_ids[triangles] (the
 * original id of each leaf-order triangle), roots[ngeom], coffs[ngeom],
 * cells[2 * cell records] (first list entry, clearance bits), list[list
 * entries] (leaf-order triangle of each entry).  No reference counterpart:
 * the tests check the merged layout on the CPU. */
int lp_debug_multi_mesh(const float *tri, const int32_t *med, const int64_t *tri_count, int32_t ngeom,
                        int32_t lds_per_part, float r0, float cell, int64_t sizes[6], uint32_t *nodes,
                        int32_t *tri_ids, int32_t *roots, uint32_t *coffs, uint32_t *cells, int32_t *list);

/* Launch shape override: persistent work-groups (0 = automatic) and BVH
 * nodes staged in LDS (-1 = automatic), safety grid cells per axis (0 =
 * off, -1 = automatic).  Replaces the fixed 6 400 work-items per launch of
 * the reference (CalculatePropagation.cpp:76). */
int lp_set_launch(lp_context *ctx, int32_t grid_blocks, int32_t lds_nodes, int32_t safety_cells);

/* Tuning options (no effect on results, only on speed; no reference
 * counterpart). */
#define LP_OPT_GRID_BLOCKS 0    /* persistent work-groups, 0 = automatic            */
#define LP_OPT_LDS_NODES 1      /* BVH nodes staged in LDS (0..2048)                */
#define LP_OPT_SAFETY_CELLS 2   /* safety grid cells on the longest axis (0 = off)  */
#define LP_OPT_DIST_CAP 3       /* free-radius query: search cap in length units (0 = off) */
#define LP_OPT_DIST_CAP_STEPS 4 /* free-radius query: cap in units of the step length */
#define LP_OPT_BLOCKS_PER_CU 5  /* work-groups per CU when GRID_BLOCKS is automatic (0,
                                   default: as many as one CU holds at once, so the
                                   persistent grid fills the chip at the variant's
                                   occupancy)                                       */
#define LP_OPT_FREE_ITERS 6     /* wavefront iterations per round (search batches are
                                   forced and finished lanes refilled at its end)  */
#define LP_OPT_WAVES_PER_SIMD 7 /* register budget of the kernel variant: 1 (compiler's
                                   choice), 4 (default), 5 or 6 waves per SIMD     */
#define LP_OPT_NEAR_R0 8        /* near-face lists: radius r0 (0 = off, default 0.07) */
#define LP_OPT_NEAR_CELL 9      /* near-face lists: cell edge (0 = r0 / 0.7)        */
#define LP_OPT_SEARCH_MIN 10    /* run a BVH search batch once this fraction of a
                                   wavefront's live lanes waits for one (0..1)     */
#define LP_OPT_LIST_MIN 11      /* run a near-list search batch once this fraction
                                   of a wavefront's live lanes waits for one (0..1) */
#define LP_OPT_REFILL_MIN 12    /* start new photons once this many lanes of a
                                   wavefront have finished (1..64)                 */
#define LP_OPT_GRAPH 13         /* lp_run_batch through a captured HIP graph (1, default)
                                   or eager launches (0)                           */
#define LP_OPT_CLEAR_STEP 14    /* check a step that leaves its free region against the
                                   near-grid cell clearance when it is at least this
                                   long (0, default: every step; < 0: never)        */
#define LP_OPT_LIST_LONG 16     /* steps longer than the near-face radius try the lists
                                   first (a face within the radius is the nearest)
                                   before the BVH (1, default) or go to the BVH at
                                   once (0)                                         */
#define LP_OPT_BOUNDARY_RUN 15  /* end a photon as capped after this many boundary events
                                   in a row, no interaction between (100 000, default;
                                   0: no limit).  No reference counterpart: a photon
                                   caught on the edge of two faces crosses from one to
                                   the other at t ~ 0 for ever, as the reference's loop
                                   would; this ends it (and changes a result) only
                                   after that many crossings                        */
#define LP_OPT_NO_SKIP 17      /* 1: search every step for its face (no free region, cell
                                   clearance, safety grid or distance query); the
                                   lists and the BVH still answer the searches.
                                   For tests: the exhaustive form of the face
                                   search, whose results the shortcuts must equal */
#define LP_OPT_MARCH 18         /* a step longer than the near-face radius (or a flight)
                                   whose face the lists do not find within the radius
                                   marches on along its ray through the near-face grid
                                   -- cell clearances over empty space, list queries
                                   near faces -- (1, default) or searches the BVH (0) */
#define LP_OPT_SECOND_STEP 19   /* extra steps per wavefront iteration (0..8, default
                                   3): a lane whose step was an interaction draws its
                                   next step at once and takes it there when it stays
                                   in its free region, up to this many times; else the
                                   drawn step waits for the next iteration           */
#define LP_OPT_CRAWL 20         /* replay edge crawls (1, default): a photon on the
                                   shared edge of two faces, crossing them in turn
                                   without moving, takes the repeats of its last few
                                   crossings from per-lane records within one
                                   iteration (same draws, states and tallies) instead
                                   of one crossing per wavefront iteration; 0: off  */
int lp_set_option(lp_context *ctx, int32_t key, double value);

/* ---- batched launch (replaces ExecKernel's loop of 6 400-photon launches,
 * CalculatePropagation.cpp:1446-1488, when many optical-property sets share
 * one mesh) ----
 * Runs photons [first, first + count) of each of ninst instances of the
 * scene in one persistent launch.  Instance i has the tissue optics
 * nmuasg[i][ntissue][4] = (n, mua, mus, g) (ntissue and the fetal id as set
 * by lp_set_materials; Rsp per instance) and draws photon j from the same
 * (seed, j) stream as lp_run, so instance i's detector and region tallies
 * equal those of lp_set_materials(instance i) + lp_run bit for bit.
 * Outputs, 2^32 fixed point, either may be NULL: det[ninst][ndet][2] (R, fetal
 * R, as LP_TALLY_DET), region[ninst][ntissue] (as LP_TALLY_REGION).  The
 * context's own tallies are untouched; event counters add to LP_TALLY_MISC.
 * The launch (tally clears + transport) is captured once as a HIP graph and
 * replayed while the scene, options and (ninst, first, count) stay the same;
 * new optics only refill the device table.  ninst * (ntissue + 1) <= 2048. */
int lp_run_batch(lp_context *ctx, const float *nmuasg, int32_t ninst, int64_t first, int64_t count, int64_t *det,
                 int64_t *region);

/* lp_run_batch over uneven geometries in the same launch (no reference
 * counterpart): instance i transports in geometry geom[i] of
 * lp_set_geometries (geom NULL: geometry 0), so its tallies equal those of
 * lp_run_batch of instance i on a mesh of that geometry alone, bit for bit. */
int lp_run_batch_geom(lp_context *ctx, const float *nmuasg, int32_t ninst, const int32_t *geom, int64_t first,
                      int64_t count, int64_t *det, int64_t *region);

/* lp_run_batch_geom that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The scene
 * must already be on the device (an lp_run or lp_run_batch after the last
 * change of mesh, source or search options; else LP_ESTATE).  Outputs go to
 * the caller's device buffers, cleared by the call: det_dev[ninst][ndet][2]
 * (may be NULL without detectors) and region_dev[ninst][ntissue], 2^32
 * fixed point, bit-equal to lp_run_batch_geom's det and region.  Event
 * counters are not kept.  The instance table and the work counter live in a
 * static device array of 256 + 16 * 2048 bytes per GPU (held until the
 * process ends), so ninst * (ntissue + 1) <= 2048 (else LP_EINVAL).  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * threads, other contexts) race.  Blocks until done. */
int lp_run_batch_noalloc(lp_context *ctx, const float *nmuasg, int32_t ninst, const int32_t *geom, int64_t first,
                         int64_t count, int64_t *det_dev, int64_t *region_dev);

/* Launch-option autotuner (no reference counterpart: the reference fixes
 * its launch at 6 400 work-items, CalculatePropagation.cpp:76): coordinate
 * search over waves per SIMD (the grid follows the variant's occupancy
 * unless LP_OPT_BLOCKS_PER_CU is set), free iterations, refill threshold and
 * the two search batch thresholds, for up to `passes` passes.
 * A setting's score is the marginal time of `sample` photons: a launch of
 * 2 sample photons less a launch of sample photons, so that the launch tail
 * cancels and the choice suits long launches.  Trial launches carry the
 * context's tallies (into stand-in buffers of the same sizes: the context's
 * tallies are untouched) and transport in geometry 0.  A setting must win by
 * 1 %, and the final choice must win by 1 % again when it and the starting
 * options are scored anew, alternately, twice each; else the starting
 * options stay.  Leaves the chosen options on the context.  The options
 * change how faces are found, never which, so results do not depend on the
 * outcome. */
typedef struct lp_tune_result {
    int64_t sample_photons;
    int32_t trials;
    double ms_before, ms_after; /* score (marginal ms of sample photons) of the initial and the chosen options */
    int32_t waves, blocks_per_cu, free_iters, refill_min; /* blocks_per_cu: in effect */
    double list_min, search_min;
} lp_tune_result;
int lp_autotune(lp_context *ctx, int64_t sample, int32_t passes, lp_tune_result *out);

#ifdef __cplusplus
}
#endif
#endif /* LIGHTPROP_H */
