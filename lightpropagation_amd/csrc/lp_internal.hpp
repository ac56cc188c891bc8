// lp_internal.hpp -- data layout shared by the host runtime and the kernels.
#pragma once
#include <cstring>
#include <stdint.h>

#include <string>
#include <vector>

namespace lp {

// Threaded (stackless) BVH node, 32 bytes.  Traversal starts at the root
// (node 0 for a single geometry; MultiMesh::root for several) and follows
// `hit` when the ray segment meets the (padded) box and `miss` otherwise;
// miss == -1 ends the walk.  For a leaf, hit < 0 encodes the
// triangle range: first = hit & 0xffffff, count = (hit >> 24) & 0x7f, and
// after the leaf's triangles are tested the walk continues at `miss`.
struct BvhNode {
    float lo[3];
    int32_t hit;
    float hi[3];
    int32_t miss;
};
static_assert(sizeof(BvhNode) == 32, "node layout");
constexpr int kMaxLeaf = 127;

// Leaf-ordered triangle record, 48 bytes: (M, original id), (E = K - M,
// packed inner | outer << 16), (F = L - M, 0).  E and F are the exact float
// differences FindTriangleInPath forms (MCLightPropagation.cl:1058-1060).
struct TriRec {
    float m[3];
    int32_t id;
    float e[3];
    int32_t med;
    float f[3];
    int32_t pad;
};
static_assert(sizeof(TriRec) == 48, "triangle layout");

struct Bvh {
    std::vector<BvhNode> nodes;  // first 2048 breadth first (top levels), then subtrees depth first; root at 0
    std::vector<TriRec> tris;    // leaf order
    int32_t depth = 0;
    float pad = 0.0f;
};

// Builds a binned-SAH BVH over ntri triangles (9 floats each).  Child boxes
// are widened by `pad` (absolute) so that culling stays conservative with
// respect to the rounding of the exact triangle test.
Bvh build_bvh(const float* tri, const int32_t* med, int64_t ntri, int max_leaf = 4);

// Conservative lower bound of the distance from any point of a grid cell to
// the mesh, used to skip the triangle search for steps that cannot reach a
// surface.  Cells hold a float distance; 0 where a surface is close.
struct SafetyGrid {
    float origin[3] = {0, 0, 0};
    float inv_cell[3] = {0, 0, 0};
    int32_t n[3] = {0, 0, 0};
    std::vector<float> dist;
};
SafetyGrid build_safety_grid(const Bvh& bvh, const float* tri, int64_t ntri, int target_cells_per_axis);

// Near-face lists: a uniform grid whose cell c lists (leaf-order indices of)
// every face within r0 of some point that the kernel's float cell lookup can
// place in c (exact point-triangle distances from the cell centre, grown by
// the half diagonal and the rounding pad).  A nearest-face query of length
// <= r0 and a distance query capped at r0 are therefore answered exactly by
// one pass over that list.  The grid extends r0 beyond every face, so a point
// outside it is more than r0 from every face.
struct alignas(8) CellRec {
    uint32_t first;       // first list entry of the cell
    uint32_t clear_bits;  // float bits of the cell's clearance
};
struct NearGrid {
    float origin[3] = {0, 0, 0};
    float inv_cell[3] = {0, 0, 0};
    int32_t n[3] = {0, 0, 0};
    float r0 = 0.0f;
    // per cell c (and one sentinel): cells[2c] = first list entry of c,
    // cells[2c + 1] = bits of the float clearance of c: no face lies within
    // that distance of any point of c (from a Chebyshev distance transform
    // over the cells the faces pass through), so faces missing from c's list
    // are farther than max(r0, clearance)
    std::vector<uint32_t> cells;
    std::vector<int32_t> list;
    uint32_t start(int64_t c) const { return cells[2 * c]; }
    float clear(int64_t c) const
    {
        float f;
        std::memcpy(&f, &cells[2 * c + 1], sizeof f);
        return f;
    }
};
// Default cell edge of the near-face grid in units of r0: a list query costs
// a fixed chain of dependent loads and lane exchanges per 64 (photon, face)
// pairs, so shorter lists (smaller cells) pay until the grid outgrows the
// caches (measured on the maternal-fetal mesh: 0.1 cm cells at r0 = 0.07 cm
// are 3 % faster than 0.14 cm).
constexpr float kNearCellPerR0 = 10.0f / 7.0f;
// A frame given, the grid takes its origin and cell counts (it must reach
// r0 + cell beyond every face of bvh, as near_grid_frame makes it);
// otherwise the grid's own frame is fitted to the faces.
NearGrid build_near_grid(const Bvh& bvh, float r0, float cell, const NearGrid* frame = nullptr);
// Frame of a near grid over faces inside the box [lo, hi] whose BVH pads are
// at most pad: what build_near_grid fits for such faces.
NearGrid near_grid_frame(const float lo[3], const float hi[3], float r0, float cell, float pad);

// Several geometries in one set of arrays: part g is the triangles
// [first[g], first[g + 1]) with a BVH and near-face grid of its own.  The
// merged BVH holds the top nodes of every part first (each part's first
// min(n_g, lds_per_part) nodes, breadth first, so one LDS stage covers the
// top of every tree), then the rest of each part; root[g] is where part g's
// walk starts.  Triangle ids and leaf ranges are global.  The parts' near
// grids share one frame (fitted to all faces) and are stored part after
// part, each with its sentinel: part g's cell c is cell coff[g] + c of the
// merged grid, whose lists index the merged leaf order.
struct MultiMesh {
    Bvh bvh;
    NearGrid ng;
    std::vector<int32_t> root;
    std::vector<uint32_t> coff;
};
// Returns false (with a reason in why) when the merged arrays would overflow
// their index fields.
bool build_multi_mesh(const float* tri, const int32_t* med, const std::vector<int64_t>& first, int lds_per_part,
                      float r0, float cell, MultiMesh& out, std::string& why);

constexpr int kMaxTissues = 64;
constexpr int kMaxDetectors = 64;
constexpr int kBlock = 256;

// Kernel arguments (passed by value).
struct KArgs {
    // scene
    const TriRec* tris;
    const BvhNode* nodes;
    int32_t nnodes;
    int32_t nlds;  // nodes [0, nlds) are staged in LDS
    const float* nrm;     // 3 per original triangle
    const int32_t* med;   // 2 per original triangle
    int32_t ntri;
    const float* safety;  // optional safety grid
    float sg_origin[3], sg_inv[3];
    int32_t sg_n[3];
    const CellRec* near_cells;  // near-face lists (null = off)
    const TriRec* near_tris;  // the triangle records of every cell's list, cell after cell
    float ng_origin[3], ng_inv[3];
    int32_t ng_n[3];
    float near_r0;
    float dist_cap;        // per-photon free-radius search: cap (0 = off)
    float dist_cap_steps;  // ... and cap in units of the current step length
    float safe_margin;     // absolute rounding margin of the free radius
    int32_t free_iters;    // wavefront iterations per round (batches forced, lanes refilled at its end)
    float search_min;      // run a BVH search batch when this fraction of live lanes waits for one
    float list_min;        // run a near-list search batch when this fraction of live lanes waits for one
    float clear_step;      // cell-clearance free check for steps of at least this length (< 0: off)
    int32_t list_long;     // steps longer than near_r0 try the near-face lists (a face within near_r0) before the BVH
    int32_t no_skip;       // 1: search every step (no free region, clearance, safety grid or distance query)
    int32_t march;         // 1: a long step the lists miss marches on through the grid (cell clearances and
                           // further list queries) instead of searching the BVH
    int32_t second_step;   // extra steps a lane that took an interaction may take in the same iteration
    int32_t refill_min;    // end the round once this many lanes of a wavefront have finished
    int32_t waves;         // occupancy target of the kernel variant (1, 4, 5, 6)
    float nmuasg[4 * kMaxTissues];
    int32_t ntissue, fetal_id;
    float pstart[3];
    const float* vstart;
    const int32_t* vocc;
    int32_t nstart;
    float poi[3], ex[3], ey[3], ez[3];  // eX, eY, eZ already normalised
    float ez_raw[3];                    // eZ as given (sign of the near-axis override)
    float det_grid[6];
    float rsp, dist_min;
    int32_t dimx, dimy, dimz, dimR, dimAlpha, sft_nx, sft_ny, arz_nr, arz_nz;
    float factor, chance, w_th;
    uint64_t seed;
    int64_t max_steps;
    uint32_t max_bnd_run;  // boundary events in a row that end a photon as capped (0: no limit)
    int32_t rng_mode;
    // tallies (null = off)
    unsigned long long *sft, *rtd, *arz, *phirz, *region, *vox, *rad, *det, *misc;
    float vox_origin[3], vox_size[3], vox_inv[3];  // vox_inv = 1 / vox_size in float
    int32_t vox_n[3];
    float vox_nf[3];  // vox_n as floats
    float rad_dr;
    int32_t rad_n;
    int32_t ndet;
    float det_rlo[kMaxDetectors], det_rhi[kMaxDetectors];
    // per-photon outputs (null = off)
    float* exit_rec;
    float* trace;
    int32_t* trace_n;
    int32_t trace_max;
    // detected photons (forward launch; null = off): one record per (photon,
    // detector) pair whose exit radius lies in the detector's annulus
    struct DetRec* det_list;
    unsigned long long* det_count;
    int64_t det_cap;
    // adjoint replay (the ADJ kernel): photons come from adj_list[0, count);
    // gradients by the path-length estimator, 2^32 fixed point
    const struct DetRec* adj_list;
    unsigned long long* grad;  // ndet * ntissue * 2: (d/dmua, d/dmus)
    unsigned long long* jac;   // ndet * jac voxels (null = no voxel Jacobian)
    unsigned long long* jac_coll;  // ndet * jac voxels: W / mus of each interaction in the voxel
                                   // (null = none); d/dmus of voxel v = jac[v] + jac_coll[v]
    float jac_origin[3], jac_inv[3];  // jac_inv = 1 / voxel size in float
    int32_t jac_n[3];
    // per-pair path records of the replay (null = none): the path length and
    // interaction count in each tissue of the photon of adj_list[k], at
    // [k * ntissue + t]
    float* pair_len;
    uint32_t* pair_int;
    // batched launch (inst_mat null = one instance with nmuasg and rsp):
    // ninst instances differing in their tissue optics, instance i's
    // materials at inst_mat[i * (ntissue + 1) + t] and its (rsp, root, coff,
    // 0) at [i * (ntissue + 1) + ntissue] (root, coff: int bits of the BVH
    // root and near-grid cell offset of its geometry, read when ngeom > 1);
    // photon slot k of [0, count) is index first + k % inst_count of
    // instance k / inst_count; det and region tallies are per instance
    const float* inst_mat;  // 4 floats per row, 16-byte aligned
    int32_t ninst;
    int32_t ngeom;  // > 1: instances select their geometry (root, coff); else every walk starts at node 0, cell 0
    int32_t lmat_n;  // float4 rows of the material table in LDS (>= ninst * (ntissue + 1))
    int64_t inst_count;
    // work
    int64_t first, count;
    unsigned long long* counter;
    // per-lane rings of the last kCrawlSlots crossings that left p unchanged
    // (null: the crawl replay is off); lanes at or past crawl_lanes do without
    float* crawl;
    int64_t crawl_lanes;
};

// A detected photon: absolute photon index, detector, R exit weight.
struct DetRec {
    int64_t photon;
    int32_t det;
    float w;
};
static_assert(sizeof(DetRec) == 16, "detection record layout");

constexpr int kAdjMaxTissues = 16;  // tissues the adjoint replay tracks per lane (LDS)
// A crossing record of an edge crawl, 32-bit words: photon index (lo, hi),
// step number, the input direction (3), last face and tissue, then the
// outcome: face, t, Fresnel coefficient, output direction (3), tissue, normal
// code, kind (1 transmit, 2 reflect), and mua + mus of the input tissue.
// Each lane keeps the records of its last kCrawlSlots steps (a ring indexed
// by step number).
constexpr int kCrawlWords = 18;
constexpr int kCrawlSlots = 4;  // a power of 2

enum MiscSlot {  // slots of the misc tally (kMiscN int64 words)
    kMiscPhotons = 0, kMiscSteps, kMiscInteract, kMiscBoundary, kMiscEscaped, kMiscCapped,
    kMiscArzOut, kMiscVoxOut, kMiscNodeVisits, kMiscTriTests, kMiscSkipped, kMiscNan, kMiscDistQueries,
    kMiscDistVisits, kMiscListTests, kMiscIters, kMiscRounds, kMiscBvhRounds, kMiscReplayMismatch, kMiscMarches,
    kMiscCrawl, kMiscUsed,
    kMiscN = 32
};

static_assert(kMiscUsed <= kMiscN, "misc tally too small for its counters");

// Launchers (defined in lp_kernel.hip).
int launch_transport(const KArgs& a, int grid, void* stream, char* err, int errlen);
// work-groups of the transport variant these arguments select that one CU
// holds at once (registers and LDS), at least 1
int transport_resident_blocks(const KArgs& a);
int launch_debug_math(int op, const float* in, float* out, int64_t n, void* stream);
int launch_debug_copy(const void* src, void* dst, int64_t bytes, void* stream);
int launch_zero_u64(void* p, int64_t n, void* stream);
// rows of the per-instance material table a batched launch may stage in LDS
constexpr int kMaxInstRows = 2048;
// BVH nodes staged in LDS by default in a mesh of several geometries (split
// evenly between their trees); 512 with one geometry
constexpr int kMultiLdsNodes = 256;
// geometries one mesh may hold (lp_set_geometries)
constexpr int kMaxGeoms = 64;

// Static scratch of the allocation-free batched launch, one per GPU for the
// life of the process: [0, kScratchCounterBytes) the work counter, then the
// instance material table (kMaxInstRows float4 rows).
constexpr size_t kScratchCounterBytes = 256;
constexpr size_t kStaticScratchBytes = kScratchCounterBytes + 16 * (size_t)kMaxInstRows;
constexpr int kMaxDevices = 64;
// Device address of that scratch on `device`, which must be current; looked
// up once per device and kept.  Returns nonzero on failure.
int static_scratch(int device, void** out);

}  // namespace lp
