// lp_capi.cpp -- host runtime behind include/lightprop.h.
//
// Owns the device copies of the scene (leaf-ordered triangles, threaded
// BVH, safety grid, normals, tissue ids), the tallies and the launch.  It
// replaces the OpenCL host runtime of src/Simulation/CalculatePropagation.cpp
// (buffer packing :1167-1222, the 6400-photon launch/wait/copy loop with host
// binning :1389-1759) by one persistent launch that bins on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lightprop.h"
#include "lp_internal.hpp"

using namespace lp;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct lp_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err = "";
    int num_cus = 0;

    // host state
    std::vector<float> tri, nrm;
    std::vector<int32_t> med;
    int64_t ntri = -1;
    // geometries: triangles [part_first[g], part_first[g + 1]) (empty: one)
    std::vector<int64_t> part_first;
    std::vector<float> nmuasg;
    int32_t ntissue = 0, fetal_id = -1;
    float pstart[3] = {0, 0, 0};
    std::vector<float> vstart;
    std::vector<int32_t> vocc;
    float poi_basis[12] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};
    bool have_source = false, have_grid = false;
    float model_dims[3] = {1, 1, 1}, det_grid[6] = {1, 1, 1, 1, 1, 1};
    float chance = 0.1f, w_th = 1e-4f, factor = 1.0f;
    uint64_t seed = 1;
    int32_t rng_mode = 0;
    int64_t max_steps = 1000000;
    int32_t max_bnd_run = 100000;
    float vox_origin[3] = {0, 0, 0}, vox_size[3] = {1, 1, 1};
    int32_t vox_n[3] = {0, 0, 0};
    float rad_dr = 0.05f;
    int32_t rad_n = 0;
    std::vector<float> det_rlo, det_rhi;
    uint32_t tally_mask = 0xffffffffu;
    int32_t record_exit = 0;
    int32_t grid_override = 0, lds_override = -1, safety_cells = -1;
    float dist_cap = 0.2f, dist_cap_steps = 16.0f;
    int32_t blocks_per_cu = 0;  // 0: as many as one CU holds at once (the variant's occupancy)
    int32_t free_iters = 32;
    float search_min = 1.0f;
    float list_min = 0.7f;
    int32_t refill_min = 16;
    float clear_step = 0.0f;
    int32_t list_long = 1;
    int32_t no_skip = 0;
    int32_t march = 1;
    int32_t second_step = 3;
    int32_t crawl = 1;  // replay edge crawls from the lanes' crossing records
    int32_t waves = 4;
    float bvh_extent = 0.0f;  // largest |coordinate| of the mesh
    float jac_origin[3] = {0, 0, 0}, jac_size[3] = {1, 1, 1};
    int32_t jac_n[3] = {0, 0, 0};

    // derived
    lp_layout L{};
    Bvh bvh;
    SafetyGrid sg;
    NearGrid ng;
    float near_r0 = 0.07f, near_cell = 0.0f;  // cell 0 = kNearCellPerR0 r0
    std::vector<int32_t> part_root;   // BVH root of each geometry
    std::vector<uint32_t> part_coff;  // near-grid cell offset of each geometry
    DevBuf d_near_cells, d_near_tris;
    bool mesh_dirty = true, src_dirty = true, tallies_dirty = true;

    // device
    DevBuf d_tris, d_nodes, d_nrm, d_med, d_safety, d_vstart, d_vocc, d_counter, d_exit;
    DevBuf d_crawl;  // per-lane records of edge-crawl crossings (kCrawlSlots x kCrawlWords words per lane)
    int64_t crawl_lanes = 0;
    DevBuf d_tally[LP_NTALLY];
    DevBuf d_tune[LP_NTALLY];  // the autotuner's stand-ins for the tallies (same sizes)
    DevBuf d_det_list, d_det_count, d_grad, d_jac, d_jac_coll, d_adj_misc;
    int64_t exit_count = 0;
    // per-pair path records of the last lp_adjoint (lp_set_adjoint_paths)
    int32_t adj_paths = 0;
    DevBuf d_pair_len, d_pair_int;
    std::vector<DetRec> pair_list;  // the replayed pairs, in the order of the records
    int32_t pair_nt = 0;
    lp_stats stats{};

    // batched launches: instance material table, per-instance tallies, and
    // the captured graph (zero the tallies and the work counter, transport)
    // with the kernel arguments and grid it was captured for
    DevBuf d_inst_mat, d_bdet, d_bregion;
    int32_t use_graph = 1;
    void* scratch = nullptr;  // this device's static scratch (lp_run_batch_noalloc)
    hipGraph_t bgraph = nullptr;
    hipGraphExec_t bexec = nullptr;
    KArgs bargs{};
    int bgrid = 0;
};

namespace {

int fail(lp_context* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) return fail((c), LP_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

int dev_alloc(lp_context* c, DevBuf& b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (b.p && b.bytes >= bytes) return LP_OK;
    if (b.p) {
        (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(c, LP_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    b.bytes = bytes;
    return LP_OK;
}

void dev_free(DevBuf& b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

int upload(lp_context* c, DevBuf& b, const void* src, size_t bytes)
{
    int rc = dev_alloc(c, b, bytes);
    if (rc) return rc;
    if (bytes) HIPCHK(c, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    return LP_OK;
}

// x86 cvtt semantics of the reference host's int()
int32_t cvt_host(double x)
{
    if (!(x > -2147483649.0 && x < 2147483648.0)) return INT32_MIN;
    return (int32_t)x;
}

// normalize(p) = p * (1/sqrt(dot(p,p))) in float, as on the device
void normalize3(const float* in, float* out)
{
    float d = (in[0] * in[0] + in[1] * in[1]) + in[2] * in[2];
    if (d == 0.0f) {
        std::memcpy(out, in, 3 * sizeof(float));
        return;
    }
    float inv = 1.0f / std::sqrt(d);
    for (int k = 0; k < 3; ++k) out[k] = inv * in[k];
}

// Rspecular (MCLightPropagation.cl:273-299)
float rspecular(const float* nm, int32_t ntissue)
{
    float n1 = nm[0], n2 = nm[4];
    float temp = (n1 - n2) / (n1 + n2);
    float r1 = temp * temp;
    if (nm[5] == 0.0f && nm[6] == 0.0f && ntissue > 2) {
        float n3 = nm[8];
        temp = (n2 - n3) / (n2 + n3);
        float r2 = temp * temp;
        r1 = r1 + (1 - r1) * (1 - r1) * r2 / (1 - r1 * r2);
    }
    return r1;
}

void compute_layout(lp_context* c)
{
    lp_layout& L = c->L;
    std::memset(&L, 0, sizeof L);
    L.dimx = cvt_host(c->model_dims[0] / c->det_grid[0]);
    L.dimy = cvt_host(c->model_dims[1] / c->det_grid[1]);
    L.dimz = cvt_host(c->model_dims[2] / c->det_grid[2]);
    L.dimAlpha = cvt_host(90 / c->det_grid[5]);
    L.dimR = (int32_t)std::ceil(std::sqrt(std::pow((double)(L.dimx / 2), 2) + std::pow((double)(L.dimy / 2), 2)));
    L.sft_nx = 2 * (L.dimx / 2) + 1;
    L.sft_ny = 2 * (L.dimy / 2) + 1;
    L.arz_nr = 2 * L.dimR + 1;
    L.arz_nz = L.dimz + 1;
    L.dist_min = c->det_grid[2] * (float)L.dimz / 2.0f;
    L.rsp = c->ntissue >= 2 ? rspecular(c->nmuasg.data(), c->ntissue) : 0.0f;
    L.n[LP_TALLY_SFT] = (int64_t)L.sft_nx * L.sft_ny * 3;
    L.n[LP_TALLY_RTD] = (L.dimR > 0 && L.dimAlpha > 0) ? (int64_t)L.dimR * L.dimAlpha * 3 : 0;
    L.n[LP_TALLY_ARZ] = (int64_t)L.arz_nr * L.arz_nz;
    L.n[LP_TALLY_PHIRZ] = L.n[LP_TALLY_ARZ];
    L.n[LP_TALLY_REGION] = c->ntissue;
    L.n[LP_TALLY_VOX] = (c->vox_n[0] > 0 && c->vox_n[1] > 0 && c->vox_n[2] > 0)
                            ? (int64_t)c->vox_n[0] * c->vox_n[1] * c->vox_n[2] : 0;
    L.n[LP_TALLY_RAD] = c->rad_n > 0 ? (int64_t)c->rad_n * 3 : 0;
    L.n[LP_TALLY_DET] = (int64_t)c->det_rlo.size() * 2;
    L.n[LP_TALLY_MISC] = kMiscN;
}

bool tally_on(const lp_context* c, int k)
{
    if (k == LP_TALLY_MISC || k == LP_TALLY_REGION) return true;
    return (c->tally_mask >> k) & 1u;
}

int check_ready(lp_context* c)
{
    if (c->ntri < 0) return fail(c, LP_ESTATE, "no mesh (lp_set_mesh)");
    if (c->ntissue < 2) return fail(c, LP_ESTATE, "no materials (lp_set_materials)");
    if (!c->have_source) return fail(c, LP_ESTATE, "no source (lp_set_source)");
    for (int64_t t = 0; t < c->ntri; ++t)
        for (int s = 0; s < 2; ++s)
            if (c->med[2 * t + s] < 0 || c->med[2 * t + s] >= c->ntissue)
                return fail(c, LP_EINVAL, "triangle %lld has tissue id %d outside [0, %d)", (long long)t,
                            c->med[2 * t + s], c->ntissue);
    return LP_OK;
}

int prepare(lp_context* c)
{
    int rc = check_ready(c);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->mesh_dirty) {
        c->bvh_extent = 0.0f;
        for (float x : c->tri) c->bvh_extent = std::max(c->bvh_extent, std::fabs(x));
        int cells = c->safety_cells < 0 ? 0 : c->safety_cells;  // off unless asked: the free radius covers it
        const float cell = c->near_cell > 0 ? c->near_cell : kNearCellPerR0 * c->near_r0;
        if (c->part_first.size() > 2) {
            if (cells > 0) return fail(c, LP_EINVAL, "the safety grid needs a single geometry");
            const int ng = (int)c->part_first.size() - 1;
            const int lds = c->lds_override >= 0 ? c->lds_override : kMultiLdsNodes;
            MultiMesh mm;
            std::string why;
            if (!build_multi_mesh(c->tri.data(), c->med.data(), c->part_first, std::max(1, lds / ng), c->near_r0, cell,
                                  mm, why))
                return fail(c, LP_EINVAL, "lp_set_geometries: %s", why.c_str());
            c->bvh = std::move(mm.bvh);
            c->ng = std::move(mm.ng);
            c->part_root = std::move(mm.root);
            c->part_coff = std::move(mm.coff);
            c->sg = SafetyGrid();
        } else {
            c->bvh = build_bvh(c->tri.data(), c->med.data(), c->ntri, 4);
            c->sg = cells > 0 ? build_safety_grid(c->bvh, c->tri.data(), c->ntri, cells) : SafetyGrid();
            c->ng = c->near_r0 > 0 ? build_near_grid(c->bvh, c->near_r0, cell) : NearGrid();
            c->part_root.assign(1, 0);
            c->part_coff.assign(1, 0u);
        }
        if (!c->ng.list.empty()) {
            if ((rc = upload(c, c->d_near_cells, c->ng.cells.data(), c->ng.cells.size() * sizeof(uint32_t)))) return rc;
            // the lists hold copies of the triangle records, so a query reads
            // one contiguous run of records instead of chasing face indices
            std::vector<TriRec> recs(c->ng.list.size());
            for (size_t k = 0; k < recs.size(); ++k) recs[k] = c->bvh.tris[c->ng.list[k]];
            if ((rc = upload(c, c->d_near_tris, recs.data(), recs.size() * sizeof(TriRec)))) return rc;
        }
        if ((rc = upload(c, c->d_tris, c->bvh.tris.data(), c->bvh.tris.size() * sizeof(TriRec)))) return rc;
        if ((rc = upload(c, c->d_nodes, c->bvh.nodes.data(), c->bvh.nodes.size() * sizeof(BvhNode)))) return rc;
        if ((rc = upload(c, c->d_nrm, c->nrm.data(), c->nrm.size() * sizeof(float)))) return rc;
        if ((rc = upload(c, c->d_med, c->med.data(), c->med.size() * sizeof(int32_t)))) return rc;
        if (!c->sg.dist.empty()) {
            if ((rc = upload(c, c->d_safety, c->sg.dist.data(), c->sg.dist.size() * sizeof(float)))) return rc;
        }
        c->mesh_dirty = false;
    }
    if (c->src_dirty) {
        if ((rc = upload(c, c->d_vstart, c->vstart.data(), c->vstart.size() * sizeof(float)))) return rc;
        if ((rc = upload(c, c->d_vocc, c->vocc.data(), c->vocc.size() * sizeof(int32_t)))) return rc;
        c->src_dirty = false;
    }
    compute_layout(c);
    if (c->tallies_dirty) {
        for (int k = 0; k < LP_NTALLY; ++k) {
            size_t bytes = (size_t)std::max<int64_t>(c->L.n[k], 1) * sizeof(int64_t);
            if ((rc = dev_alloc(c, c->d_tally[k], bytes))) return rc;
            HIPCHK(c, hipMemsetAsync(c->d_tally[k].p, 0, bytes, c->stream));
        }
        c->tallies_dirty = false;
    }
    if ((rc = dev_alloc(c, c->d_counter, 64))) return rc;
    if (!c->d_crawl.p) {  // room for every lane of the largest automatic grid (8 work-groups per CU)
        const int64_t lanes = (int64_t)std::max(c->num_cus, 1) * 8 * kBlock;
        if ((rc = dev_alloc(c, c->d_crawl, (size_t)lanes * kCrawlSlots * kCrawlWords * sizeof(float)))) return rc;
        c->crawl_lanes = lanes;
    }
    return LP_OK;
}

void fill_args(lp_context* c, KArgs& a)
{
    std::memset(&a, 0, sizeof a);
    a.tris = (const TriRec*)c->d_tris.p;
    a.nodes = (const BvhNode*)c->d_nodes.p;
    a.crawl = c->crawl ? (float*)c->d_crawl.p : nullptr;
    a.crawl_lanes = c->crawl_lanes;
    a.nnodes = (int32_t)c->bvh.nodes.size();
    // a launch over several geometries stages fewer BVH nodes, leaving LDS
    // for the larger material tables of its many instances
    int lds = c->lds_override >= 0 ? c->lds_override : (c->part_root.size() > 1 ? kMultiLdsNodes : 512);
    a.nlds = std::min<int32_t>(a.nnodes, lds);
    a.nrm = (const float*)c->d_nrm.p;
    a.med = (const int32_t*)c->d_med.p;
    a.ntri = (int32_t)c->ntri;
    if (!c->ng.list.empty()) {
        a.near_cells = (const CellRec*)c->d_near_cells.p;
        a.near_tris = (const TriRec*)c->d_near_tris.p;
        for (int k = 0; k < 3; ++k) {
            a.ng_origin[k] = c->ng.origin[k];
            a.ng_inv[k] = c->ng.inv_cell[k];
            a.ng_n[k] = c->ng.n[k];
        }
        a.near_r0 = c->ng.r0;
    }
    a.dist_cap = c->dist_cap;
    a.dist_cap_steps = c->dist_cap_steps;
    a.free_iters = c->free_iters;
    a.search_min = c->search_min;
    a.list_min = c->list_min;
    a.clear_step = c->clear_step;
    a.list_long = c->list_long;
    a.no_skip = c->no_skip;
    a.march = c->march;
    a.second_step = c->second_step;
    a.refill_min = c->refill_min;
    a.waves = c->waves;
    a.safe_margin = 2e-6f * c->bvh_extent;
    if (!c->sg.dist.empty()) {
        a.safety = (const float*)c->d_safety.p;
        for (int k = 0; k < 3; ++k) {
            a.sg_origin[k] = c->sg.origin[k];
            a.sg_inv[k] = c->sg.inv_cell[k];
            a.sg_n[k] = c->sg.n[k];
        }
    }
    std::memcpy(a.nmuasg, c->nmuasg.data(), c->nmuasg.size() * sizeof(float));
    a.ntissue = c->ntissue;
    a.fetal_id = c->fetal_id;
    std::memcpy(a.pstart, c->pstart, sizeof a.pstart);
    a.vstart = (const float*)c->d_vstart.p;
    a.vocc = (const int32_t*)c->d_vocc.p;
    a.nstart = (int32_t)c->vocc.size();
    std::memcpy(a.poi, c->poi_basis, 3 * sizeof(float));
    normalize3(c->poi_basis + 3, a.ex);
    normalize3(c->poi_basis + 6, a.ey);
    normalize3(c->poi_basis + 9, a.ez);
    std::memcpy(a.ez_raw, c->poi_basis + 9, 3 * sizeof(float));
    std::memcpy(a.det_grid, c->det_grid, sizeof a.det_grid);
    a.rsp = c->L.rsp;
    a.dist_min = c->L.dist_min;
    a.dimx = c->L.dimx;
    a.dimy = c->L.dimy;
    a.dimz = c->L.dimz;
    a.dimR = c->L.dimR;
    a.dimAlpha = c->L.dimAlpha;
    a.sft_nx = c->L.sft_nx;
    a.sft_ny = c->L.sft_ny;
    a.arz_nr = c->L.arz_nr;
    a.arz_nz = c->L.arz_nz;
    a.factor = c->factor;
    a.chance = c->chance;
    a.w_th = c->w_th;
    a.seed = c->seed;
    a.max_steps = c->max_steps;
    a.max_bnd_run = (uint32_t)c->max_bnd_run;
    a.rng_mode = c->rng_mode;
    auto T = [&](int k) -> unsigned long long* {
        return (tally_on(c, k) && c->L.n[k] > 0) ? (unsigned long long*)c->d_tally[k].p : nullptr;
    };
    a.sft = T(LP_TALLY_SFT);
    a.rtd = T(LP_TALLY_RTD);
    a.arz = T(LP_TALLY_ARZ);
    a.phirz = a.arz ? (unsigned long long*)c->d_tally[LP_TALLY_PHIRZ].p : nullptr;
    a.region = T(LP_TALLY_REGION);
    a.vox = T(LP_TALLY_VOX);
    a.rad = T(LP_TALLY_RAD);
    a.det = c->det_rlo.empty() ? nullptr : T(LP_TALLY_DET);
    a.misc = (unsigned long long*)c->d_tally[LP_TALLY_MISC].p;
    std::memcpy(a.vox_origin, c->vox_origin, sizeof a.vox_origin);
    std::memcpy(a.vox_size, c->vox_size, sizeof a.vox_size);
    for (int k = 0; k < 3; ++k) a.vox_inv[k] = 1.0f / c->vox_size[k];
    std::memcpy(a.vox_n, c->vox_n, sizeof a.vox_n);
    for (int k = 0; k < 3; ++k) a.vox_nf[k] = (float)c->vox_n[k];
    a.rad_dr = c->rad_dr;
    a.rad_n = c->rad_n;
    a.ndet = (int32_t)c->det_rlo.size();
    for (int k = 0; k < a.ndet; ++k) {
        a.det_rlo[k] = c->det_rlo[k];
        a.det_rhi[k] = c->det_rhi[k];
    }
    a.counter = (unsigned long long*)c->d_counter.p;
    a.ninst = 1;
    a.ngeom = 1;
    a.lmat_n = kMaxTissues + 1;
}

// Work-groups per CU of a launch with these arguments: the option, or the
// number one CU holds at once, so that the persistent grid fills the chip
// at the occupancy the kernel variant allows.
int effective_blocks_per_cu(const lp_context* c, const KArgs& a)
{
    return c->blocks_per_cu > 0 ? c->blocks_per_cu : transport_resident_blocks(a);
}

int grid_for(const lp_context* c, const KArgs& a)
{
    int grid = c->grid_override;
    if (grid <= 0) grid = std::max(1, c->num_cus) * effective_blocks_per_cu(c, a);
    const int64_t count = a.count;
    int64_t max_grid = (count + kBlock - 1) / kBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(grid, std::max<int64_t>(max_grid, 1)));
}

int launch(lp_context* c, KArgs& a)
{
    const int grid = grid_for(c, a);
    HIPCHK(c, hipMemsetAsync(c->d_counter.p, 0, 64, c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    char err[256] = {0};
    if (launch_transport(a, grid, c->stream, err, sizeof err)) return fail(c, LP_EDEVICE, "%s", err);
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.kernel_ms = ms;
    c->stats.photons = a.count;
    c->stats.grid_blocks = grid;
    c->stats.bvh_nodes = (int64_t)c->bvh.nodes.size();
    c->stats.bvh_depth = c->bvh.depth;
    return LP_OK;
}

}  // namespace

extern "C" {

int lp_create(lp_context** out, int32_t device)
{
    if (!out) return LP_EINVAL;
    *out = nullptr;
    lp_context* c = new (std::nothrow) lp_context();
    if (!c) return LP_ENOMEM;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0 || device < 0 || device >= n) {
        delete c;
        return LP_EDEVICE;
    }
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return LP_EDEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    // the device address of this GPU's static scratch (device current)
    if (static_scratch(device, &c->scratch)) c->scratch = nullptr;
    *out = c;
    return LP_OK;
}

void lp_destroy(lp_context* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->bexec) (void)hipGraphExecDestroy(c->bexec);
    if (c->bgraph) (void)hipGraphDestroy(c->bgraph);
    DevBuf* bufs[] = {&c->d_tris, &c->d_nodes, &c->d_nrm, &c->d_med, &c->d_safety, &c->d_vstart,
                      &c->d_vocc, &c->d_counter, &c->d_exit, &c->d_near_cells, &c->d_near_tris, &c->d_det_list,
                      &c->d_det_count, &c->d_grad, &c->d_jac, &c->d_jac_coll, &c->d_adj_misc, &c->d_inst_mat,
                      &c->d_bdet, &c->d_bregion, &c->d_pair_len, &c->d_pair_int, &c->d_crawl};
    for (DevBuf* b : bufs) dev_free(*b);
    for (auto& b : c->d_tally) dev_free(b);
    for (auto& b : c->d_tune) dev_free(b);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* lp_last_error(const lp_context* c) { return c ? c->err.c_str() : "null context"; }

int lp_set_mesh(lp_context* c, const float* tri, const float* nrm, const int32_t* med, int64_t ntri)
{
    if (!c) return LP_EINVAL;
    if (ntri < 0 || ntri >= (1 << 24) || (ntri > 0 && (!tri || !nrm || !med)))
        return fail(c, LP_EINVAL, "lp_set_mesh: bad arguments (ntri=%lld)", (long long)ntri);
    c->tri.assign(tri, tri + 9 * ntri);
    c->nrm.assign(nrm, nrm + 3 * ntri);
    c->med.assign(med, med + 2 * ntri);
    for (int64_t i = 0; i < 2 * ntri; ++i)
        if (med[i] < 0 || med[i] > 0xffff) return fail(c, LP_EINVAL, "lp_set_mesh: tissue id out of range");
    c->ntri = ntri;
    c->part_first.clear();
    c->mesh_dirty = true;
    return LP_OK;
}

int lp_set_geometries(lp_context* c, const int64_t* tri_count, int32_t ngeom)
{
    if (!c) return LP_EINVAL;
    if (c->ntri < 0) return fail(c, LP_ESTATE, "lp_set_geometries: no mesh (lp_set_mesh)");
    if (!tri_count || ngeom < 1 || ngeom > kMaxGeoms)
        return fail(c, LP_EINVAL, "lp_set_geometries: need 1 to %d geometries", kMaxGeoms);
    std::vector<int64_t> f(1, 0);
    for (int32_t g = 0; g < ngeom; ++g) {
        if (tri_count[g] < 1) return fail(c, LP_EINVAL, "lp_set_geometries: geometry %d has no triangles", g);
        f.push_back(f.back() + tri_count[g]);
    }
    if (f.back() != c->ntri)
        return fail(c, LP_EINVAL, "lp_set_geometries: the counts add up to %lld of the mesh's %lld triangles",
                    (long long)f.back(), (long long)c->ntri);
    if (ngeom == 1) f.clear();
    if (f != c->part_first) c->mesh_dirty = true;
    c->part_first = std::move(f);
    return LP_OK;
}

int lp_set_materials(lp_context* c, const float* nmuasg, int32_t ntissue, int32_t fetal_id)
{
    if (!c) return LP_EINVAL;
    if (!nmuasg || ntissue < 2 || ntissue > kMaxTissues)
        return fail(c, LP_EINVAL, "lp_set_materials: need 2..%d tissues", kMaxTissues);
    c->nmuasg.assign(nmuasg, nmuasg + 4 * ntissue);
    c->ntissue = ntissue;
    c->fetal_id = fetal_id;
    c->tallies_dirty = true;
    return LP_OK;
}

int lp_set_source(lp_context* c, const float pstart[3], const float* vstart, const int32_t* vocc, int32_t nstart,
                  const float poi_basis[12])
{
    if (!c) return LP_EINVAL;
    if (!pstart || !vstart || !vocc || nstart < 1 || !poi_basis)
        return fail(c, LP_EINVAL, "lp_set_source: bad arguments");
    std::memcpy(c->pstart, pstart, 3 * sizeof(float));
    c->vstart.assign(vstart, vstart + 3 * nstart);
    c->vocc.assign(vocc, vocc + nstart);
    std::memcpy(c->poi_basis, poi_basis, 12 * sizeof(float));
    c->have_source = true;
    c->src_dirty = true;
    return LP_OK;
}

int lp_set_grid(lp_context* c, const float model_dims[3], const float det_grid[6])
{
    if (!c) return LP_EINVAL;
    if (!model_dims || !det_grid) return fail(c, LP_EINVAL, "lp_set_grid: null");
    for (int k = 0; k < 6; ++k)
        if (!(det_grid[k] > 0)) return fail(c, LP_EINVAL, "lp_set_grid: detection grid must be > 0");
    std::memcpy(c->model_dims, model_dims, 3 * sizeof(float));
    std::memcpy(c->det_grid, det_grid, 6 * sizeof(float));
    c->have_grid = true;
    c->tallies_dirty = true;
    return LP_OK;
}

int lp_set_params(lp_context* c, float chance, float w_th, float factor, uint64_t seed, int32_t rng_mode,
                  int64_t max_steps)
{
    if (!c) return LP_EINVAL;
    if (!(chance > 0 && chance <= 1) || !(w_th >= 0) || !(factor > 0) || (rng_mode != 0 && rng_mode != 1))
        return fail(c, LP_EINVAL, "lp_set_params: bad arguments");
    c->chance = chance;
    c->w_th = w_th;
    c->factor = factor;
    c->seed = seed;
    c->rng_mode = rng_mode;
    c->max_steps = max_steps > 0 ? std::min<int64_t>(max_steps, 0x7fffffff) : 1000000;  // steps count in 32 bits
    return LP_OK;
}

int lp_set_voxel_grid(lp_context* c, const float origin[3], const float size[3], const int32_t n[3])
{
    if (!c) return LP_EINVAL;
    if (!origin || !size || !n) return fail(c, LP_EINVAL, "lp_set_voxel_grid: null");
    for (int k = 0; k < 3; ++k)
        if (n[k] < 0 || n[k] > (1 << 24) || (n[k] > 0 && !(size[k] > 0)))
            return fail(c, LP_EINVAL, "lp_set_voxel_grid: bad size (at most 2^24 voxels per axis)");
    if ((int64_t)n[0] * n[1] * n[2] >= ((int64_t)1 << 31))
        return fail(c, LP_EINVAL, "lp_set_voxel_grid: fewer than 2^31 voxels");
    std::memcpy(c->vox_origin, origin, sizeof c->vox_origin);
    std::memcpy(c->vox_size, size, sizeof c->vox_size);
    std::memcpy(c->vox_n, n, sizeof c->vox_n);
    c->tallies_dirty = true;
    return LP_OK;
}

int lp_set_radial(lp_context* c, float dr, int32_t nbins)
{
    if (!c) return LP_EINVAL;
    if (nbins < 0 || (nbins > 0 && !(dr > 0))) return fail(c, LP_EINVAL, "lp_set_radial: bad arguments");
    c->rad_dr = dr;
    c->rad_n = nbins;
    c->tallies_dirty = true;
    return LP_OK;
}

int lp_set_detectors(lp_context* c, const float* rlo, const float* rhi, int32_t ndet)
{
    if (!c) return LP_EINVAL;
    if (ndet < 0 || ndet > kMaxDetectors || (ndet > 0 && (!rlo || !rhi)))
        return fail(c, LP_EINVAL, "lp_set_detectors: 0..%d detectors", kMaxDetectors);
    c->det_rlo.assign(rlo, rlo + ndet);
    c->det_rhi.assign(rhi, rhi + ndet);
    c->tallies_dirty = true;
    return LP_OK;
}

int lp_set_tallies(lp_context* c, uint32_t mask, int32_t record_exit)
{
    if (!c) return LP_EINVAL;
    c->tally_mask = mask;
    c->record_exit = record_exit;
    return LP_OK;
}

int lp_set_launch(lp_context* c, int32_t grid_blocks, int32_t lds_nodes, int32_t safety_cells)
{
    if (!c) return LP_EINVAL;
    if (lds_nodes > 2048) return fail(c, LP_EINVAL, "lp_set_launch: at most 2048 LDS nodes");
    c->grid_override = grid_blocks;
    c->lds_override = lds_nodes;
    if (safety_cells != c->safety_cells) c->mesh_dirty = true;
    c->safety_cells = safety_cells;
    return LP_OK;
}

int lp_set_option(lp_context* c, int32_t key, double v)
{
    if (!c) return LP_EINVAL;
    switch (key) {
    case LP_OPT_GRID_BLOCKS:
        if (v < 0 || v > 1e7) return fail(c, LP_EINVAL, "grid blocks out of range");
        c->grid_override = (int32_t)v;
        return LP_OK;
    case LP_OPT_LDS_NODES:
        if (v < -1 || v > 2048) return fail(c, LP_EINVAL, "LDS nodes out of range");
        c->lds_override = (int32_t)v;
        return LP_OK;
    case LP_OPT_SAFETY_CELLS:
        if (v < -1 || v > 1024) return fail(c, LP_EINVAL, "safety cells out of range");
        if ((int32_t)v != c->safety_cells) c->mesh_dirty = true;
        c->safety_cells = (int32_t)v;
        return LP_OK;
    case LP_OPT_DIST_CAP:
        if (!(v >= 0)) return fail(c, LP_EINVAL, "distance cap must be >= 0");
        c->dist_cap = (float)v;
        return LP_OK;
    case LP_OPT_DIST_CAP_STEPS:
        if (!(v >= 0)) return fail(c, LP_EINVAL, "distance cap steps must be >= 0");
        c->dist_cap_steps = (float)v;
        return LP_OK;
    case LP_OPT_FREE_ITERS:
        if (v < 1 || v > 4096) return fail(c, LP_EINVAL, "free iterations out of range");
        c->free_iters = (int32_t)v;
        return LP_OK;
    case LP_OPT_WAVES_PER_SIMD:
        if (!(v == 1 || v == 4 || v == 5 || v == 6)) return fail(c, LP_EINVAL, "waves per SIMD must be 1, 4, 5 or 6");
        c->waves = (int32_t)v;
        return LP_OK;
    case LP_OPT_NEAR_R0:
        if (!(v >= 0 && v < 1e6)) return fail(c, LP_EINVAL, "near radius must be >= 0");
        if ((float)v != c->near_r0) c->mesh_dirty = true;
        c->near_r0 = (float)v;
        return LP_OK;
    case LP_OPT_NEAR_CELL:
        if (!(v >= 0 && v < 1e6)) return fail(c, LP_EINVAL, "near cell must be >= 0");
        if ((float)v != c->near_cell) c->mesh_dirty = true;
        c->near_cell = (float)v;
        return LP_OK;
    case LP_OPT_CLEAR_STEP:
        if (!(v > -1e30 && v < 1e30)) return fail(c, LP_EINVAL, "clearance step out of range");
        c->clear_step = (float)v;
        return LP_OK;
    case LP_OPT_GRAPH:
        if (!(v == 0 || v == 1)) return fail(c, LP_EINVAL, "graph option must be 0 or 1");
        c->use_graph = (int32_t)v;
        return LP_OK;
    case LP_OPT_SEARCH_MIN:
        if (!(v >= 0 && v <= 1)) return fail(c, LP_EINVAL, "search fraction must be in [0, 1]");
        c->search_min = (float)v;
        return LP_OK;
    case LP_OPT_LIST_MIN:
        if (!(v >= 0 && v <= 1)) return fail(c, LP_EINVAL, "list search fraction must be in [0, 1]");
        c->list_min = (float)v;
        return LP_OK;
    case LP_OPT_REFILL_MIN:
        if (v < 1 || v > 64) return fail(c, LP_EINVAL, "refill lanes must be in [1, 64]");
        c->refill_min = (int32_t)v;
        return LP_OK;
    case LP_OPT_LIST_LONG:
        if (v != 0 && v != 1) return fail(c, LP_EINVAL, "list_long is 0 or 1");
        c->list_long = (int32_t)v;
        return LP_OK;
    case LP_OPT_MARCH:
        if (v != 0 && v != 1) return fail(c, LP_EINVAL, "march is 0 or 1");
        c->march = (int32_t)v;
        return LP_OK;
    case LP_OPT_SECOND_STEP:
        if (!(v >= 0 && v <= 8 && v == (int32_t)v)) return fail(c, LP_EINVAL, "second_step is 0..8");
        c->second_step = (int32_t)v;
        return LP_OK;
    case LP_OPT_CRAWL:
        if (!(v == 0 || v == 1)) return fail(c, LP_EINVAL, "crawl is 0 or 1");
        c->crawl = (int32_t)v;
        return LP_OK;
    case LP_OPT_NO_SKIP:
        if (v != 0 && v != 1) return fail(c, LP_EINVAL, "no_skip is 0 or 1");
        c->no_skip = (int32_t)v;
        return LP_OK;
    case LP_OPT_BOUNDARY_RUN:
        if (v < 0 || v > 1e9) return fail(c, LP_EINVAL, "boundary run must be in [0, 10^9] (0: no limit)");
        c->max_bnd_run = (int32_t)v;
        return LP_OK;
    case LP_OPT_BLOCKS_PER_CU:
        if (v < 0 || v > 32) return fail(c, LP_EINVAL, "blocks per CU out of range (0: automatic)");
        c->blocks_per_cu = (int32_t)v;
        return LP_OK;
    default:
        return fail(c, LP_EINVAL, "unknown option %d", key);
    }
}

int lp_get_layout(lp_context* c, lp_layout* out)
{
    if (!c || !out) return LP_EINVAL;
    compute_layout(c);
    *out = c->L;
    return LP_OK;
}

int lp_reset_tallies(lp_context* c)
{
    if (!c) return LP_EINVAL;
    c->tallies_dirty = true;
    return LP_OK;
}

int lp_run(lp_context* c, int64_t first, int64_t count)
{
    if (!c) return LP_EINVAL;
    if (first < 0 || count < 0) return fail(c, LP_EINVAL, "lp_run: negative range");
    int rc = prepare(c);
    if (rc) return rc;
    if (count == 0) return LP_OK;
    KArgs a;
    fill_args(c, a);
    a.first = first;
    a.count = count;
    if (c->record_exit) {
        if ((rc = dev_alloc(c, c->d_exit, (size_t)count * 8 * sizeof(float)))) return rc;
        a.exit_rec = (float*)c->d_exit.p;
        c->exit_count = count;
    } else {
        c->exit_count = 0;
    }
    return launch(c, a);
}

}  // extern "C"

namespace {

// Checks a batched launch's arguments and writes its material table: per
// instance the ntissue rows (n, mua, mus, g), then (rsp, BVH root bits, cell
// offset bits, 0) of its geometry; ninst * (ntissue + 1) float4 rows.
int build_inst_table(lp_context* c, const float* nmuasg, int32_t ninst, const int32_t* geom, int64_t count,
                     const char* who, float* tab)
{
    const int32_t ngeom = (int32_t)c->part_root.size();
    if (geom && ngeom > 1 && c->rng_mode != 0)
        return fail(c, LP_EINVAL, "%s: several geometries need the Philox streams (rng_mode 0)", who);
    if (geom)
        for (int32_t i = 0; i < ninst; ++i)
            if (geom[i] < 0 || geom[i] >= ngeom)
                return fail(c, LP_EINVAL, "%s: instance %d: geometry %d outside [0, %d)", who, i, geom[i], ngeom);
    const int32_t nt = c->ntissue;
    if ((int64_t)ninst * (nt + 1) > kMaxInstRows)
        return fail(c, LP_EINVAL, "%s: %d instances of %d tissues exceed the %d-row material table", who, ninst, nt,
                    kMaxInstRows);
    if (count > 0 && (int64_t)ninst > INT64_MAX / 2 / count) return fail(c, LP_EINVAL, "%s: too many photons", who);
    for (int32_t i = 0; i < ninst; ++i) {
        const float* m = nmuasg + (size_t)i * nt * 4;
        for (int32_t t = 0; t < nt; ++t)
            if (!(m[4 * t] > 0) || !(m[4 * t + 1] >= 0) || !(m[4 * t + 2] >= 0) || !(m[4 * t + 3] >= -1 && m[4 * t + 3] <= 1))
                return fail(c, LP_EINVAL, "%s: instance %d tissue %d: need n > 0, mua, mus >= 0, |g| <= 1", who, i, t);
        std::memcpy(&tab[(size_t)i * (nt + 1) * 4], m, (size_t)nt * 4 * sizeof(float));
        float* row = &tab[((size_t)i * (nt + 1) + nt) * 4];
        row[0] = rspecular(m, nt);
        row[3] = 0.0f;
        const int32_t g = geom ? geom[i] : 0;
        std::memcpy(&row[1], &c->part_root[g], sizeof(int32_t));
        std::memcpy(&row[2], &c->part_coff[g], sizeof(uint32_t));
    }
    return LP_OK;
}

// Captures [zero the instance tallies, zero the work counter, transport
// launch] on the context's stream as a graph, unless the graph held was
// captured for the same kernel arguments and grid.  Memory is cleared by a
// kernel node: memset nodes are not used in captured graphs.
int batch_graph(lp_context* c, const KArgs& a, int grid, int64_t ndet_words, int64_t nreg_words)
{
    if (c->bexec && c->bgrid == grid && std::memcmp(&c->bargs, &a, sizeof a) == 0) return LP_OK;
    if (c->bexec) (void)hipGraphExecDestroy(c->bexec);
    if (c->bgraph) (void)hipGraphDestroy(c->bgraph);
    c->bexec = nullptr;
    c->bgraph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    char err[256] = {0};
    int bad = 0;
    if (a.det) bad |= launch_zero_u64(a.det, ndet_words, c->stream);
    bad |= launch_zero_u64(a.region, nreg_words, c->stream);
    bad |= launch_zero_u64(a.counter, 8, c->stream);
    if (!bad) bad = launch_transport(a, grid, c->stream, err, sizeof err);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (bad || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return fail(c, LP_EDEVICE, "lp_run_batch: graph capture failed (%s)", err[0] ? err : hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(c, LP_EDEVICE, "lp_run_batch: graph instantiation: %s", hipGetErrorString(e));
    }
    c->bgraph = g;
    c->bexec = x;
    c->bargs = a;
    c->bgrid = grid;
    return LP_OK;
}

}  // namespace

extern "C" {

int lp_run_batch(lp_context* c, const float* nmuasg, int32_t ninst, int64_t first, int64_t count, int64_t* det,
                 int64_t* region)
{
    return lp_run_batch_geom(c, nmuasg, ninst, nullptr, first, count, det, region);
}

int lp_run_batch_geom(lp_context* c, const float* nmuasg, int32_t ninst, const int32_t* geom, int64_t first,
                      int64_t count, int64_t* det, int64_t* region)
{
    if (!c) return LP_EINVAL;
    if (!nmuasg || ninst < 1 || first < 0 || count < 0) return fail(c, LP_EINVAL, "lp_run_batch: bad arguments");
    int rc = prepare(c);
    if (rc) return rc;
    const int32_t nt = c->ntissue, nd = (int32_t)c->det_rlo.size();
    std::vector<float> tab((size_t)ninst * (nt + 1) * 4, 0.0f);
    if ((rc = build_inst_table(c, nmuasg, ninst, geom, count, "lp_run_batch", tab.data()))) return rc;
    const int32_t ngeom = (int32_t)c->part_root.size();
    const int64_t ndw = (int64_t)ninst * nd * 2, nrw = (int64_t)ninst * nt;
    if ((rc = dev_alloc(c, c->d_inst_mat, tab.size() * sizeof(float)))) return rc;
    if ((rc = dev_alloc(c, c->d_bdet, (size_t)std::max<int64_t>(ndw, 1) * sizeof(int64_t)))) return rc;
    if ((rc = dev_alloc(c, c->d_bregion, (size_t)nrw * sizeof(int64_t)))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_inst_mat.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    KArgs a;
    fill_args(c, a);
    a.sft = a.rtd = a.arz = a.phirz = a.vox = a.rad = nullptr;
    a.det = nd > 0 ? (unsigned long long*)c->d_bdet.p : nullptr;
    a.region = (unsigned long long*)c->d_bregion.p;
    a.inst_mat = (const float*)c->d_inst_mat.p;
    a.ninst = ninst;
    a.ngeom = geom ? ngeom : 1;
    a.lmat_n = std::max<int32_t>(kMaxTissues + 1, ninst * (nt + 1));
    a.inst_count = std::max<int64_t>(count, 1);
    a.first = first;
    a.count = (int64_t)ninst * count;
    const int grid = grid_for(c, a);
    if (a.count == 0) {
        if (det) std::memset(det, 0, (size_t)ndw * sizeof(int64_t));
        if (region) std::memset(region, 0, (size_t)nrw * sizeof(int64_t));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return LP_OK;
    }
    if (c->use_graph) {
        if ((rc = batch_graph(c, a, grid, ndw, nrw))) return rc;
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        HIPCHK(c, hipGraphLaunch(c->bexec, c->stream));
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    } else {
        if (a.det) HIPCHK(c, hipMemsetAsync(a.det, 0, (size_t)ndw * sizeof(int64_t), c->stream));
        HIPCHK(c, hipMemsetAsync(a.region, 0, (size_t)nrw * sizeof(int64_t), c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_counter.p, 0, 64, c->stream));
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        char err[256] = {0};
        if (launch_transport(a, grid, c->stream, err, sizeof err)) return fail(c, LP_EDEVICE, "%s", err);
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    }
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (det && ndw) HIPCHK(c, hipMemcpy(det, c->d_bdet.p, (size_t)ndw * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (region) HIPCHK(c, hipMemcpy(region, c->d_bregion.p, (size_t)nrw * sizeof(int64_t), hipMemcpyDeviceToHost));
    c->stats.kernel_ms = ms;
    c->stats.photons = a.count;
    c->stats.grid_blocks = grid;
    return LP_OK;
}

int lp_run_batch_noalloc(lp_context* c, const float* nmuasg, int32_t ninst, const int32_t* geom, int64_t first,
                         int64_t count, int64_t* det_dev, int64_t* region_dev)
{
    if (!c) return LP_EINVAL;
    if (!nmuasg || ninst < 1 || first < 0 || count < 0 || !region_dev)
        return fail(c, LP_EINVAL, "lp_run_batch_noalloc: bad arguments");
    if (!c->scratch) return fail(c, LP_EDEVICE, "lp_run_batch_noalloc: no static scratch on device %d", c->device);
    // nothing may be built or uploaded here: the scene must already be on the device
    if (c->ntri < 0 || c->mesh_dirty || c->src_dirty || !c->d_counter.p)
        return fail(c, LP_ESTATE,
                    "lp_run_batch_noalloc: the scene is not on the device (run lp_run or lp_run_batch once after "
                    "the last change of mesh, source or search options)");
    int rc = check_ready(c);
    if (rc) return rc;
    const int32_t nt = c->ntissue, nd = (int32_t)c->det_rlo.size();
    if (nd > 0 && !det_dev) return fail(c, LP_EINVAL, "lp_run_batch_noalloc: detectors are set but det_dev is NULL");
    if ((int64_t)ninst * (nt + 1) > kMaxInstRows)
        return fail(c, LP_EINVAL,
                    "lp_run_batch_noalloc: %d instances of %d tissues need %lld material rows; the static table "
                    "holds %d (%zu bytes)",
                    ninst, nt, (long long)ninst * (nt + 1), kMaxInstRows, (size_t)16 * kMaxInstRows);
    float tab[4 * kMaxInstRows];  // host staging of the table (stack, no allocation)
    if ((rc = build_inst_table(c, nmuasg, ninst, geom, count, "lp_run_batch_noalloc", tab))) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    compute_layout(c);
    const int32_t ngeom = (int32_t)c->part_root.size();
    const int64_t ndw = (int64_t)ninst * nd * 2, nrw = (int64_t)ninst * nt;
    char* sp = (char*)c->scratch;
    KArgs a;
    fill_args(c, a);
    a.sft = a.rtd = a.arz = a.phirz = a.vox = a.rad = nullptr;
    a.misc = nullptr;  // no event counters: the context's tallies are not touched
    a.crawl = nullptr;  // nor the context's crawl records (the crawl replay is off)
    a.det = nd > 0 ? (unsigned long long*)det_dev : nullptr;
    a.region = (unsigned long long*)region_dev;
    a.counter = (unsigned long long*)sp;
    a.inst_mat = (const float*)(sp + kScratchCounterBytes);
    a.ninst = ninst;
    a.ngeom = geom ? ngeom : 1;
    a.lmat_n = std::max<int32_t>(kMaxTissues + 1, ninst * (nt + 1));
    a.inst_count = std::max<int64_t>(count, 1);
    a.first = first;
    a.count = (int64_t)ninst * count;
    const int grid = grid_for(c, a);
    // the scratch keeps the last call's contents: clear the counter (a fill
    // kernel, not a memset node) and the caller's outputs, then the table
    int bad = launch_zero_u64(a.counter, 8, c->stream);
    if (a.det) bad |= launch_zero_u64(a.det, ndw, c->stream);
    bad |= launch_zero_u64(a.region, nrw, c->stream);
    if (bad) return fail(c, LP_EDEVICE, "lp_run_batch_noalloc: clearing the outputs failed");
    HIPCHK(c, hipMemcpyAsync(sp + kScratchCounterBytes, tab, (size_t)ninst * (nt + 1) * 4 * sizeof(float),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (a.count > 0) {
        char err[256] = {0};
        if (launch_transport(a, grid, c->stream, err, sizeof err)) return fail(c, LP_EDEVICE, "%s", err);
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.kernel_ms = ms;
    c->stats.photons = a.count;
    c->stats.grid_blocks = grid;
    return LP_OK;
}

int lp_autotune(lp_context* c, int64_t sample, int32_t passes, lp_tune_result* out)
{
    if (!c) return LP_EINVAL;
    if (sample < 1 || passes < 1 || passes > 8) return fail(c, LP_EINVAL, "lp_autotune: sample >= 1, 1..8 passes");
    int rc = prepare(c);
    if (rc) return rc;
    if ((rc = dev_alloc(c, c->d_adj_misc, kMiscN * sizeof(int64_t)))) return rc;
    // The trial launches carry the context's tallies -- their atomics are a
    // large part of the cost -- into stand-in buffers of the same sizes, so
    // the context's own tallies are untouched (counters into a scratch row).
    for (int k = 0; k < LP_NTALLY; ++k) {
        const bool on = tally_on(c, k) || (k == LP_TALLY_PHIRZ && tally_on(c, LP_TALLY_ARZ));  // phirz rides with arz
        if (k != LP_TALLY_MISC && on && c->L.n[k] > 0)
            if ((rc = dev_alloc(c, c->d_tune[k], (size_t)c->L.n[k] * sizeof(int64_t)))) return rc;
    }
    auto launch_n = [&](int64_t n, double* ms) -> int {
        KArgs a;
        fill_args(c, a);
        auto stand_in = [&](unsigned long long* p, int k) {
            return p ? (unsigned long long*)c->d_tune[k].p : nullptr;
        };
        a.sft = stand_in(a.sft, LP_TALLY_SFT);
        a.rtd = stand_in(a.rtd, LP_TALLY_RTD);
        a.phirz = a.arz ? (unsigned long long*)c->d_tune[LP_TALLY_PHIRZ].p : nullptr;
        a.arz = stand_in(a.arz, LP_TALLY_ARZ);
        a.region = stand_in(a.region, LP_TALLY_REGION);
        a.vox = stand_in(a.vox, LP_TALLY_VOX);
        a.rad = stand_in(a.rad, LP_TALLY_RAD);
        a.det = stand_in(a.det, LP_TALLY_DET);
        a.misc = (unsigned long long*)c->d_adj_misc.p;
        a.first = 0;
        a.count = n;
        int r = launch(c, a);
        *ms = c->stats.kernel_ms;
        return r;
    };
    // A setting's score is the marginal time of `sample` photons -- a launch
    // of 2 sample photons less one of sample photons -- so that the launch
    // tail, which a short launch shortens by a smaller grid and a long one
    // does not, cancels: the score is what a long launch pays per photon.
    auto trial = [&](double* ms) -> int {
        double t1 = 0, t2 = 0;
        int r = launch_n(sample, &t1);
        if (!r) r = launch_n(2 * sample, &t2);
        *ms = t2 - t1;
        return r;
    };
    struct Knob {
        const char* name;
        std::vector<double> values;
        double* cur;
    };
    double w = c->waves, fi = c->free_iters, rm = c->refill_min, lm = c->list_min,
           sm = c->search_min;
    const double w0 = w, fi0 = fi, rm0 = rm, lm0 = lm, sm0 = sm;
    std::vector<Knob> knobs = {{"waves", {1, 4, 5, 6}, &w},
                               {"free_iters", {16, 32, 64}, &fi},
                               {"refill_min", {8, 16, 24, 32}, &rm},
                               {"list_min", {0.4, 0.5, 0.6, 0.7, 0.75}, &lm},
                               {"search_min", {0.5, 0.75, 1.0}, &sm}};
    auto apply = [&]() {
        c->waves = (int32_t)w;
        c->free_iters = (int32_t)fi;
        c->refill_min = (int32_t)rm;
        c->list_min = (float)lm;
        c->search_min = (float)sm;
    };
    double ms0 = 0, best = 0, tmp = 0;
    if ((rc = trial(&tmp))) return rc;  // warm-up (first launch of the variant)
    if ((rc = trial(&ms0))) return rc;
    best = ms0;
    int trials = 2;
    for (int pass = 0; pass < passes; ++pass) {
        bool changed = false;
        for (Knob& k : knobs) {
            const double keep = *k.cur;
            double kbest = keep;
            for (double v : k.values) {
                if (v == keep) continue;
                *k.cur = v;
                apply();
                double ms = 0;
                if ((rc = trial(&ms))) {
                    *k.cur = keep;
                    apply();
                    return rc;
                }
                ++trials;
                if (ms < best * 0.99) {  // 1 % better than the best so far
                    best = ms;
                    kbest = v;
                }
            }
            *k.cur = kbest;
            apply();
            changed |= kbest != keep;
        }
        if (!changed) break;
    }
    // A choice made on single timings must beat the starting options again
    // when both are timed anew, alternately, twice each; else they stay.
    const double w1 = w, fi1 = fi, rm1 = rm, lm1 = lm, sm1 = sm;
    if (best < ms0) {
        double t0 = 0, t1 = 0;
        for (int r = 0; r < 2; ++r) {
            for (int side = 0; side < 2; ++side) {
                if (side == 0) {
                    w = w0; fi = fi0; rm = rm0; lm = lm0; sm = sm0;
                } else {
                    w = w1; fi = fi1; rm = rm1; lm = lm1; sm = sm1;
                }
                apply();
                double ms = 0;
                if ((rc = trial(&ms))) return rc;
                ++trials;
                (side == 0 ? t0 : t1) += 0.5 * ms;
            }
        }
        ms0 = t0;
        best = t1;
        if (!(t1 < 0.99 * t0)) {
            w = w0; fi = fi0; rm = rm0; lm = lm0; sm = sm0;
            best = t0;
        } else {
            w = w1; fi = fi1; rm = rm1; lm = lm1; sm = sm1;
        }
        apply();
    }
    if (out) {
        out->sample_photons = sample;
        out->trials = trials;
        out->ms_before = ms0;
        out->ms_after = best;
        out->waves = c->waves;
        KArgs a;
        fill_args(c, a);
        out->blocks_per_cu = effective_blocks_per_cu(c, a);  // in effect (the automatic value when the option is 0)
        out->free_iters = c->free_iters;
        out->refill_min = c->refill_min;
        out->list_min = c->list_min;
        out->search_min = c->search_min;
    }
    return LP_OK;
}

int lp_get_tally(lp_context* c, int32_t which, int64_t* out, int64_t n)
{
    if (!c) return LP_EINVAL;
    if (which < 0 || which >= LP_NTALLY || (!out && n > 0)) return fail(c, LP_EINVAL, "lp_get_tally: bad tally");
    compute_layout(c);
    if (c->tallies_dirty || !c->d_tally[which].p) {
        std::memset(out, 0, (size_t)n * sizeof(int64_t));
        return LP_OK;
    }
    if (n > c->L.n[which]) return fail(c, LP_EINVAL, "lp_get_tally: %lld > %lld elements", (long long)n,
                                       (long long)c->L.n[which]);
    HIPCHK(c, hipMemcpyAsync(out, c->d_tally[which].p, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LP_OK;
}

int lp_get_exit_records(lp_context* c, float* out, int64_t n)
{
    if (!c || (!out && n > 0)) return LP_EINVAL;
    if (n > c->exit_count) return fail(c, LP_ESTATE, "lp_get_exit_records: only %lld records", (long long)c->exit_count);
    HIPCHK(c, hipMemcpyAsync(out, c->d_exit.p, (size_t)n * 8 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LP_OK;
}

int lp_trace(lp_context* c, int64_t first, int32_t count, int32_t max_rec, float* out, int32_t* nrec)
{
    if (!c) return LP_EINVAL;
    if (first < 0 || count <= 0 || max_rec <= 0 || !out || !nrec) return fail(c, LP_EINVAL, "lp_trace: bad arguments");
    int rc = prepare(c);
    if (rc) return rc;
    DevBuf tr, tn;
    size_t tbytes = (size_t)count * max_rec * 12 * sizeof(float);
    if ((rc = dev_alloc(c, tr, tbytes)) || (rc = dev_alloc(c, tn, (size_t)count * sizeof(int32_t)))) {
        dev_free(tr);
        dev_free(tn);
        return rc;
    }
    KArgs a;
    fill_args(c, a);
    a.first = first;
    a.count = count;
    a.trace = (float*)tr.p;
    a.trace_n = (int32_t*)tn.p;
    a.trace_max = max_rec;
    if (c->record_exit) {
        if ((rc = dev_alloc(c, c->d_exit, (size_t)count * 8 * sizeof(float)))) {
            dev_free(tr);
            dev_free(tn);
            return rc;
        }
        a.exit_rec = (float*)c->d_exit.p;
        c->exit_count = count;
    }
    rc = launch(c, a);
    if (!rc) {
        hipError_t e1 = hipMemcpy(out, tr.p, tbytes, hipMemcpyDeviceToHost);
        hipError_t e2 = hipMemcpy(nrec, tn.p, (size_t)count * sizeof(int32_t), hipMemcpyDeviceToHost);
        if (e1 != hipSuccess || e2 != hipSuccess) rc = fail(c, LP_EDEVICE, "lp_trace: copy failed");
    }
    dev_free(tr);
    dev_free(tn);
    return rc;
}

int lp_set_jacobian_grid(lp_context* c, const float origin[3], const float size[3], const int32_t n[3])
{
    if (!c) return LP_EINVAL;
    if (!origin || !size || !n) return fail(c, LP_EINVAL, "lp_set_jacobian_grid: null argument");
    int64_t nv = 1;
    for (int k = 0; k < 3; ++k) {
        if (n[k] < 0 || !(size[k] > 0)) return fail(c, LP_EINVAL, "lp_set_jacobian_grid: bad grid");
        nv *= n[k];
    }
    if (nv > ((int64_t)1 << 28)) return fail(c, LP_EINVAL, "lp_set_jacobian_grid: at most 2^28 voxels");
    for (int k = 0; k < 3; ++k) {
        c->jac_origin[k] = origin[k];
        c->jac_size[k] = size[k];
        c->jac_n[k] = n[k];
    }
    return LP_OK;
}

int lp_adjoint(lp_context* c, int64_t first, int64_t count, int64_t* signals, int64_t* grad, int64_t* jac,
               int64_t* npairs)
{
    return lp_adjoint_voxels(c, first, count, signals, grad, jac, nullptr, npairs);
}

int lp_adjoint_voxels(lp_context* c, int64_t first, int64_t count, int64_t* signals, int64_t* grad, int64_t* jac,
                      int64_t* jac_mus, int64_t* npairs)
{
    if (!c) return LP_EINVAL;
    if (first < 0 || count < 0) return fail(c, LP_EINVAL, "lp_adjoint: negative range");
    int rc = prepare(c);
    if (rc) return rc;
    const int32_t nd = (int32_t)c->det_rlo.size(), nt = c->ntissue;
    if (nd < 1) return fail(c, LP_ESTATE, "lp_adjoint: no detectors (lp_set_detectors)");
    if (nt > kAdjMaxTissues) return fail(c, LP_EINVAL, "lp_adjoint: at most %d tissues", kAdjMaxTissues);
    // The estimator differentiates the sampling density mu_t exp(-mu_t s) of
    // a step length s = -ln(xi) / mu_t: with PixelPerUnit factor != 1 a
    // fresh step is -ln(xi) factor / mu_t while a step carried over a
    // boundary is not scaled, and that density no longer holds.
    if (c->factor != 1.0f)
        return fail(c, LP_EINVAL, "lp_adjoint: needs PixelPerUnit factor 1 (the path-length estimator assumes it)");
    const int64_t nvox = (int64_t)c->jac_n[0] * c->jac_n[1] * c->jac_n[2];
    // capacity: photons x the most detectors any exit radius can lie in
    int32_t overlap = 0;
    for (int32_t i = 0; i < nd; ++i) {
        int32_t o = 0;
        for (int32_t j = 0; j < nd; ++j)
            if (c->det_rlo[j] <= c->det_rlo[i] && c->det_rlo[i] < c->det_rhi[j]) ++o;
        overlap = std::max(overlap, o);
    }
    const int64_t cap = std::max<int64_t>(1, count * std::max(overlap, 1));
    c->pair_list.clear();
    c->pair_nt = 0;

    // 1. forward, recording the detected photons
    if ((rc = dev_alloc(c, c->d_det_list, (size_t)cap * sizeof(DetRec)))) return rc;
    if ((rc = dev_alloc(c, c->d_det_count, 64))) return rc;
    HIPCHK(c, hipMemsetAsync(c->d_det_count.p, 0, 8, c->stream));
    KArgs a;
    fill_args(c, a);
    a.first = first;
    a.count = count;
    a.det_list = (DetRec*)c->d_det_list.p;
    a.det_count = (unsigned long long*)c->d_det_count.p;
    a.det_cap = cap;
    c->exit_count = 0;
    if (count > 0 && (rc = launch(c, a))) return rc;
    const double fwd_ms = count > 0 ? c->stats.kernel_ms : 0.0;
    // every copy below is ordered on the context's (non-blocking) stream and
    // waited for, also when no launch ran to synchronise it
    unsigned long long n = 0;
    HIPCHK(c, hipMemcpyAsync(&n, c->d_det_count.p, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((int64_t)n > cap) return fail(c, LP_EDEVICE, "lp_adjoint: detection list overflow");
    std::vector<DetRec> list(n);
    if (n) {
        HIPCHK(c, hipMemcpyAsync(list.data(), c->d_det_list.p, n * sizeof(DetRec), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (signals) {
        std::fill(signals, signals + nd, 0);
        for (const DetRec& r : list) signals[r.det] += (int64_t)(r.w * 4294967296.0f);
    }

    // 2. replay the detected photons, accumulating the gradients
    const size_t gbytes = (size_t)nd * nt * 2 * sizeof(int64_t);
    const size_t jbytes = ((jac || jac_mus) && nvox > 0) ? (size_t)nd * nvox * sizeof(int64_t) : 0;
    const bool want_mus = jac_mus && jbytes;
    if ((rc = dev_alloc(c, c->d_grad, gbytes))) return rc;
    HIPCHK(c, hipMemsetAsync(c->d_grad.p, 0, gbytes, c->stream));
    if (jbytes) {
        if ((rc = dev_alloc(c, c->d_jac, jbytes))) return rc;
        HIPCHK(c, hipMemsetAsync(c->d_jac.p, 0, jbytes, c->stream));
    }
    if (want_mus) {
        if ((rc = dev_alloc(c, c->d_jac_coll, jbytes))) return rc;
        HIPCHK(c, hipMemsetAsync(c->d_jac_coll.p, 0, jbytes, c->stream));
    }
    if ((rc = dev_alloc(c, c->d_adj_misc, kMiscN * sizeof(int64_t)))) return rc;
    HIPCHK(c, hipMemsetAsync(c->d_adj_misc.p, 0, kMiscN * sizeof(int64_t), c->stream));
    KArgs b;
    fill_args(c, b);
    b.sft = b.rtd = b.arz = b.phirz = b.region = b.vox = b.rad = b.det = nullptr;
    b.misc = (unsigned long long*)c->d_adj_misc.p;
    b.exit_rec = nullptr;
    b.first = 0;
    b.count = (int64_t)n;
    b.adj_list = (const DetRec*)c->d_det_list.p;
    b.grad = (unsigned long long*)c->d_grad.p;
    b.jac = jbytes ? (unsigned long long*)c->d_jac.p : nullptr;
    b.jac_coll = want_mus ? (unsigned long long*)c->d_jac_coll.p : nullptr;
    for (int k = 0; k < 3; ++k) {
        b.jac_origin[k] = c->jac_origin[k];
        b.jac_inv[k] = 1.0f / c->jac_size[k];
        b.jac_n[k] = c->jac_n[k];
    }
    if (c->adj_paths && n > 0) {
        const size_t pbytes = (size_t)n * nt * sizeof(float);
        if ((rc = dev_alloc(c, c->d_pair_len, pbytes)) || (rc = dev_alloc(c, c->d_pair_int, pbytes))) return rc;
        b.pair_len = (float*)c->d_pair_len.p;
        b.pair_int = (uint32_t*)c->d_pair_int.p;
    }
    double rep_ms = 0;
    if (n > 0) {
        if ((rc = launch(c, b))) return rc;
        rep_ms = c->stats.kernel_ms;
    }
    int64_t misc[kMiscN];
    HIPCHK(c, hipMemcpyAsync(misc, c->d_adj_misc.p, sizeof misc, hipMemcpyDeviceToHost, c->stream));
    if (grad) HIPCHK(c, hipMemcpyAsync(grad, c->d_grad.p, gbytes, hipMemcpyDeviceToHost, c->stream));
    if (jbytes && jac) HIPCHK(c, hipMemcpyAsync(jac, c->d_jac.p, jbytes, hipMemcpyDeviceToHost, c->stream));
    if (want_mus) {  // d/dmus of a voxel = the mua term (-L_v W) + its interactions' W / mus
        HIPCHK(c, hipMemcpyAsync(jac_mus, c->d_jac_coll.p, jbytes, hipMemcpyDeviceToHost, c->stream));
        if (!jac) {
            std::vector<int64_t> ja((size_t)nd * nvox);
            HIPCHK(c, hipMemcpyAsync(ja.data(), c->d_jac.p, jbytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (size_t i = 0; i < ja.size(); ++i) jac_mus[i] += ja[i];
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (want_mus && jac)
        for (int64_t i = 0; i < (int64_t)nd * nvox; ++i) jac_mus[i] += jac[i];
    if (npairs) *npairs = (int64_t)n;
    if (c->adj_paths) {
        c->pair_list = std::move(list);
        c->pair_nt = nt;
    }
    c->stats.kernel_ms = fwd_ms;
    c->stats.photons = count;
    c->stats.replay_ms = rep_ms;
    c->stats.replayed = (int64_t)n;
    c->stats.replay_mismatch = misc[kMiscReplayMismatch];
    return LP_OK;
}

int lp_set_adjoint_paths(lp_context* c, int32_t enable)
{
    if (!c) return LP_EINVAL;
    c->adj_paths = enable != 0;
    if (!c->adj_paths) {
        dev_free(c->d_pair_len);
        dev_free(c->d_pair_int);
        c->pair_list.clear();
        c->pair_list.shrink_to_fit();
        c->pair_nt = 0;
    }
    return LP_OK;
}

int lp_get_adjoint_paths(lp_context* c, int64_t npairs, int64_t* photon, int32_t* det, float* w, float* length,
                         uint32_t* interactions)
{
    if (!c) return LP_EINVAL;
    if (!c->adj_paths) return fail(c, LP_ESTATE, "lp_get_adjoint_paths: path records are off (lp_set_adjoint_paths)");
    const int64_t n = (int64_t)c->pair_list.size();
    if (npairs != n)
        return fail(c, LP_EINVAL, "lp_get_adjoint_paths: the last lp_adjoint replayed %lld pairs, not %lld",
                    (long long)n, (long long)npairs);
    if (n == 0) return LP_OK;
    if (!photon || !det || !w || !length || !interactions) return fail(c, LP_EINVAL, "lp_get_adjoint_paths: null");
    const int32_t nt = c->pair_nt;
    std::vector<float> L((size_t)n * nt);
    std::vector<uint32_t> N((size_t)n * nt);
    HIPCHK(c, hipMemcpyAsync(L.data(), c->d_pair_len.p, L.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(N.data(), c->d_pair_int.p, N.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the detection list holds the pairs in the order the forward launch
    // found them, which varies from run to run: hand them out by (photon, detector)
    std::vector<int64_t> order((size_t)n);
    for (int64_t i = 0; i < n; ++i) order[i] = i;
    const std::vector<DetRec>& r = c->pair_list;
    std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
        return r[x].photon != r[y].photon ? r[x].photon < r[y].photon : r[x].det < r[y].det;
    });
    for (int64_t j = 0; j < n; ++j) {
        const int64_t i = order[j];
        photon[j] = r[i].photon;
        det[j] = r[i].det;
        w[j] = r[i].w;
        std::memcpy(length + j * nt, L.data() + i * nt, (size_t)nt * sizeof(float));
        std::memcpy(interactions + j * nt, N.data() + i * nt, (size_t)nt * sizeof(uint32_t));
    }
    return LP_OK;
}

int lp_get_stats(lp_context* c, lp_stats* out)
{
    if (!c || !out) return LP_EINVAL;
    int64_t bytes = 0;
    DevBuf* bufs[] = {&c->d_tris, &c->d_nodes, &c->d_nrm, &c->d_med, &c->d_safety, &c->d_vstart,
                      &c->d_vocc, &c->d_counter, &c->d_exit, &c->d_near_cells, &c->d_near_tris, &c->d_det_list,
                      &c->d_det_count, &c->d_grad, &c->d_jac, &c->d_jac_coll, &c->d_adj_misc, &c->d_inst_mat,
                      &c->d_bdet, &c->d_bregion, &c->d_pair_len, &c->d_pair_int, &c->d_crawl};
    for (DevBuf* b : bufs) bytes += (int64_t)b->bytes;
    for (auto& b : c->d_tally) bytes += (int64_t)b.bytes;
    c->stats.device_bytes = bytes;
    c->stats.bvh_nodes = (int64_t)c->bvh.nodes.size();
    c->stats.bvh_depth = c->bvh.depth;
    *out = c->stats;
    return LP_OK;
}

int lp_debug_math(lp_context* c, int32_t op, const float* xy, float* out, int64_t n)
{
    if (!c || !xy || !out || n < 0) return LP_EINVAL;
    if (n == 0) return LP_OK;
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf din, dout;
    int rc;
    if ((rc = dev_alloc(c, din, (size_t)n * 2 * sizeof(float))) || (rc = dev_alloc(c, dout, (size_t)n * sizeof(float)))) {
        dev_free(din);
        dev_free(dout);
        return rc;
    }
    rc = LP_OK;
    if (hipMemcpy(din.p, xy, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        launch_debug_math(op, (const float*)din.p, (float*)dout.p, n, nullptr) != 0 ||
        hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out, dout.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(c, LP_EDEVICE, "lp_debug_math failed");
    dev_free(din);
    dev_free(dout);
    return rc;
}

int lp_debug_copy(lp_context* c, int64_t bytes, double* ms)
{
    if (!c || bytes < 16 || bytes % 16) return c ? fail(c, LP_EINVAL, "lp_debug_copy: bytes must be a multiple of 16")
                                                 : LP_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf s, d;
    int rc;
    if ((rc = dev_alloc(c, s, (size_t)bytes)) || (rc = dev_alloc(c, d, (size_t)bytes))) {
        dev_free(s);
        dev_free(d);
        return rc;
    }
    rc = LP_OK;
    float t = 0;
    if (hipMemsetAsync(s.p, 1, (size_t)bytes, c->stream) != hipSuccess ||
        hipEventRecord(c->ev0, c->stream) != hipSuccess ||
        launch_debug_copy(s.p, d.p, bytes, c->stream) != 0 || hipEventRecord(c->ev1, c->stream) != hipSuccess ||
        hipEventSynchronize(c->ev1) != hipSuccess || hipEventElapsedTime(&t, c->ev0, c->ev1) != hipSuccess)
        rc = fail(c, LP_EDEVICE, "lp_debug_copy failed");
    if (ms) *ms = t;
    dev_free(s);
    dev_free(d);
    return rc;
}

}  // extern "C"
