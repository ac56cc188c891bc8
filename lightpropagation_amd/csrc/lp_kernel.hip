// lp_kernel.hip -- persistent-thread photon transport on triangle meshes
// for gfx950 (MI355X).
//
// One lane carries one photon packet through the loop body of the reference
// kernel (src/Simulation/MCLightPropagation.cl:1375-1575): nearest face
// ahead, step length, boundary branch (Fresnel reflect/refract) or
// interaction branch (absorb, Henyey-Greenstein spin), Russian roulette.
// The grid is persistent: each wavefront refills lanes whose photon died
// from a global photon counter with one atomic per refill (wave-level
// compaction), so lanes stay busy until the photon pool is drained.
//
// Differences of design from the reference (results are the same as the
// oracle's restatement of the reference algorithm, bit for bit):
//  * the nearest-face query runs after the step length is drawn and is
//    bounded by it (it draws no random numbers, so the stream is unchanged);
//    a step shorter than the conservative distance to the mesh skips it;
//  * faces are found through a threaded BVH whose top levels sit in LDS;
//  * per-photon counter-based (Philox) or per-photon ran3 streams replace
//    the work-group-shared, lock-ordered ran3 state (.cl:1245-1259);
//  * all tallies are 64-bit fixed-point integer atomics.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lp_devmath.hpp"
#include "lp_internal.hpp"

namespace lp {


struct Ran3Lds {
    // ran3 state of each lane, lane-minor so neighbouring lanes use
    // neighbouring banks: s[i * kBlock + tid]
    int32_t s[56 * kBlock];
};

// ---------------------------------------------------------------------------
// random numbers
// ---------------------------------------------------------------------------
struct PhiloxState {
    uint32_t id0, id1, block;                  // photon index and the next counter block
    uint32_t b0, b1, b2, b3, b4, b5, b6;       // unused words, b0 first
    int left;                                  // how many of b0..b6 are unused
    __device__ void init(uint64_t photon)
    {
        id0 = (uint32_t)photon;
        id1 = (uint32_t)(photon >> 32);
        block = 0;
        left = 0;
    }
    __device__ __forceinline__ void gen(uint32_t c[4], uint32_t k0, uint32_t k1)
    {
        c[0] = block;
        c[1] = id0;
        c[2] = id1;
        c[3] = 0u;
        philox10(c, k0, k1);
        block++;
    }
    // Leaves at least 4 unused words: one block is appended when fewer
    // remain.  Called once per wavefront iteration, so the wavefront runs one
    // Philox block per iteration instead of one at every draw site at which
    // some lane has run dry.  Registers only (value selects, no indexing).
    __device__ __forceinline__ void top_up(uint32_t k0, uint32_t k1)
    {
        if (left < 4) {
            uint32_t c[4];
            gen(c, k0, k1);
            const int l = left;
            b0 = l == 0 ? c[0] : b0;
            b1 = l == 0 ? c[1] : l == 1 ? c[0] : b1;
            b2 = l == 0 ? c[2] : l == 1 ? c[1] : l == 2 ? c[0] : b2;
            b3 = l == 0 ? c[3] : l == 1 ? c[2] : l == 2 ? c[1] : c[0];
            b4 = l == 1 ? c[3] : l == 2 ? c[2] : l == 3 ? c[1] : b4;
            b5 = l == 2 ? c[3] : l == 3 ? c[2] : b5;
            b6 = l == 3 ? c[3] : b6;
            left = l + 4;
        }
    }
    __device__ __forceinline__ float next(uint32_t k0, uint32_t k1)
    {
        if (left == 0) {  // more draws in one iteration than top_up provides (rare)
            uint32_t c[4];
            gen(c, k0, k1);
            b0 = c[0];
            b1 = c[1];
            b2 = c[2];
            b3 = c[3];
            left = 4;
        }
        // take b0 and shift: no runtime-indexed array (it would live in scratch)
        uint32_t u = b0;
        b0 = b1;
        b1 = b2;
        b2 = b3;
        b3 = b4;
        b4 = b5;
        b5 = b6;
        left--;
        return (float)(u >> 8) * 5.96046448e-08f;
    }
    // the next two words in stream order, with one shift of the buffer
    __device__ __forceinline__ void next2(uint32_t k0, uint32_t k1, float& x0, float& x1)
    {
        if (left < 2) {  // more draws in one iteration than top_up provides (rare)
            x0 = next(k0, k1);
            x1 = next(k0, k1);
            return;
        }
        const uint32_t u0 = b0, u1 = b1;
        b0 = b2;
        b1 = b3;
        b2 = b4;
        b3 = b5;
        b4 = b6;
        left -= 2;
        x0 = (float)(u0 >> 8) * 5.96046448e-08f;
        x1 = (float)(u1 >> 8) * 5.96046448e-08f;
    }
    // the value next() returns next, not consumed (after top_up: left >= 4)
    __device__ __forceinline__ float peek() const { return (float)(b0 >> 8) * 5.96046448e-08f; }
};

// PSNR::InitRand3 (src/PSNR/PSNR.cpp:48-80) in int32 (all values < 2^31),
// its first draw, then the float conversion the host does (GetRan3Array).
// The state is kept as 32-bit words; float words are bit casts.
__device__ void ran3_init_lds(int32_t* m, int tid, uint64_t seed, uint64_t photon, int& inext, int& inextp)
{
    const int32_t MBIG = 1000000000, MSEED = 161803398;
    int32_t sd = (int32_t)((seed + photon) % (uint64_t)MSEED);
    int32_t mj = (MSEED - sd) % MBIG;
    m[0 * kBlock + tid] = 0;
    m[55 * kBlock + tid] = mj;
    int32_t mk = 1;
    for (int i = 1; i <= 54; i++) {
        int ii = (21 * i) % 55;
        m[ii * kBlock + tid] = mk;
        mk = mj - mk;
        if (mk < 0) mk += MBIG;
        mj = m[ii * kBlock + tid];
    }
    for (int k = 1; k <= 4; k++)
        for (int i = 1; i <= 55; i++) {
            int32_t v = m[i * kBlock + tid] - m[(1 + (i + 30) % 55) * kBlock + tid];
            if (v < 0) v += MBIG;
            m[i * kBlock + tid] = v;
        }
    inext = 1;
    inextp = 32;
    int32_t d = m[1 * kBlock + tid] - m[32 * kBlock + tid];
    if (d < 0) d += MBIG;
    m[1 * kBlock + tid] = d;
    for (int i = 0; i < 56; ++i) m[i * kBlock + tid] = __float_as_int((float)m[i * kBlock + tid]);
}

// DoRand3 (.cl:1204-1243), one stream per lane.
__device__ __forceinline__ float ran3f_next(int32_t* m, int tid, int& inext, int& inextp)
{
    int i = inext + 1, j = inextp + 1;
    if (i == 56) i = 1;
    if (j == 56) j = 1;
    float r = __int_as_float(m[i * kBlock + tid]) - __int_as_float(m[j * kBlock + tid]);
    if (r < 0.0f) r = r + 1000000000.0f;
    m[i * kBlock + tid] = __float_as_int(r);
    inext = i;
    inextp = j;
    return r / 1000000000.0f;
}

template <int RNG, bool COUNT>
struct Rng {
    PhiloxState ph;
    uint32_t k0, k1;  // Philox key = the seed (the same for every lane)
    int inext, inextp;
    int32_t* s;
    int tid;
    int64_t ndraw;  // draws consumed (kept for the step trace only)
    __device__ void init(uint64_t seed, uint64_t photon)
    {
        if (COUNT) ndraw = 0;
        if (RNG == 0)
            ph.init(photon);
        else
            ran3_init_lds(s, tid, seed, photon, inext, inextp);
    }
    __device__ __forceinline__ void top_up()
    {
        if (RNG == 0) ph.top_up(k0, k1);
    }
    __device__ __forceinline__ float next()
    {
        if (COUNT) ndraw++;
        if (RNG == 0) return ph.next(k0, k1);
        return ran3f_next(s, tid, inext, inextp);
    }
    // the value next() returns next, not consumed (Philox: after top_up)
    __device__ __forceinline__ float peek() const
    {
        if (RNG == 0) return ph.peek();
        int i = inext + 1, j = inextp + 1;
        if (i == 56) i = 1;
        if (j == 56) j = 1;
        float r = __int_as_float(s[i * kBlock + tid]) - __int_as_float(s[j * kBlock + tid]);
        if (r < 0.0f) r = r + 1000000000.0f;
        return r / 1000000000.0f;
    }
    __device__ __forceinline__ void next2(float& x0, float& x1)
    {
        if (COUNT) ndraw += 2;
        if (RNG == 0) {
            ph.next2(k0, k1, x0, x1);
        } else {
            x0 = ran3f_next(s, tid, inext, inextp);
            x1 = ran3f_next(s, tid, inext, inextp);
        }
    }
};

// ---------------------------------------------------------------------------
// nearest face ahead
// ---------------------------------------------------------------------------
struct Hit {
    float t;
    int32_t id;
    int32_t med;
};

// Nearest face along p + t v with 0 < t < 10000 and t <= tmax, lowest
// original index among equal t, ref excluded: the argmin FindTriangleInPath
// (.cl:1037-1093) computes, restricted to t <= tmax.
//
// The running best is one 64-bit key (t bits, face id): t > 0, so the bits
// of t order as t does, and the ordered tie rule of the reference loop --
// accept t < best, or t == best with a lower id, the initial (tmax, none)
// counting as id +inf -- is the single comparison key < best.  Written as
// three separate conditional updates (t, id, med), hipcc (ROCm 7.2, gfx950)
// kept the previous med when a face was taken through the t == best branch
// (a hit at exactly t = tmax), so the key form is also what keeps med right.
__device__ __forceinline__ void find_face(const KArgs& a, const float4* lnodes, int32_t root, f3 p, f3 v, int32_t ref,
                                          float tmax, Hit& h, uint32_t& visits, uint32_t& tests)
{
    h.t = tmax;
    h.id = -1;
    h.med = 0;
    if (a.nnodes == 0) return;
    // A non-finite ray meets no face (every exact test compares NaN); return
    // at once rather than let NaN slab tests open every box.
    if (!(__builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z) &&
          __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z)))
        return;
    uint64_t best = ((uint64_t)__float_as_uint(tmax) << 32) | 0xffffffffu;
    int32_t med = 0;
    const float ix = (v.x != 0.0f) ? 1.0f / v.x : 3.0e38f;
    const float iy = (v.y != 0.0f) ? 1.0f / v.y : 3.0e38f;
    const float iz = (v.z != 0.0f) ? 1.0f / v.z : 3.0e38f;
    const float4* gn = reinterpret_cast<const float4*>(a.nodes);
    int32_t n = root;
    while (n >= 0) {
        float4 lo, hi;
        if (n < a.nlds) {
            lo = lnodes[2 * n];
            hi = lnodes[2 * n + 1];
        } else {
            lo = gn[2 * n];
            hi = gn[2 * n + 1];
        }
        visits++;
        const float bt = __uint_as_float((uint32_t)(best >> 32));
        float tx0 = (lo.x - p.x) * ix, tx1 = (hi.x - p.x) * ix;
        float ty0 = (lo.y - p.y) * iy, ty1 = (hi.y - p.y) * iy;
        float tz0 = (lo.z - p.z) * iz, tz1 = (hi.z - p.z) * iz;
        float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
        float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), bt));
        int32_t hit = __float_as_int(lo.w), miss = __float_as_int(hi.w);
        if (tn <= tf) {
            if (hit >= 0) {
                n = hit;
                continue;
            }
            int32_t first = hit & 0xffffff, cnt = (hit >> 24) & 0x7f;
            const float4* tr = reinterpret_cast<const float4*>(a.tris) + 3 * (int64_t)first;
            for (int32_t j = 0; j < cnt; ++j) {
                float4 A = tr[3 * j], B = tr[3 * j + 1], C = tr[3 * j + 2];
                int32_t id = __float_as_int(A.w);
                tests++;
                if (id == ref) continue;
                f3 E = mk(B.x, B.y, B.z), F = mk(C.x, C.y, C.z);
                f3 G = sub(p, mk(A.x, A.y, A.z));
                f3 GxE = cross(G, E), vxF = cross(v, F);
                float w = dot(vxF, E);
                if (w == 0.0f) continue;  // t = +-inf or NaN: never accepted
                float tx = dot(GxE, F);
                bool wp = w > 0.0f;
                if ((tx > 0.0f) != wp) continue;  // t = tx/w > 0 needs equal signs
                float ty = dot(vxF, G), tz = dot(GxE, v);
                float tiny = __builtin_fabsf(w) * 1e-30f;
                if (!(ty == 0.0f || (ty > 0.0f) == wp || __builtin_fabsf(ty) < tiny)) continue;
                if (!(tz == 0.0f || (tz > 0.0f) == wp || __builtin_fabsf(tz) < tiny)) continue;
                float t = tx / w, k = ty / w, l = tz / w;
                if (t > 0.0f && k >= 0.0f && l >= 0.0f && (k + l) < 1.0f && t < 10000.0f) {
                    const uint64_t key = ((uint64_t)__float_as_uint(t) << 32) | (uint32_t)id;
                    med = key < best ? __float_as_int(B.w) : med;
                    best = key < best ? key : best;
                }
            }
        }
        n = miss;
    }
    if ((uint32_t)best != 0xffffffffu) {
        h.t = __uint_as_float((uint32_t)(best >> 32));
        h.id = (int32_t)(uint32_t)best;
        h.med = med;
    }
}

__device__ __forceinline__ float safety_at(const KArgs& a, f3 p)
{
    if (!a.safety) return 0.0f;
    float fx = (p.x - a.sg_origin[0]) * a.sg_inv[0];
    float fy = (p.y - a.sg_origin[1]) * a.sg_inv[1];
    float fz = (p.z - a.sg_origin[2]) * a.sg_inv[2];
    if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f)) return 0.0f;
    int ix = (int)fminf(fx, 1e9f), iy = (int)fminf(fy, 1e9f), iz = (int)fminf(fz, 1e9f);
    if (ix >= a.sg_n[0] || iy >= a.sg_n[1] || iz >= a.sg_n[2]) return 0.0f;
    return a.safety[((int64_t)iz * a.sg_n[1] + iy) * a.sg_n[0] + ix];
}

// Squared distance from p to the triangle (a, a + ab, a + ac): the
// closest-point construction of C. Ericson, Real-Time Collision Detection
// (2005), section 5.1.5, by Voronoi regions of the vertices and edges.
__device__ __forceinline__ float point_tri_dist2(f3 p, f3 a, f3 ab, f3 ac)
{
    f3 ap = sub(p, a);
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return dot(ap, ap);
    f3 bp = sub(ap, ab);
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return dot(bp, bp);
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        f3 q = sub(ap, mul(d1 / (d1 - d3), ab));
        return dot(q, q);
    }
    f3 cp = sub(ap, ac);
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return dot(cp, cp);
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        f3 q = sub(ap, mul(d2 / (d2 - d6), ac));
        return dot(q, q);
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        f3 q = sub(bp, mul(w, sub(ac, ab)));
        return dot(q, q);
    }
    float den = 1.0f / (va + vb + vc);
    f3 q = sub(ap, add(mul(vb * den, ab), mul(vc * den, ac)));
    return dot(q, q);
}

// Distance from p to the nearest face, searched out to `cap` (returns cap
// when nothing is nearer).  Boxes are padded, so box distances are lower
// bounds; the caller subtracts a rounding margin.
__device__ __forceinline__ float mesh_distance(const KArgs& a, const float4* lnodes, int32_t root, f3 p, float cap,
                                               uint32_t& visits)
{
    float best2 = cap * cap;
    const float4* gn = reinterpret_cast<const float4*>(a.nodes);
    int32_t n = root;
    while (n >= 0) {
        float4 lo, hi;
        if (n < a.nlds) {
            lo = lnodes[2 * n];
            hi = lnodes[2 * n + 1];
        } else {
            lo = gn[2 * n];
            hi = gn[2 * n + 1];
        }
        visits++;
        float dx = fmaxf(fmaxf(lo.x - p.x, p.x - hi.x), 0.0f);
        float dy = fmaxf(fmaxf(lo.y - p.y, p.y - hi.y), 0.0f);
        float dz = fmaxf(fmaxf(lo.z - p.z, p.z - hi.z), 0.0f);
        int32_t hit = __float_as_int(lo.w), miss = __float_as_int(hi.w);
        if ((dx * dx + dy * dy) + dz * dz < best2) {
            if (hit >= 0) {
                n = hit;
                continue;
            }
            int32_t first = hit & 0xffffff, cnt = (hit >> 24) & 0x7f;
            const float4* tr = reinterpret_cast<const float4*>(a.tris) + 3 * (int64_t)first;
            for (int32_t j = 0; j < cnt; ++j) {
                float4 A = tr[3 * j], B = tr[3 * j + 1], C = tr[3 * j + 2];
                float d2 = point_tri_dist2(p, mk(A.x, A.y, A.z), mk(B.x, B.y, B.z), mk(C.x, C.y, C.z));
                best2 = fminf(best2, d2);
            }
        }
        n = miss;
    }
    return __builtin_sqrtf(best2);
}

// Cell of the near-face grid holding p, or -1 outside it (then no face lies
// within near_r0 of p), in the frame all geometries share (callers add the
// cell offset of the photon's geometry).
// (The grid holds fewer than 2^31 cells, so 32-bit index arithmetic.)
__device__ __forceinline__ int32_t near_cell(const KArgs& a, f3 p)
{
    float fx = (p.x - a.ng_origin[0]) * a.ng_inv[0];
    float fy = (p.y - a.ng_origin[1]) * a.ng_inv[1];
    float fz = (p.z - a.ng_origin[2]) * a.ng_inv[2];
    if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f)) return -1;
    int32_t ix = (int32_t)fminf(fx, 2.0e9f), iy = (int32_t)fminf(fy, 2.0e9f), iz = (int32_t)fminf(fz, 2.0e9f);
    if (ix >= a.ng_n[0] || iy >= a.ng_n[1] || iz >= a.ng_n[2]) return -1;
    return (iz * a.ng_n[1] + iy) * a.ng_n[0] + ix;
}

// BVH root and near-grid cell offset of an instance's geometry: the int bits
// in the y and z words of its (rsp, root, coff, 0) material row in a launch
// over several geometries (MG), node 0 and cell offset 0 otherwise.
struct Geom {
    int32_t root;
    uint32_t coff;
};
template <bool MG>
__device__ __forceinline__ Geom geom_of(const KArgs& a, const float4* lmat, int32_t inst)
{
    Geom g{0, 0u};
    if (MG) {
        const float4 r = lmat[inst * (a.ntissue + 1) + a.ntissue];
        g.root = __float_as_int(r.y);
        g.coff = __float_as_uint(r.z);
    }
    return g;
}

// Per-wavefront LDS rows of the near-face list query: one slot per lane,
// holding that lane's results while the pairs of its list are merged in.
struct NearRed {
    unsigned long long key[64];  // (t bits, face id) of the nearest hit
    unsigned long long d1[64];   // (squared distance bits, list record) of the nearest listed face
    uint32_t d2[64];             // bits of the second-smallest squared distance
    int32_t med[64];             // packed tissues of the nearest hit
};

// Near-face list query, wave-cooperative, for the lanes with `part` set: the
// (photon, listed face) pairs of all those lanes are numbered by a prefix sum
// over their list lengths and tested 64 at a time, one pair per lane, so a
// pass costs one face test whatever the lengths of the individual lists.
// Each pair's outcome is merged into its photon's slot of R by LDS atomic
// minima: the nearest hit is the lexicographic minimum of (t, face id) over
// the hits with t <= tmax (the nearest face and tie rule of find_face) and,
// when want, the two smallest distances from p to the listed faces with the
// record of the smaller (a value that loses a minimum, or is displaced from
// it, is a candidate for the second smallest); unlisted = a lower bound on
// the distance to every face not in the list.
// The list is that of the cell holding cq (a point on the ray: p, or where
// a march along it has got to); rays and distances are from p.  With at_p
// (cq is p) the cell's record comes from the caller: cell pc (< 0: outside
// the grid), list start pfirst, clearance bits pbits, next list start pnext.
// Must be called by the whole wavefront.
__device__ __forceinline__ void near_batch(const KArgs& a, NearRed& R, bool part, uint32_t coff, f3 cq, f3 p, f3 v,
                                           int32_t ref, float tmax, bool want, bool at_p, int32_t pc, uint32_t pfirst,
                                           uint32_t pbits, uint32_t pnext, Hit& h, float& d1sq, float& d2sq,
                                           int32_t& i1, float& unlisted, uint32_t& pairs)
{
    const int lane = threadIdx.x & 63;
    h.t = tmax;
    h.id = -1;
    h.med = 0;
    d1sq = d2sq = __builtin_inff();
    i1 = -1;
    unlisted = a.near_r0;  // the grid reaches r0 beyond every face
    uint32_t b = 0, n = 0;
    if (part && at_p) {
        if (pc >= 0) {
            b = pfirst;
            n = pnext - pfirst;
            unlisted = fmaxf(a.near_r0, __uint_as_float(pbits));
        }
    } else if (part) {
        const int32_t c0 = near_cell(a, cq);
        if (c0 >= 0) {
            const uint32_t c = (uint32_t)c0 + coff;
            const CellRec cr = a.near_cells[c];
            b = cr.first;
            n = a.near_cells[c + 1].first - b;
            unlisted = fmaxf(a.near_r0, __uint_as_float(cr.clear_bits));
        }
    }
    // inclusive prefix sum of the list lengths over the wavefront, by DPP:
    // shifts by 1, 2, 4, 8 within each row of 16 lanes, then lane 15 of a row
    // added to the next row and lane 31 to the upper two rows (lanes a
    // shift reads from outside its row take 0)
    int incl = (int)n;
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xf, 0xf, false);  // row_shr:1
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xf, 0xf, false);  // row_shr:2
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xf, 0xf, false);  // row_shr:4
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xf, 0xf, false);  // row_shr:8
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x142, 0xa, 0xf, false);  // row_bcast:15
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x143, 0xc, 0xf, false);  // row_bcast:31
    const uint32_t excl = (uint32_t)incl - n;
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane(incl, 63);
    pairs = total;
    R.key[lane] = ~0ull;
    R.d1[lane] = ~0ull;
    R.d2[lane] = 0x7f800000u;  // +inf
    R.med[lane] = 0;
    const float4* tr = reinterpret_cast<const float4*>(a.near_tris);
    for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t g = base + lane;
        // owner of pair g: the last lane whose first pair is <= g
        int j = 0;
#pragma unroll
        for (int st = 32; st >= 1; st >>= 1) {
            uint32_t ex = __shfl(excl, j + st);
            if (ex <= g) j += st;
        }
        const f3 op = mk(__shfl(p.x, j), __shfl(p.y, j), __shfl(p.z, j));
        const f3 ov = mk(__shfl(v.x, j), __shfl(v.y, j), __shfl(v.z, j));
        const int32_t oref = __shfl(ref, j);
        const float otmax = __shfl(tmax, j);
        const int owant = __shfl((int)want, j);
        const uint32_t k = __shfl(b, j) + (g - __shfl(excl, j));
        uint64_t key = ~0ull;
        int32_t med = 0;
        if (g < total) {
            const float4 A = tr[3 * (int64_t)k], B = tr[3 * (int64_t)k + 1], C = tr[3 * (int64_t)k + 2];
            const f3 M = mk(A.x, A.y, A.z), E = mk(B.x, B.y, B.z), F = mk(C.x, C.y, C.z);
            if (owant) {
                const float dd = point_tri_dist2(op, M, E, F);
                const unsigned long long dk = ((unsigned long long)__float_as_uint(dd) << 32) | k;
                const unsigned long long old = atomicMin(&R.d1[j], dk);
                atomicMin(&R.d2[j], (uint32_t)((old > dk ? old : dk) >> 32));
            }
            const int32_t id = __float_as_int(A.w);
            const f3 G = sub(op, M);
            const f3 GxE = cross(G, E), vxF = cross(ov, F);
            const float w = dot(vxF, E);
            const float tx = dot(GxE, F);
            const bool wp = w > 0.0f;
            const float ty = dot(vxF, G), tz = dot(GxE, ov);
            const float tiny = __builtin_fabsf(w) * 1e-30f;
            bool ok = id != oref && w != 0.0f && (tx > 0.0f) == wp;
            ok = ok && (ty == 0.0f || (ty > 0.0f) == wp || __builtin_fabsf(ty) < tiny);
            ok = ok && (tz == 0.0f || (tz > 0.0f) == wp || __builtin_fabsf(tz) < tiny);
            if (ok) {
                const float t = tx / w, kk = ty / w, l = tz / w;
                if (t > 0.0f && kk >= 0.0f && l >= 0.0f && (kk + l) < 1.0f && t < 10000.0f && t <= otmax) {
                    key = ((uint64_t)__float_as_uint(t) << 32) | (uint32_t)id;
                    med = __float_as_int(B.w);
                    atomicMin(&R.key[j], (unsigned long long)key);
                }
            }
        }
        // every lane's minimum of this pass is in place (a wavefront's LDS
        // operations complete in order); the pair holding its owner's
        // current nearest hit writes that hit's tissues
        __builtin_amdgcn_wave_barrier();
        if (key != ~0ull && R.key[j] == key) R.med[j] = med;
        __builtin_amdgcn_wave_barrier();
    }
    if (part) {
        const unsigned long long best = R.key[lane];
        if (best != ~0ull) {
            h.t = __uint_as_float((uint32_t)(best >> 32));
            h.id = (int32_t)(uint32_t)best;
            h.med = R.med[lane];
        }
        const unsigned long long d1k = R.d1[lane];
        if (want && d1k != ~0ull) {
            d1sq = __uint_as_float((uint32_t)(d1k >> 32));
            i1 = (int32_t)(uint32_t)d1k;
            d2sq = __uint_as_float(R.d2[lane]);
        }
    }
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void atomic_add_u64(unsigned long long* p, unsigned long long v)
{
    if (v) atomicAdd(p, v);
}

// ---------------------------------------------------------------------------
// the transport kernel
// ---------------------------------------------------------------------------
struct Photon {
    f3 p, v;
    float w, unfin;
    // last boundary normal as (face id << 1) | near-axis override, -1 for the
    // initial (0, 0, -1); the exit angle recomputes it from the face record
    int32_t nvc;
    // free region: the ball of radius safe around anc, cut by the
    // half-space dot(pu, x) > pc (pu = 0, pc = -1: no cut); no face meets
    // the region, p lies inside it, and it is convex, so a step whose end
    // point stays inside needs no face search
    f3 anc, pu;
    float safe, pc;
    int32_t ref, curr;
    // fetal: bit 0 = traversed the fetal tissue; bits 1.. = boundary events
    // since the last interaction
    int32_t fetal, roulette;
    int32_t inst;  // instance of a batched launch (0 otherwise)
    uint32_t steps;
    int64_t k;  // index within this run
};

// Adjoint: adds -(length in voxel) * W of the segment c0 -> c1 (length s,
// emitter frame) to each voxel of the Jacobian grid it crosses --
// Amanatides & Woo's traversal in grid units, clipped to the grid, one axis
// advanced per crossing (x before y before z on ties); the same float
// operations as the oracle's restatement, so the int64 sums agree bit for bit.
__device__ __forceinline__ void jac_dda(const KArgs& a, unsigned long long* jd, f3 c0, f3 c1, float s, float W)
{
    const int32_t n0 = a.jac_n[0], n1 = a.jac_n[1], n2 = a.jac_n[2];
    const float g0x = (c0.x - a.jac_origin[0]) * a.jac_inv[0];
    const float g0y = (c0.y - a.jac_origin[1]) * a.jac_inv[1];
    const float g0z = (c0.z - a.jac_origin[2]) * a.jac_inv[2];
    const float dx = (c1.x - a.jac_origin[0]) * a.jac_inv[0] - g0x;
    const float dy = (c1.y - a.jac_origin[1]) * a.jac_inv[1] - g0y;
    const float dz = (c1.z - a.jac_origin[2]) * a.jac_inv[2] - g0z;
    float t0 = 0.0f, t1 = 1.0f;
    auto clip = [&](float g, float d, int32_t n) -> bool {
        if (d == 0.0f) return g >= 0.0f && g < (float)n;
        float ta = (0.0f - g) / d, tb = ((float)n - g) / d;
        if (ta > tb) {
            float x = ta;
            ta = tb;
            tb = x;
        }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
        return true;
    };
    if (!clip(g0x, dx, n0) || !clip(g0y, dy, n1) || !clip(g0z, dz, n2)) return;
    if (!(t0 < t1)) return;
    const float inf = __builtin_inff();
    auto setup = [&](float g, float d, int32_t n, int32_t& i, int32_t& st, float& tmax, float& tdel) {
        int32_t k = (int32_t)__builtin_floorf(g + t0 * d);
        if (k < 0) k = 0;
        if (k > n - 1) k = n - 1;
        i = k;
        if (d > 0.0f) {
            st = 1;
            tmax = ((float)(k + 1) - g) / d;
            tdel = 1.0f / d;
        } else if (d < 0.0f) {
            st = -1;
            tmax = ((float)k - g) / d;
            tdel = -1.0f / d;
        } else {
            st = 0;
            tmax = inf;
            tdel = inf;
        }
    };
    int32_t ix, iy, iz, sx, sy, sz;
    float mx, my, mz, ex, ey, ez;
    setup(g0x, dx, n0, ix, sx, mx, ex);
    setup(g0y, dy, n1, iy, sy, my, ey);
    setup(g0z, dz, n2, iz, sz, mz, ez);
    float t = t0;
    for (int32_t it = 0; it < n0 + n1 + n2 + 3; ++it) {
        const float tn = fminf(fminf(mx, my), fminf(mz, t1));
        if (tn > t) {
            const int64_t v = ((int64_t)iz * n1 + iy) * n0 + ix;
            atomic_add_u64(&jd[v], fx32(-(((tn - t) * s) * W)));
            t = tn;
        }
        if (t >= t1) break;
        if (mx <= my && mx <= mz) {
            ix += sx;
            if (ix < 0 || ix >= n0) break;
            mx += ex;
        } else if (my <= mz) {
            iy += sy;
            if (iy < 0 || iy >= n1) break;
            my += ey;
        } else {
            iz += sz;
            if (iz < 0 || iz >= n2) break;
            mz += ez;
        }
    }
}

// boundary events in a row after which a lane's searches no longer wait
// for a full batch
constexpr uint32_t kUrgentRun = 16;
// live lanes at or below which a wavefront that has no photons left to start
// runs its searches at once (its last photons are the launch's tail)
constexpr int kTailLanes = 16;

// per-wavefront event counters in LDS
enum WaveCount {
    wPhot, wInt, wBnd, wEsc, wCap, wSkip, wDq, wNan, wLst, wIter, wLRounds, wBRounds, wArzOut, wVoxOut, wMismatch,
    wMarch, wCrawl, wVis, wTst, wDvis, kWc
};

// WAVES: minimum waves per SIMD the register allocator must allow (1 =
// the compiler's choice); a tuning knob, results do not depend on it.
// ADJ: the adjoint replay -- photons come from the detection list and every
// move adds to the path-length gradients instead of the tallies.
// MG: a batched launch over several geometries -- each photon's searches
// start at its instance's BVH root and near-grid cell offset.
template <int RNG, bool TRACE, int WAVES, bool ADJ, bool MG>
__global__ __launch_bounds__(kBlock, WAVES) void transport_kernel(KArgs a_arg)
{
    // The arguments are read through a pointer that each iteration makes
    // opaque to the compiler: a field is then loaded from the argument
    // segment (scalar cache) where it is used instead of being held in a
    // scalar register for the whole loop, which would exceed the scalar
    // register file and spill to vector lanes.
    // (the argument is the whole kernel-argument segment, from offset 0)
    typedef const __attribute__((address_space(4))) KArgs CKArgs;
    CKArgs* ap = (CKArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    // a fresh copy of that pointer which the compiler cannot see through
    auto launder = [&]() {
        CKArgs* p = (CKArgs*)__builtin_amdgcn_kernarg_segment_ptr();
        asm volatile("" : "+s"(p));
        ap = p;
    };
#define a (*(const KArgs*)ap)
    extern __shared__ float4 smem[];
    __shared__ unsigned long long wcnt[kBlock / 64][kWc];
    __shared__ NearRed nred[kBlock / 64];
    float4* lnodes = smem;                         // 2 * nlds
    float4* lmat = smem + 2 * a.nlds;              // lmat_n: per instance, ntissue materials + (rsp, 0, 0, 0)
    float* lalb = reinterpret_cast<float*>(lmat + a.lmat_n);  // lmat_n: mua / (mua + mus) of each row
    int32_t* lran3 = reinterpret_cast<int32_t*>(lalb + a.lmat_n);
    // adjoint: per-lane path length and interaction count per tissue, tissue-major
    float* adjL = reinterpret_cast<float*>(reinterpret_cast<char*>(lran3) + (RNG == 1 ? sizeof(Ran3Lds) : 0));
    uint32_t* adjN = reinterpret_cast<uint32_t*>(adjL + kAdjMaxTissues * kBlock);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    {
        const float4* gn = reinterpret_cast<const float4*>(a.nodes);
        for (int i = tid; i < 2 * a.nlds; i += kBlock) lnodes[i] = gn[i];
        for (int i = tid; i < (kBlock / 64) * kWc; i += kBlock) (&wcnt[0][0])[i] = 0;
        if (a.inst_mat) {
            const float4* im = reinterpret_cast<const float4*>(a.inst_mat);
            for (int i = tid; i < a.ninst * (a.ntissue + 1); i += kBlock) lmat[i] = im[i];
        } else {
            for (int i = tid; i < a.ntissue; i += kBlock)
                lmat[i] = make_float4(a.nmuasg[4 * i], a.nmuasg[4 * i + 1], a.nmuasg[4 * i + 2], a.nmuasg[4 * i + 3]);
            if (tid == 0) lmat[a.ntissue] = make_float4(a.rsp, 0.0f, 0.0f, 0.0f);
        }
    }
    __syncthreads();
    // the absorbed fraction of each material row, the same division Absorb makes
    for (int i = tid; i < a.ninst * (a.ntissue + 1); i += kBlock) {
        const float4 m = lmat[i];
        lalb[i] = (m.y > 0.0f && m.z > 0.0f) ? m.y / (m.y + m.z) : 0.0f;
    }
    __syncthreads();
    const int32_t ntp1 = a.ntissue + 1;  // material rows per instance

    // the emitter frame, read from the argument segment where used (held in
    // scalar registers for the whole loop it would spill to vector lanes)
#define POI mk(a.poi[0], a.poi[1], a.poi[2])
#define eX mk(a.ex[0], a.ex[1], a.ex[2])
#define eY mk(a.ey[0], a.ey[1], a.ey[2])
#define eZ mk(a.ez[0], a.ez[1], a.ez[2])
#define eZr mk(a.ez_raw[0], a.ez_raw[1], a.ez_raw[2])

    Rng<RNG, TRACE> rng;
    rng.s = lran3;
    rng.tid = tid;
    rng.k0 = (uint32_t)a.seed;
    rng.k1 = (uint32_t)(a.seed >> 32);
    // each lane's two records of edge-crawl crossings: those of an earlier
    // launch (other optics, other photons) are voided before any is read
    if constexpr (!ADJ) {
        const uint32_t gid = blockIdx.x * kBlock + tid;
        if (a.crawl && (int64_t)gid < a.crawl_lanes) {
            uint32_t* u = reinterpret_cast<uint32_t*>(a.crawl + (size_t)gid * kCrawlSlots * kCrawlWords);
#pragma unroll
            for (int i = 0; i < kCrawlSlots; ++i) {
                u[i * kCrawlWords] = 0xffffffffu;
                u[i * kCrawlWords + 1] = 0xffffffffu;
            }
        }
    }
    Photon ph;
    bool alive = false;
    bool exhausted = false;  // wave-uniform
    // per-lane 32-bit counters (photons, steps, interactions, boundary
    // events, escaped, capped, skipped searches, distance queries, NaN
    // photons, BVH visits, triangle tests, distance-query visits); flushed
    // to the 64-bit totals when one grows large and at the end
    // Event counters live in LDS, one row per wavefront: the first active
    // lane adds the number of active lanes (or a lane adds its own rare
    // weight), so counting costs no registers.  Only the visit counts of
    // the BVH loops are kept per lane (flushed when one grows large and at
    // the end).
    unsigned long long* wc = wcnt[tid >> 6];
    auto wcount = [&](int slot) {
        const unsigned long long m = __ballot(1);
        if (lane == __ffsll((long long)m) - 1) atomicAdd(&wc[slot], (unsigned long long)__popcll(m));
    };
    // the frequent events are counted per round in wave-uniform registers
    // (ballot popcounts) and added to the LDS row once per round
    uint32_t r_int = 0, r_bnd = 0, r_skip = 0, r_iter = 0, r_lr = 0, r_lst = 0;
    // adds a lane's own count (search statistics of the rare BVH batches)
    auto wsum = [&](int slot, uint32_t v) {
        if (v) atomicAdd(&wc[slot], (unsigned long long)v);
    };
    unsigned long long reg_acc = 0;  // fixed-point absorption in ph.curr not yet flushed
    int32_t vox_cell = -1;           // voxel whose deposits are pending in vox_acc
    unsigned long long vox_acc = 0;

    float stepsize = 0.0f;  // drawn step length, pending until used
    float mq = 0.0f;        // a long step's search: the length of the ray shown free of faces so far
    bool ballistic = false;
    bool have_step = false;  // a step length has been drawn and not yet used
    bool search = false;     // the pending step needs a face search
    bool by_list = false;    // ... from the near-face lists (else from the BVH)
    bool crawled = false;    // this iteration's boundary event left p where it was (and was recorded)

    // weights recorded at the exit (R, fetal R, T): set by the boundary event
    // that ends the photon, read by the epilogue in the same iteration
    f3 xsft = mk(0.0f, 0.0f, 0.0f);
    int32_t adj_det = 0;  // adjoint: the detector and R weight of the replayed photon
    float adj_w = 0.0f;
    int64_t adj_k = 0;    // ... and its index in the detection list
    // adjoint: one move of length s from p0 in the current tissue, ending in
    // an interaction (collision) or at a face
    auto adj_step = [&](f3 p0, float s, bool collision, float mus) {
        adjL[ph.curr * kBlock + tid] += s;
        if (collision) adjN[ph.curr * kBlock + tid] += 1u;
        if (a.jac) {
            const int64_t nvox = (int64_t)a.jac_n[0] * a.jac_n[1] * a.jac_n[2];
            const f3 d0 = sub(p0, POI), d1 = sub(ph.p, POI);
            const f3 c1 = mk(dot(eX, d1), dot(eY, d1), dot(eZ, d1));
            jac_dda(a, a.jac + adj_det * nvox, mk(dot(eX, d0), dot(eY, d0), dot(eZ, d0)), c1, s, adj_w);
            // the interaction's mus term goes to the voxel holding it
            if (collision && a.jac_coll && mus != 0.0f) {
                const float gx = (c1.x - a.jac_origin[0]) * a.jac_inv[0];
                const float gy = (c1.y - a.jac_origin[1]) * a.jac_inv[1];
                const float gz = (c1.z - a.jac_origin[2]) * a.jac_inv[2];
                if (gx >= 0.0f && gy >= 0.0f && gz >= 0.0f && gx < (float)a.jac_n[0] && gy < (float)a.jac_n[1] &&
                    gz < (float)a.jac_n[2]) {
                    const int64_t v = ((int64_t)(int32_t)gz * a.jac_n[1] + (int32_t)gy) * a.jac_n[0] + (int32_t)gx;
                    atomic_add_u64(&a.jac_coll[adj_det * nvox + v], fx32(adj_w / mus));
                }
            }
        }
    };

    // interaction branch (.cl:1513-1537)
    auto interaction = [&](float s) {
        const float4 cm = lmat[ph.inst * ntp1 + ph.curr];
        const f3 p_old = ph.p;
        ph.p = add(ph.p, mul(s, ph.v));
        if constexpr (ADJ) adj_step(p_old, s, true, cm.z);
        ph.ref = -1;
        // Absorb (.cl:841-847); skipped only in a start tissue with mua or
        // mus 0, which a photon never re-enters and where nothing was absorbed
        float absorbed = 0.0f;
        if (cm.y > 0.0f && cm.z > 0.0f) {
            absorbed = ph.w * lalb[ph.inst * ntp1 + ph.curr];  // ph.w * (mua / (mua + mus))
            ph.w -= absorbed;
        }
        float xt, xp;
        rng.next2(xt, xp);
            {  // VScatter (.cl:597-655)
            const f3 v = ph.v;
            const float g = cm.w;
            f3 e_rot = mk(0.0f, 0.0f, 0.0f);
            if (v.z != 0.0f) {
                float z = -(v.x + v.y) / v.z;
                e_rot = mk(1.0f, 1.0f, z);
            } else if (v.y != 0.0f) {
                float y = -(v.x) / v.y;
                e_rot = mk(1.0f, y, 0.0f);
            } else if (v.x != 0.0f) {
                e_rot = mk(0.0f, 1.0f, 0.0f);
            }
            e_rot = normalize_n(e_rot);  // squared norm >= 2
            float cos_t;
            float g1 = (1 - g * g) / (1 - g + 2 * g * xt);
            if (g != 0.0f) {
                cos_t = (1 + g * g - g1 * g1) / (2 * g);
                if (cos_t < -1) cos_t = -1;
                else if (cos_t > 1) cos_t = 1;
            } else {
                cos_t = 2 * xt - 1;
            }
            float sin_t = sqrt_n(1.0f - cos_t * cos_t);
            float psi = 6.28318548f * xp;
            float cos_p = cosf_lp(psi);
            float sin_p = sqrt_n(1.0f - cos_p * cos_p);
            if (psi > 3.14159274f) sin_p *= -1;
            // Rotate (.cl:697-708) twice
            float d1 = dot(e_rot, v);
            f3 c1 = cross(e_rot, v);
            f3 c1n = cross(c1, e_rot);
            // rt and the new direction are sums of orthogonal parts of norm
            // near 1 (e_rot is normal to v)
            f3 rt = normalize_n(add(add(mul(d1, e_rot), mul(cos_t, c1n)), mul(sin_t, c1)));
            float d2 = dot(v, rt);
            f3 c2 = cross(v, rt);
            f3 c2n = cross(c2, v);
            ph.v = normalize_n(add(add(mul(d2, v), mul(cos_p, c2n)), mul(sin_p, c2)));
        }
            // absorption tallies at the new position (replaces RecordAbsorbtionRZ, .cl:254-271)
        const f3 d = sub(ph.p, POI);
        const f3 c = mk(dot(eX, d), dot(eY, d), dot(eZ, d));
        unsigned long long q = fx32_nn(absorbed);
        reg_acc += q;
        if (a.arz) {
            float rho = __builtin_sqrtf(c.x * c.x + c.y * c.y) * sign_cl(c.x) / a.det_grid[3];
            int32_t ir = cvt_i32f(rho);
            int32_t iz = cvt_i32f(c.z / a.det_grid[2]);
            if (ir >= -a.dimR && ir <= a.dimR && iz >= 0 && iz <= a.dimz) {
                int64_t cell = (int64_t)iz * a.arz_nr + (ir + a.dimR);
                atomic_add_u64(&a.arz[cell], q);
                atomic_add_u64(&a.phirz[cell], fx24(absorbed / (cm.y == 0.0f ? 1.0f : cm.y)));
            } else {
                if (q) atomicAdd(&wc[wArzOut], q);
            }
        }
        if (a.vox) {
            float fx = (c.x - a.vox_origin[0]) * a.vox_inv[0];
            float fy = (c.y - a.vox_origin[1]) * a.vox_inv[1];
            float fz = (c.z - a.vox_origin[2]) * a.vox_inv[2];
            // the truncation lies in [0, n) exactly when x does (n < 2^24 is
            // exact in float): the host's int() binning with its bounds check
            const bool ok = fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < a.vox_nf[0] && fy < a.vox_nf[1] &&
                            fz < a.vox_nf[2];
            // consecutive deposits in one voxel are summed in a register
            // (value selects, so the accumulator stays in registers); fewer
            // than 2^31 voxels, so 32-bit indices
            // (clamped first, so that the conversions are defined off the grid too)
            const int32_t ix = (int32_t)__builtin_amdgcn_fmed3f(fx, 0.0f, 16777216.0f);
            const int32_t iy = (int32_t)__builtin_amdgcn_fmed3f(fy, 0.0f, 16777216.0f);
            const int32_t iz = (int32_t)__builtin_amdgcn_fmed3f(fz, 0.0f, 16777216.0f);
            const int32_t cell = (iz * a.vox_n[1] + iy) * a.vox_n[0] + ix;
            bool same = ok && cell == vox_cell;
            if (ok && !same && vox_cell >= 0) atomic_add_u64(&a.vox[vox_cell], vox_acc);
            vox_acc = same ? vox_acc + q : (ok ? q : vox_acc);
            vox_cell = ok ? cell : vox_cell;
            if (!ok && q) atomicAdd(&wc[wVoxOut], q);
        }
    };

    // boundary branch (.cl:1462-1511) at face h.id, distance stb
    auto boundary_event = [&](const Hit& h, float s, float stb) -> int {
        const float4 cm = lmat[ph.inst * ntp1 + ph.curr];
        int kind;
        const f3 p_old = ph.p;
        ph.p = add(ph.p, mul(stb, ph.v));  // MoveStep
        if constexpr (ADJ) adj_step(p_old, stb, false, 0.0f);
        // A move below the float spacing of p leaves p unchanged (a photon on
        // the shared edge of two faces): such a crossing is recorded, input
        // state first, in the lane's ring slot for its step number.
        float* crec = nullptr;
        if constexpr (!ADJ) {
            const uint32_t gid = blockIdx.x * kBlock + tid;
            if (a.crawl && h.id >= 0 && (int64_t)gid < a.crawl_lanes &&
                __float_as_uint(ph.p.x) == __float_as_uint(p_old.x) &&
                __float_as_uint(ph.p.y) == __float_as_uint(p_old.y) &&
                __float_as_uint(ph.p.z) == __float_as_uint(p_old.z)) {
                crec = a.crawl + ((size_t)gid * kCrawlSlots + (ph.steps & (kCrawlSlots - 1))) * kCrawlWords;
                uint32_t* u = reinterpret_cast<uint32_t*>(crec);
                u[0] = (uint32_t)ph.k;
                u[1] = (uint32_t)((uint64_t)ph.k >> 32);
                u[2] = ph.steps;
                crec[3] = ph.v.x;
                crec[4] = ph.v.y;
                crec[5] = ph.v.z;
                u[6] = (uint32_t)ph.ref;
                u[7] = (uint32_t)ph.curr;
            }
        }
        ph.ref = -1;
        ph.safe = 0.0f;
        if (h.id < 0) {  // nothing ahead: the photon has left the mesh
            kind = 3;
            ph.w = 0.0f;
            wcount(wEsc);
        } else {
            float x = rng.next();
            int32_t in_id = h.med & 0xffff, out_id = (h.med >> 16) & 0xffff;
            f3 nv = normalize(mk(a.nrm[3 * h.id], a.nrm[3 * h.id + 1], a.nrm[3 * h.id + 2]));
            int32_t nx_id = (ph.curr == out_id) ? in_id : out_id;
            const float4 nm = lmat[ph.inst * ntp1 + nx_id];
            const bool axis = __builtin_fabsf(__builtin_sqrtf(ph.p.x * ph.p.x + ph.p.y * ph.p.y)) < 0.01f;
            if (axis) nv = mul(sign_cl(dot(nv, eZr)), eZ);
            ph.nvc = (h.id << 1) | (axis ? 1 : 0);
            // TransmitOrReflect (.cl:921-1018)
            float cos_in = dot(nv, ph.v);
            float abs_cos = __builtin_fabsf(cos_in);
            float n1 = cm.x, n2 = nm.x;
            float cos_crit = 0.0f;
            if (n1 > n2) {
                float q = n2 / n1;
                cos_crit = __builtin_sqrtf(1.0f - q * q);
            }
            float r;
            if (abs_cos <= cos_crit) {
                r = 1.0f;
            } else if (n1 == n2) {  // GetFresnel (.cl:857-907)
                r = 0.0f;
            } else if (abs_cos >= 1.0f) {  // normal incidence (> 1 only by rounding)
                r = (n1 - n2) / (n1 + n2);
                r *= r;
            } else if (cos_in == 0.0f) {
                r = 1.0f;
            } else {
                float sin_in = __builtin_sqrtf(clamp0(1 - cos_in * cos_in));
                float sin_out = n1 * sin_in / n2;
                if (sin_out >= 1.0f) {
                    r = 1.0f;
                } else {
                    float cos_out = __builtin_sqrtf(1 - sin_out * sin_out);
                    float cap = cos_in * cos_out - sin_in * sin_out;
                    float cam = cos_in * cos_out + sin_in * sin_out;
                    float sap = sin_in * cos_out + cos_in * sin_out;
                    float sam = sin_in * cos_out - cos_in * sin_out;
                    r = 0.5f * (sam * sam) * ((cam * cam) + (cap * cap)) / ((sap * sap) * (cam * cam));
                }
            }
            f3 ey = normalize(cross(nv, cross(nv, ph.v)));
            int32_t new_id;
            float m_a, m_s;
            if (x > r) {  // Transmit (.cl:506-529)
                float sin_t = __builtin_sqrtf(clamp0(1 - cos_in * cos_in)) * n1 / n2;
                float cos_t = -__builtin_sqrtf(clamp0(1 - sin_t * sin_t));
                if (cos_in > 0.0f) cos_t = __builtin_fabsf(cos_t);
                ph.v = normalize(sub(mul(cos_t, nv), mul(sin_t, ey)));
                new_id = (cos_in > 0) ? out_id : in_id;
                m_a = nm.y;
                m_s = nm.z;
                kind = 1;
            } else {  // Reflect (.cl:532-549)
                float sin_r = __builtin_sqrtf(clamp0(1 - cos_in * cos_in));
                ph.v = neg(normalize(add(mul(cos_in, nv), mul(sin_r, ey))));
                new_id = ph.curr;
                m_a = cm.y;
                m_s = cm.z;
                kind = 2;
            }
            if (m_a == 0 || m_s == 0) {  // leaves the tissue: record and kill
                // transmitted past distMin, else reflected (and fetal-tagged);
                // written as value selects so that sft stays in registers
                float dist = __builtin_fabsf(dot(eZ, sub(ph.p, POI)));
                float wr = ph.w * (1 - lmat[ph.inst * ntp1 + a.ntissue].x);  // (1 - Rsp)
                bool trans = dist > a.dist_min;
                xsft.z = trans ? wr : 0.0f;
                xsft.x = trans ? 0.0f : wr;
                xsft.y = (!trans && (ph.fetal & 1)) ? wr : 0.0f;
                ph.w = 0.0f;
            }
            if (new_id != ph.curr) {
                if (a.region) atomic_add_u64(&a.region[ph.inst * a.ntissue + ph.curr], reg_acc);
                reg_acc = 0;
            }
            ph.curr = new_id;
            ph.ref = h.id;
            if (ph.curr == a.fetal_id) ph.fetal |= 1;
            ph.unfin = (s - stb) * (cm.y + cm.z);
            if constexpr (!ADJ) {
                if (crec) {
                    uint32_t* u = reinterpret_cast<uint32_t*>(crec);
                    u[8] = (uint32_t)h.id;
                    crec[9] = stb;
                    crec[10] = r;
                    crec[11] = ph.v.x;
                    crec[12] = ph.v.y;
                    crec[13] = ph.v.z;
                    u[14] = (uint32_t)ph.curr;
                    u[15] = (uint32_t)ph.nvc;
                    u[16] = (uint32_t)kind;
                    crec[17] = cm.y + cm.z;
                    crawled = true;
                }
            }
        }
        return kind;
    };

    // Roulette (.cl:1020-1035, 1539-1545), trace, step cap; false when the
    // photon has ended
    // (inlined at each of its call sites: an out-of-line call would take the
    // photon state by address, into scratch memory)
    auto end_step = [&](int kind, int hid, float s) __attribute__((always_inline)) -> bool {
        if (ph.w < a.w_th && ph.w > 0.0f) {
            float x = rng.next();
            if (x < a.chance) {
                ph.w /= a.chance;
                ph.roulette = 1;
            } else {
                ph.w = 0;
                ph.roulette = 0;
            }
        }
        ph.steps++;
        if (TRACE && ph.steps <= (uint32_t)a.trace_max) {
            float* t = a.trace + ((int64_t)ph.k * a.trace_max + ((int64_t)ph.steps - 1)) * 12;
            t[0] = (float)kind;
            t[1] = (float)hid;
            t[2] = (float)ph.curr;
            t[3] = ph.w;
            t[4] = ph.p.x;
            t[5] = ph.p.y;
            t[6] = ph.p.z;
            t[7] = ph.v.x;
            t[8] = ph.v.y;
            t[9] = ph.v.z;
            t[10] = s;
            t[11] = (float)(TRACE ? rng.ndraw : 0);
        }
        // step cap, and the boundary-run cap: a photon caught on the edge of
        // two faces crosses from one to the other at t ~ 0 for ever (the
        // last-face exclusion cannot separate them), which the reference
        // loop never ends
        ph.fetal = kind == 0 ? (ph.fetal & 1) : ph.fetal + 2;
        if (ph.w > 0.0f && ((a.max_steps > 0 && (int64_t)ph.steps >= a.max_steps) ||
                            (a.max_bnd_run > 0 && ((uint32_t)ph.fetal >> 1) >= a.max_bnd_run))) {
            wcount(wCap);
            ph.w = 0.0f;
        }
        return ph.w > 0.0f && (ph.w > a.w_th || ph.roulette);
    };

    // An edge crawl, fast-forwarded.  A photon on the shared edge of two
    // faces crosses them in turn without moving, its state (direction, last
    // face, tissue) running through a short cycle.  When the crossings of
    // steps n - L .. n - 1 (L <= kCrawlSlots) are all recorded and the first
    // one's input state is the present one, the crossings that follow repeat
    // them in order: the nearest face ahead, its t, the Fresnel coefficient,
    // the outgoing direction, tissue and normal are functions of that state
    // and of the launch's optics alone.  The step length is recomputed from
    // the unfinished optical depth as the step draw computes it, and each
    // crossing consumes its Fresnel draw from the stream as the boundary event
    // does.  The replay stops, leaving the next crossing to the general path,
    // when the step no longer reaches the face, when the draw picks the other
    // outcome, when the weight lies below the roulette threshold (roulette
    // would draw), or when no optical depth is left; end_step ends the photon
    // as it would (step or run cap).  Draws, states, tallies and trace rows are
    // those of taking the crossings one iteration at a time.  False when the
    // photon has ended.
    auto crawl_forward = [&]() -> bool {
        const uint32_t gid = blockIdx.x * kBlock + tid;
        const float* ring = a.crawl + (size_t)gid * kCrawlSlots * kCrawlWords;
        const uint32_t k0 = (uint32_t)ph.k, k1 = (uint32_t)((uint64_t)ph.k >> 32);
        // the shortest cycle: a record of this photon (in this launch: the
        // kernel voids older ones) for each of the last L steps, the first of
        // them leaving from the present state
        uint32_t L = 0;
#pragma unroll
        for (uint32_t l = 2; l <= (uint32_t)kCrawlSlots; ++l) {
            const uint32_t* u =
                reinterpret_cast<const uint32_t*>(ring + ((ph.steps - l) & (kCrawlSlots - 1)) * kCrawlWords);
            if (u[0] != k0 || u[1] != k1 || u[2] != ph.steps - l) break;  // not recorded: no cycle
            if (u[3] == __float_as_uint(ph.v.x) && u[4] == __float_as_uint(ph.v.y) &&
                u[5] == __float_as_uint(ph.v.z) && (int32_t)u[6] == ph.ref && (int32_t)u[7] == ph.curr) {
                L = l;
                break;
            }
        }
        if (L == 0) return true;
        const uint32_t first = ph.steps - L;  // the step whose record is replayed first
        uint32_t j = 0;                       // ... and the place in the cycle of the next crossing
        uint32_t nff = 0;
        bool live = true;
        for (;;) {
            const float* rx = ring + ((first + j) & (kCrawlSlots - 1)) * kCrawlWords;
            const uint32_t* ux = reinterpret_cast<const uint32_t*>(rx);
            const float t = rx[9], r = rx[10], m = rx[17];
            if (ph.unfin == 0.0f || !(m > 0.0f) || ph.w < a.w_th) break;
            const float s = ph.unfin / m;  // the step draw's division
            if (!(t <= s)) break;           // the face lies beyond the step: an interaction
            const int32_t kd = (int32_t)ux[16];
            rng.top_up();
            if ((rng.peek() > r) != (kd == 1)) break;  // the other outcome
            (void)rng.next();
            const int32_t c = (int32_t)ux[14];
            if (c != ph.curr) {
                if (a.region && reg_acc) atomic_add_u64(&a.region[ph.inst * a.ntissue + ph.curr], reg_acc);
                reg_acc = 0;
            }
            ph.v = mk(rx[11], rx[12], rx[13]);
            ph.nvc = (int32_t)ux[15];
            ph.curr = c;
            const int32_t f = (int32_t)ux[8];
            ph.ref = f;
            if (ph.curr == a.fetal_id) ph.fetal |= 1;
            ph.unfin = (s - t) * m;
            ++nff;
            if (!end_step(kd, f, s)) {
                live = false;
                break;
            }
            j = (j + 1 == L) ? 0u : j + 1;
        }
        wsum(wBnd, nff);
        wsum(wCrawl, nff);
        return live;
    };

    // epilogue (.cl:1586-1594) + host binning (cpp:1553-1630)
    auto epilogue = [&]() {
        if (ph.p.x != ph.p.x || ph.p.y != ph.p.y || ph.p.z != ph.p.z || ph.v.x != ph.v.x || ph.v.y != ph.v.y ||
            ph.v.z != ph.v.z)
            wcount(wNan);
        if constexpr (ADJ) {
            // the replay must end as the forward photon did
            if (xsft.x != adj_w) wcount(wMismatch);
            const float W = adj_w;
            for (int32_t t = 0; t < a.ntissue; ++t) {
                const float L = adjL[t * kBlock + tid];
                const uint32_t n = adjN[t * kBlock + tid];
                if (a.pair_len) {
                    a.pair_len[adj_k * a.ntissue + t] = L;
                    a.pair_int[adj_k * a.ntissue + t] = n;
                }
                if (L == 0.0f && n == 0u) continue;
                const float mus = lmat[t].z;
                const float gm = -(L * W);
                const float gs = mus != 0.0f ? (((float)n / mus) - L) * W : gm;
                unsigned long long* g = a.grad + ((int64_t)adj_det * a.ntissue + t) * 2;
                atomic_add_u64(&g[0], fx32(gm));
                atomic_add_u64(&g[1], fx32(gs));
            }
            return;
        }
        if (a.region) atomic_add_u64(&a.region[ph.inst * a.ntissue + ph.curr], reg_acc);
        reg_acc = 0;
        if (vox_cell >= 0) atomic_add_u64(&a.vox[vox_cell], vox_acc);
        vox_cell = -1;
        vox_acc = 0;
        const f3 dq = sub(ph.p, POI);
        const f3 dn = sub(POI, ph.p);
        const float e0 = dot(eX, dq), e1 = dot(eY, dq), e2 = dot(eZ, dq);
        const float e3 = xsft.x, e4 = xsft.y, e5 = xsft.z;
        f3 nv = mk(0.0f, 0.0f, -1.0f);  // the last boundary normal, as the boundary event computed it
        if (ph.nvc >= 0) {
            const int32_t id = ph.nvc >> 1;
            nv = normalize(mk(a.nrm[3 * id], a.nrm[3 * id + 1], a.nrm[3 * id + 2]));
            if (ph.nvc & 1) nv = mul(sign_cl(dot(nv, eZr)), eZ);
        }
        const float e6 = __builtin_fabsf(acosf_lp(dot(ph.v, nv)));
        const float e7 = __builtin_sqrtf((dn.x * dn.x + dn.y * dn.y) + dn.z * dn.z);
        if (a.exit_rec) {
            float4* o = reinterpret_cast<float4*>(a.exit_rec + 8 * ph.k);
            o[0] = make_float4(e0, e1, e2, e3);
            o[1] = make_float4(e4, e5, e6, e7);
        }
        if (TRACE) a.trace_n[ph.k] = (int32_t)(ph.steps < (uint32_t)a.trace_max ? ph.steps : (uint32_t)a.trace_max);
        if (e3 != 0.0f || e4 != 0.0f || e5 != 0.0f) {
            unsigned long long q0 = fx32(e3), q1 = fx32(e4), q2 = fx32(e5);
            if (a.sft) {
                int32_t ex0 = -(a.dimx / 2), ex1 = a.dimx / 2, ey0 = -(a.dimy / 2), ey1 = a.dimy / 2;
                int32_t x = cvt_i32f(e0 / a.det_grid[0]);
                int32_t y = cvt_i32f(e1 / a.det_grid[1]);
                y = (y < ey0) ? ey0 : y;
                y = (y > ey1) ? ey1 : y;
                if (x < ex0) x = ex0;
                else if (x > ex1) x = ex1;
                int64_t px = ((int64_t)(y - ey0) * a.sft_nx + (x - ex0)) * 3;
                atomic_add_u64(&a.sft[px], q0);
                atomic_add_u64(&a.sft[px + 1], q1);
                atomic_add_u64(&a.sft[px + 2], q2);
            }
            if (a.rtd) {
                int32_t ia = cvt_i32(((double)(e6 * 180.0f) / 3.14159265358979323846) / (double)a.det_grid[5]);
                int32_t ir = cvt_i32f(e7 / a.det_grid[3]);
                if (ir > a.dimR - 1) ir = a.dimR - 1;
                else if (ir < 0) ir = 0;
                if (ia > a.dimAlpha - 1) ia = a.dimAlpha - 1;
                else if (ia < 0) ia = 0;
                int64_t px = ((int64_t)ia * a.dimR + ir) * 3;
                atomic_add_u64(&a.rtd[px], q0);
                atomic_add_u64(&a.rtd[px + 1], q1);
                atomic_add_u64(&a.rtd[px + 2], q2);
            }
            if (a.rad) {
                int32_t ib = cvt_i32f(e7 / a.rad_dr);
                if (ib >= 0 && ib < a.rad_n) {
                    atomic_add_u64(&a.rad[3 * ib], q0);
                    atomic_add_u64(&a.rad[3 * ib + 1], q1);
                    atomic_add_u64(&a.rad[3 * ib + 2], q2);
                }
            }
            if (a.det) {
                for (int kd = 0; kd < a.ndet; ++kd) {
                    if (e7 >= a.det_rlo[kd] && e7 < a.det_rhi[kd]) {
                        const int64_t di = 2 * ((int64_t)ph.inst * a.ndet + kd);
                        atomic_add_u64(&a.det[di], q0);
                        atomic_add_u64(&a.det[di + 1], q1);
                    }
                }
            }
        }
        if (a.det_list && e3 != 0.0f) {  // detected photons, for the adjoint replay
            for (int kd = 0; kd < a.ndet; ++kd) {
                if (e7 >= a.det_rlo[kd] && e7 < a.det_rhi[kd]) {
                    const unsigned long long slot = atomicAdd(a.det_count, 1ull);
                    if (slot < (unsigned long long)a.det_cap) {
                        DetRec r;
                        r.photon = a.first + ph.k;
                        r.det = kd;
                        r.w = e3;
                        a.det_list[slot] = r;
                    }
                }
            }
        }
    };

    for (;;) {
        // ---- refill lanes whose photon has finished (one atomic per wave) ----
        if (!exhausted) {
            unsigned long long need = __ballot(!alive);
            if (need) {
                int nneed = __popcll(need);
                int leader = __ffsll((long long)need) - 1;
                unsigned long long got = 0;
                if (lane == leader) got = atomicAdd(a.counter, (unsigned long long)nneed);
                uint32_t blo = __shfl((uint32_t)got, leader), bhi = __shfl((uint32_t)(got >> 32), leader);
                unsigned long long base = ((unsigned long long)bhi << 32) | blo;
                int rank = __popcll(need & ((1ull << lane) - 1ull));
                if (base + (unsigned long long)nneed >= (unsigned long long)a.count) exhausted = true;
                if (!alive) {
                    unsigned long long k = base + (unsigned long long)rank;
                    if (k < (unsigned long long)a.count) {
                        int64_t id = a.first + (int64_t)k;
                        ph.inst = 0;
                        if (a.inst_mat) {  // batched: k = instance * inst_count + index within it
                            ph.inst = (int32_t)(k / (unsigned long long)a.inst_count);
                            id = a.first + (int64_t)(k - (unsigned long long)ph.inst * (unsigned long long)a.inst_count);
                        }
                        if constexpr (ADJ) {
                            const DetRec r = a.adj_list[k];
                            id = r.photon;
                            adj_det = r.det;
                            adj_w = r.w;
                            adj_k = (int64_t)k;
                            for (int32_t t = 0; t < a.ntissue; ++t) {
                                adjL[t * kBlock + tid] = 0.0f;
                                adjN[t * kBlock + tid] = 0u;
                            }
                        }
                        rng.init(a.seed, (uint64_t)id);
                        // GetvStart (.cl:423-467): A == 1 -> vStart[0]; A > 1 by cumulative counts
                        int dsel = 0;
                        if (a.nstart > 1) {
                            int64_t sum = 0;
                            for (int i = 0; i < a.nstart; ++i) {
                                sum += a.vocc[i];
                                dsel = i;
                                if (id < sum) break;
                            }
                        }
                        // PhotonClass::Init (.cl:344-389)
                        ph.v = mk(a.vstart[3 * dsel], a.vstart[3 * dsel + 1], a.vstart[3 * dsel + 2]);
                        ph.p = mk(a.pstart[0], a.pstart[1], a.pstart[2]);
                        ph.w = 1.0f - lmat[ph.inst * ntp1 + a.ntissue].x;
                        ph.unfin = 0.0f;
                        ph.fetal = 0;
                        ph.ref = -1;
                        ph.curr = 0;
                        ph.nvc = -1;
                        ph.roulette = 1;
                        ph.steps = 0;
                        ph.safe = 0.0f;
                        ph.anc = ph.p;
                        ph.k = (int64_t)k;
                        reg_acc = 0;
                        alive = true;
                        have_step = false;
                        search = false;
                        wcount(wPhot);
                    }
                }
            }
        }
        if (!__any(alive)) break;

        // ---- one round ----
        // Each lane draws its step length (.cl:1405-1442) when none is
        // pending.  A step whose end point stays in the lane's free region
        // needs no face search and is taken at once; any other step waits for
        // a search.  Waiting lanes are served in batches, by kind: short steps
        // by the near-face lists, long ones (and lanes with no list) by the
        // BVH.  A kind's batch runs once enough live lanes wait for it, or
        // when no lane can move otherwise, or at the iteration cap that ends
        // the round; a round also ends once enough lanes have finished, so
        // that they are refilled.
        for (int it = 0;; ++it) {
            launder();
            // the Philox key from the argument segment each iteration, so that
            // its round keys are formed in scalar registers where used rather
            // than held for the whole loop (and spilled to vector lanes)
            rng.k0 = (uint32_t)a.seed;
            rng.k1 = (uint32_t)(a.seed >> 32);
            bool go = false;     // this lane takes a step in this iteration
            bool freed = false;  // ... and the step was shown to stay in the free region
            bool skipped = false;  // ... with no face search
            bool bnd = false;
            crawled = false;
            Hit h;
            h.id = -1;
            h.t = 10000.0f;
            xsft = mk(0.0f, 0.0f, 0.0f);
            if (alive) rng.top_up();
            bool act = alive && !search;
            // the record of the near-grid cell holding p (list start, clearance,
            // the next cell's list start), loaded before the step length is
            // drawn so that the load overlaps the draw: the clearance check
            // and a list query from p use it
            int32_t pre_c = -1;
            uint32_t pre_first = 0u, pre_bits = 0u, pre_next = 0u;
            if (alive && a.near_cells) {
                pre_c = near_cell(a, ph.p);
                const uint32_t ci = (uint32_t)(pre_c >= 0 ? pre_c : 0) + geom_of<MG>(a, lmat, ph.inst).coff;
                const uint32_t* cw = reinterpret_cast<const uint32_t*>(a.near_cells) + 2 * (size_t)ci;
                pre_first = cw[0];
                pre_bits = cw[1];
                pre_next = cw[2];
            }
            if (act) {
                if (!have_step) {
                    const float4 cm = lmat[ph.inst * ntp1 + ph.curr];
                    ballistic = (cm.y == 0.0f && cm.z == 0.0f);
                    if (!ballistic) {
                        if (ph.unfin == 0.0f) {
                            float x;
                            do {
                                x = rng.next();
                            } while (x <= 0.0f);
                            stepsize = -logf_u01(x) * a.factor / (cm.y + cm.z);
                        } else {
                            stepsize = ph.unfin / (cm.y + cm.z);
                            ph.unfin = 0.0f;
                        }
                    }
                    have_step = true;
                }
                // free step: the end point stays inside the free region
                const f3 q = add(ph.p, mul(stepsize, ph.v));
                const f3 e = sub(q, ph.anc);
                freed = !a.no_skip && !ballistic && dot(e, e) < ph.safe * ph.safe && dot(ph.pu, q) > ph.pc;
                if (!freed && !a.no_skip && !ballistic && a.near_cells && a.clear_step >= 0.0f &&
                    stepsize >= a.clear_step) {
                    // else the ball of the near-grid cell's clearance around p:
                    // no face lies within the clearance of any point the cell
                    // lookup places in the cell
                    if (pre_c >= 0) {
                        const float cl = __uint_as_float(pre_bits);
                        const float rad = fmaxf(fmaxf(__builtin_fabsf(ph.p.x), __builtin_fabsf(ph.p.y)),
                                                __builtin_fabsf(ph.p.z));
                        const float sc = cl - (a.safe_margin + 1e-5f * (cl + rad));
                        const f3 eq = sub(q, ph.p);
                        if (sc > 0.0f && dot(eq, eq) < sc * sc) {
                            ph.anc = ph.p;
                            ph.safe = sc;
                            ph.pu = mk(0.0f, 0.0f, 0.0f);
                            ph.pc = -1.0f;
                            freed = true;
                        }
                    }
                }
                if (freed || (!a.no_skip && !ballistic && a.safety && stepsize < safety_at(a, ph.p))) {
                    go = true;
                    skipped = true;
                } else {
                    search = true;
                    mq = 0.0f;
                    const bool finite_p =
                        __builtin_isfinite(ph.p.x) && __builtin_isfinite(ph.p.y) && __builtin_isfinite(ph.p.z);
                    // the near-face lists answer a step of at most r0, and a
                    // longer step or a ballistic lane's flight when its face
                    // lies within r0 (a miss there falls back to the BVH)
                    by_list = a.near_cells && finite_p && (ballistic || stepsize <= a.near_r0 || a.list_long);
                }
            }
            const int nlive = __popcll(__ballot(alive));
            const int nlist = __popcll(__ballot(search && by_list));
            const int nbvh = __popcll(__ballot(search && !by_list));
            const bool stuck = !__any(go);
            // a lane in a run of boundary events (a photon crawling along the
            // shared edge of two faces advances by the rounding-level t of each
            // crossing, some 10^4-10^5 of them) is served at once rather than
            // waiting for a full batch: only how soon, never what, changes
            const bool urgent = __any(search && ((uint32_t)ph.fetal >> 1) >= kUrgentRun) ||
                                (exhausted && nlive <= kTailLanes);
            const bool at_cap = it + 1 >= a.free_iters;
            const bool do_list = nlist > 0 && (stuck || at_cap || urgent || nlist >= a.list_min * nlive);
            const bool do_bvh = nbvh > 0 && (stuck || at_cap || urgent || nbvh >= a.search_min * nlive);
            ++r_iter;
            if (do_list) {
                ++r_lr;
                // exact answer from the cells' lists, and the free region with it
                const bool part = search && by_list;
                float d1sq, d2sq, unlisted;
                int32_t i1;
                // a flight or a step longer than r0: the lists see only r0 of
                // it from the point the march along it has reached (every
                // face within r0 of that point is listed and the ray before
                // it is free of faces, so a hit there is the nearest face of
                // the whole mesh)
                const bool reach = ballistic || stepsize > a.near_r0;
                const bool want = !a.no_skip && ph.ref < 0 && !reach;
                const float tlim = ballistic ? 10000.0f : stepsize;
                const uint32_t coff = part ? geom_of<MG>(a, lmat, ph.inst).coff : 0u;
                f3 cq = ph.p;
                float wmax = reach ? a.near_r0 : stepsize;
                const bool at_p = !(reach && mq > 0.0f);
                if (part && !at_p) {
                    cq = add(ph.p, mul(mq, ph.v));
                    const float rq = fmaxf(fmaxf(__builtin_fabsf(cq.x), __builtin_fabsf(cq.y)), __builtin_fabsf(cq.z));
                    wmax = fminf(mq + (a.near_r0 - (a.safe_margin + 1e-5f * (a.near_r0 + rq))), tlim);
                }
                Hit hl;
                uint32_t pairs;
                near_batch(a, nred[tid >> 6], part, coff, cq, ph.p, ph.v, ph.ref, wmax, want, at_p, pre_c, pre_first,
                           pre_bits, pre_next, hl, d1sq, d2sq, i1, unlisted, pairs);
                r_lst += pairs;
                if (part && reach) {
                    if (hl.id >= 0) {  // the nearest face ahead
                        h = hl;
                        if (ballistic) stepsize = h.t;
                        bnd = true;
                        search = false;
                        go = true;
                    } else if (!a.march || near_cell(a, ph.p) < 0) {
                        by_list = false;  // the BVH
                    } else {
                        // March on: from each point the cell's clearance (no
                        // face within it, less a rounding margin) frees the
                        // ray ahead, until a face is near (the next list
                        // query looks from there), the step ends, or the ray
                        // leaves the grid (a box reaching beyond every face:
                        // nothing lies ahead then).
                        float cleared = wmax;
                        bool exited = false;
                        int n = 0;
                        while (cleared < tlim && n < 256) {
                            ++n;
                            const f3 q = add(ph.p, mul(cleared, ph.v));
                            const int32_t c = near_cell(a, q);
                            if (c < 0) {
                                exited = true;
                                break;
                            }
                            const float cl = __uint_as_float(a.near_cells[(uint32_t)c + coff].clear_bits);
                            const float rq =
                                fmaxf(fmaxf(__builtin_fabsf(q.x), __builtin_fabsf(q.y)), __builtin_fabsf(q.z));
                            const float sc = cl - (a.safe_margin + 1e-5f * (cl + rq));
                            if (!(sc > 0.0f)) break;
                            cleared += sc;
                        }
                        wcount(wMarch);
                        if (exited || cleared >= tlim) {
                            // no face within the step: a flight (or a step of
                            // 10^4 or more) leaves the mesh, a step ends in
                            // its tissue
                            h.id = -1;
                            h.t = 10000.0f;
                            bnd = ballistic || stepsize >= 10000.0f;
                            if (ballistic) stepsize = h.t;
                            search = false;
                            go = true;
                        } else if (n >= 256) {
                            by_list = false;  // a march this long: the BVH
                        } else {
                            mq = cleared;  // a face is near: the next list query looks from here
                        }
                    }
                } else if (part) {
                    h = hl;
                    if (want) {
                        const float rad = fmaxf(fmaxf(__builtin_fabsf(ph.p.x), __builtin_fabsf(ph.p.y)),
                                                __builtin_fabsf(ph.p.z));
                        const float margin = a.safe_margin + 1e-5f * (a.near_r0 + rad);
                        // cut by the nearest face's plane when p is clearly on
                        // one side of it: the ball may then reach the
                        // second-nearest face
                        float d = fminf(__builtin_sqrtf(d1sq), unlisted);
                        ph.pu = mk(0.0f, 0.0f, 0.0f);
                        ph.pc = -1.0f;
                        if (i1 >= 0 && d1sq < unlisted * unlisted) {
                            const float4* tr = reinterpret_cast<const float4*>(a.near_tris);
                            const float4 A = tr[3 * (int64_t)i1], B = tr[3 * (int64_t)i1 + 1],
                                         C = tr[3 * (int64_t)i1 + 2];
                            const f3 n = cross(mk(B.x, B.y, B.z), mk(C.x, C.y, C.z));
                            const f3 u = mul(1.0f / __builtin_sqrtf(dot(n, n)), n);
                            const float c0 = dot(u, mk(A.x, A.y, A.z));
                            const float sd = dot(u, ph.p) - c0;
                            if (__builtin_fabsf(sd) > 4.0f * margin) {  // false for a degenerate face (NaN)
                                const float sg = sd > 0.0f ? 1.0f : -1.0f;
                                ph.pu = mul(sg, u);
                                ph.pc = sg * c0 + margin;
                                d = fminf(__builtin_sqrtf(d2sq), unlisted);
                            }
                        }
                        ph.safe = fmaxf(0.0f, d - (a.safe_margin + 1e-5f * (d + rad)));  // >= 0: the test squares it
                        ph.anc = ph.p;
                        wcount(wDq);
                    }
                    bnd = h.id >= 0;
                    if (h.id < 0) h.t = 10000.0f;
                    search = false;
                    go = true;
                }
            }
            if (do_bvh) {
                if (lane == 0) atomicAdd(&wc[wBRounds], 1ull);
                const bool sb = search && !by_list;
                // refresh the free radius of waiting lanes -- the nearest-face
                // distance less a rounding margin (a face found by the exact
                // test at t <= s lies within s + margin of p)
                const bool finite_p =
                    __builtin_isfinite(ph.p.x) && __builtin_isfinite(ph.p.y) && __builtin_isfinite(ph.p.z);
                // a distance query capped below the step cannot free it
                bool dq = sb && !a.no_skip && !ballistic && a.dist_cap > 0.0f && ph.ref < 0 && finite_p &&
                          stepsize < fminf(a.dist_cap, a.dist_cap_steps * stepsize);
                if (__any(dq)) {
                    if (dq) {
                        float cap = fminf(a.dist_cap, a.dist_cap_steps * stepsize);
                        uint32_t dvis = 0;
                        float d = mesh_distance(a, lnodes, geom_of<MG>(a, lmat, ph.inst).root, ph.p, cap, dvis);
                        wsum(wDvis, dvis);
                        float rad = fmaxf(fmaxf(__builtin_fabsf(ph.p.x), __builtin_fabsf(ph.p.y)),
                                          __builtin_fabsf(ph.p.z));
                        ph.safe = fmaxf(0.0f, d - (a.safe_margin + 1e-5f * (d + rad)));  // >= 0: the test squares it
                        ph.anc = ph.p;
                        ph.pu = mk(0.0f, 0.0f, 0.0f);
                        ph.pc = -1.0f;
                        wcount(wDq);
                        if (stepsize < ph.safe) {
                            search = false;
                            go = true;
                            skipped = true;
                        }
                    }
                }
                // nearest face within the pending step
                if (search && !by_list) {
                    uint32_t vis = 0, tst = 0;
                    find_face(a, lnodes, geom_of<MG>(a, lmat, ph.inst).root, ph.p, ph.v, ph.ref,
                              ballistic ? 10000.0f : stepsize, h, vis, tst);
                    wsum(wVis, vis);
                    wsum(wTst, tst);
                    bnd = h.id >= 0 || ballistic || stepsize >= 10000.0f;  // no face ahead at all: escape
                    if (h.id < 0) h.t = 10000.0f;
                    if (ballistic) stepsize = h.t;
                    search = false;
                    go = true;
                }
            }
            // ---- the step itself (the reference loop body, .cl:1375-1575) ----
            if (go) {
                int kind = 0;
                if (bnd) {
                    kind = boundary_event(h, stepsize, h.t);
                } else {
                    interaction(stepsize);
                    // keep "p lies in the free region" true: a searched step
                    // may end outside it, and a later step from outside could
                    // cross a face even though it ends inside
                    const f3 e = sub(ph.p, ph.anc);
                    if (!freed && !(dot(e, e) < ph.safe * ph.safe && dot(ph.pu, ph.p) > ph.pc)) ph.safe = 0.0f;
                }
                have_step = false;
                bool live = end_step(kind, bnd ? h.id : -1, stepsize);
                if constexpr (!ADJ) {
                    if (crawled && live && ((uint32_t)ph.fetal >> 1) >= 2) live = crawl_forward();
                }
                if (!live) {
                    epilogue();
                    alive = false;
                }
            }
            r_int += (uint32_t)__popcll(__ballot(go && !bnd));
            r_bnd += (uint32_t)__popcll(__ballot(go && bnd));
            r_skip += (uint32_t)__popcll(__ballot(skipped));
            // ---- further steps in the same iteration ----
            // A lane whose step was an interaction draws its next step now and
            // takes it at once when its end point stays in the free region (the
            // same test as at the top of the iteration; the cell clearance is
            // not tried here), up to second_step times.  Otherwise the drawn
            // step waits for the next iteration, which finds have_step set.
            // The per-iteration work of the wavefront (refill, random top-up,
            // batch decisions, its scalar bookkeeping) is then shared by several
            // steps of most lanes; each photon's steps, draws and deposits are
            // the same as with one step per iteration.
            bool prev = go && !bnd;  // this lane's last step was an interaction
            // (not while a lane of the wavefront crawls along an edge: each of
            // its crossings takes an iteration, so iterations stay short)
            const int extra = __any(alive && ((uint32_t)ph.fetal >> 1) >= kUrgentRun) ? 0 : a.second_step;
            for (int k2 = 0; k2 < extra && __any(prev && alive && !ballistic); ++k2) {
                const bool cand = prev && alive && !ballistic;
                bool go2 = false;
                if (cand) rng.top_up();
                if (cand) {
                    const float4 cm = lmat[ph.inst * ntp1 + ph.curr];
                    if (ph.unfin == 0.0f) {
                        float x;
                        do {
                            x = rng.next();
                        } while (x <= 0.0f);
                        stepsize = -logf_u01(x) * a.factor / (cm.y + cm.z);
                    } else {
                        stepsize = ph.unfin / (cm.y + cm.z);
                        ph.unfin = 0.0f;
                    }
                    have_step = true;
                    const f3 q = add(ph.p, mul(stepsize, ph.v));
                    const f3 e = sub(q, ph.anc);
                    go2 = !a.no_skip && dot(e, e) < ph.safe * ph.safe && dot(ph.pu, q) > ph.pc;
                }
                if (go2) {
                    interaction(stepsize);
                    have_step = false;
                    if (!end_step(0, -1, stepsize)) {
                        epilogue();
                        alive = false;
                    }
                }
                const uint32_t n2 = (uint32_t)__popcll(__ballot(go2));
                r_int += n2;
                r_skip += n2;
                prev = go2;
            }
            // end the round at the cap (finished lanes are refilled then), or
            // once enough lanes have finished and photons are left to start
            if (at_cap || (!exhausted && 64 - __popcll(__ballot(alive)) >= a.refill_min)) break;
        }
        if (lane == 0) {  // the round's counts into the wavefront's row
            wc[wInt] += r_int;
            wc[wBnd] += r_bnd;
            wc[wSkip] += r_skip;
            wc[wIter] += r_iter;
            wc[wLRounds] += r_lr;
            wc[wLst] += r_lst;
        }
        r_int = r_bnd = r_skip = r_iter = r_lr = r_lst = 0;
    }
    if (a.misc) {
        __threadfence_block();
        if (lane == 0) {
            atomic_add_u64(&a.misc[kMiscPhotons], wc[wPhot]);
            atomic_add_u64(&a.misc[kMiscSteps], wc[wInt] + wc[wBnd]);  // every step is one or the other
            atomic_add_u64(&a.misc[kMiscInteract], wc[wInt]);
            atomic_add_u64(&a.misc[kMiscBoundary], wc[wBnd]);
            atomic_add_u64(&a.misc[kMiscEscaped], wc[wEsc]);
            atomic_add_u64(&a.misc[kMiscCapped], wc[wCap]);
            atomic_add_u64(&a.misc[kMiscSkipped], wc[wSkip]);
            atomic_add_u64(&a.misc[kMiscDistQueries], wc[wDq]);
            atomic_add_u64(&a.misc[kMiscNan], wc[wNan]);
            atomic_add_u64(&a.misc[kMiscListTests], wc[wLst]);
            atomic_add_u64(&a.misc[kMiscIters], wc[wIter]);
            atomic_add_u64(&a.misc[kMiscRounds], wc[wLRounds]);
            atomic_add_u64(&a.misc[kMiscBvhRounds], wc[wBRounds]);
            atomic_add_u64(&a.misc[kMiscArzOut], wc[wArzOut]);
            atomic_add_u64(&a.misc[kMiscVoxOut], wc[wVoxOut]);
            atomic_add_u64(&a.misc[kMiscReplayMismatch], wc[wMismatch]);
            atomic_add_u64(&a.misc[kMiscMarches], wc[wMarch]);
            atomic_add_u64(&a.misc[kMiscCrawl], wc[wCrawl]);
            atomic_add_u64(&a.misc[kMiscNodeVisits], wc[wVis]);
            atomic_add_u64(&a.misc[kMiscTriTests], wc[wTst]);
            atomic_add_u64(&a.misc[kMiscDistVisits], wc[wDvis]);
        }
    }
#undef a
#undef POI
#undef eX
#undef eY
#undef eZ
#undef eZr
}

// ---------------------------------------------------------------------------
// static scratch of the allocation-free batched launch
// ---------------------------------------------------------------------------
// Each GPU holds its own copy until the process ends.  Host code uses the
// address hipGetSymbolAddress returns for the current device, never the
// array's name (a host-side stand-in that is not a device address).
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int static_scratch(int device, void** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices || !out) return 1;
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return 1;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

namespace {

// The kernel variant a launch with these arguments runs, and its dynamic LDS.
int pick_transport(const KArgs& a, const void** fn, size_t* lds, char* err, int errlen)
{
    *lds = (size_t)(2 * a.nlds + a.lmat_n) * sizeof(float4) + (size_t)a.lmat_n * sizeof(float);
    if (a.rng_mode == 1) *lds += sizeof(Ran3Lds);
    const int w = a.waves;
#define LP_PICK(R, T, W, J, M) (const void*)transport_kernel<R, T, W, J, M>
    if (a.ngeom > 1 && (a.adj_list || a.rng_mode == 1 || a.trace)) {
        snprintf(err, errlen, "transport launch: several geometries need the Philox streams, no trace or replay");
        return 1;
    }
    if (a.ngeom > 1) {
        *fn = w == 4   ? LP_PICK(0, false, 4, false, true)
              : w == 5 ? LP_PICK(0, false, 5, false, true)
              : w == 6 ? LP_PICK(0, false, 6, false, true)
                       : LP_PICK(0, false, 1, false, true);
    } else if (a.adj_list) {
        *lds += (size_t)kAdjMaxTissues * kBlock * (sizeof(float) + sizeof(uint32_t));
        *fn = a.rng_mode == 1 ? LP_PICK(1, false, 1, true, false) : LP_PICK(0, false, 1, true, false);
    } else if (a.rng_mode == 1) {
        *fn = a.trace ? LP_PICK(1, true, 1, false, false) : LP_PICK(1, false, 1, false, false);
    } else if (a.trace) {
        *fn = LP_PICK(0, true, 1, false, false);
    } else {
        *fn = w == 4   ? LP_PICK(0, false, 4, false, false)
              : w == 5 ? LP_PICK(0, false, 5, false, false)
              : w == 6 ? LP_PICK(0, false, 6, false, false)
                       : LP_PICK(0, false, 1, false, false);
    }
#undef LP_PICK
    return 0;
}

}  // namespace

int launch_transport(const KArgs& a, int grid, void* stream, char* err, int errlen)
{
    const void* fn = nullptr;
    size_t lds = 0;
    if (pick_transport(a, &fn, &lds, err, errlen)) return 1;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) {
        void* args[] = {const_cast<KArgs*>(&a)};
        e = hipLaunchKernel(fn, dim3(grid), dim3(kBlock), args, lds, (hipStream_t)stream);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        snprintf(err, errlen, "transport launch: %s", hipGetErrorString(e));
        return 1;
    }
    return 0;
}

int transport_resident_blocks(const KArgs& a)
{
    const void* fn = nullptr;
    size_t lds = 0;
    char err[128];
    if (pick_transport(a, &fn, &lds, err, sizeof err)) return 1;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) return 1;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kBlock, lds) != hipSuccess) return 1;
    return n > 0 ? n : 1;
}

// ---------------------------------------------------------------------------
// element-wise access to the math core, for the parity tests
// ---------------------------------------------------------------------------
__global__ void debug_math_kernel(int op, const float* in, float* out, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = in[2 * i], y = in[2 * i + 1];
    float r;
    switch (op) {
    case 0: r = __builtin_sqrtf(x); break;
    case 1: r = x / y; break;
    case 2: r = logf_lp(x); break;
    case 3: r = cosf_lp(x); break;
    case 4: r = acosf_lp(x); break;
    case 5: r = x * y; break;
    case 7: r = sqrt_n(x); break;
    case 8: r = recip_n(x); break;
    case 9: r = logf_u01(x); break;
    default: r = x + y; break;
    }
    out[i] = r;
}

// Plain streaming copy (16 B per lane per iteration): the reference point for
// converting the profiler's FETCH_SIZE / WRITE_SIZE counters into bytes.
__global__ void debug_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int launch_debug_copy(const void* src, void* dst, int64_t bytes, void* stream)
{
    int64_t n = bytes / 16;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(debug_copy_kernel, dim3(4096), dim3(256), 0, (hipStream_t)stream, (const float4*)src,
                       (float4*)dst, n);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Sets n 64-bit words to zero: clears tallies and the work counter inside
// captured HIP graphs (a kernel node rather than a memset node).
__global__ void zero_u64_kernel(unsigned long long* __restrict__ p, int64_t n)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

int launch_zero_u64(void* p, int64_t n, void* stream)
{
    if (n <= 0) return 0;
    int64_t blocks = std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (unsigned long long*)p, n);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_debug_math(int op, const float* in, float* out, int64_t n, void* stream)
{
    if (n <= 0) return 0;
    int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(debug_math_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, op, in, out, n);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace lp
