// lp_model.cpp -- scene construction (model layer) for liblp.so.
//
// Emitter model after src/Model/Emitter.cpp:186-339, cell normals in the
// role of vtkPolyDataNormals (src/IO/ImportExport.cpp:372-379), and
// generators for the synthetic scenes of the benchmark configurations
// (layered slab, maternal-fetal abdomen), which the reference does not ship.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/lightprop_model.h"

namespace {

using V3 = std::array<double, 3>;

V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 crs(const V3& a, const V3& b)
{
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double dt(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 unit(const V3& a)
{
    double l = std::sqrt(dt(a, a));
    return {a[0] / l, a[1] / l, a[2] / l};
}

void put_tri(float* tri, float* nrm, int32_t* med, int64_t k, const V3& a, const V3& b, const V3& c,
             int32_t inner, int32_t outer)
{
    const V3* p[3] = {&a, &b, &c};
    for (int v = 0; v < 3; ++v)
        for (int x = 0; x < 3; ++x) tri[9 * k + 3 * v + x] = (float)(*p[v])[x];
    // normal from the stored (float) vertices, in double
    V3 fa, fb, fc;
    for (int x = 0; x < 3; ++x) {
        fa[x] = tri[9 * k + x];
        fb[x] = tri[9 * k + 3 + x];
        fc[x] = tri[9 * k + 6 + x];
    }
    V3 n = unit(crs(sub(fb, fa), sub(fc, fa)));
    for (int x = 0; x < 3; ++x) nrm[3 * k + x] = (float)n[x];
    med[2 * k] = inner;
    med[2 * k + 1] = outer;
}

// Unit-sphere geodesic polyhedron: icosahedron faces split into freq^2
// triangles, points projected to the sphere.  Points on shared icosahedron
// edges are computed once (canonical direction) so the mesh is watertight.
struct Geodesic {
    std::vector<V3> pts;
    std::vector<std::array<int32_t, 3>> faces;
};

Geodesic make_geodesic(int n)
{
    const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<V3> ico = {{-1, phi, 0}, {1, phi, 0}, {-1, -phi, 0}, {1, -phi, 0}, {0, -1, phi}, {0, 1, phi},
                           {0, -1, -phi}, {0, 1, -phi}, {phi, 0, -1}, {phi, 0, 1}, {-phi, 0, -1}, {-phi, 0, 1}};
    // generic rotation so that no vertex or edge lies on the z axis
    const double ax = 0.31, ay = 0.53, az = 0.17;
    auto rot = [&](V3 p) {
        V3 q = {p[0], std::cos(ax) * p[1] - std::sin(ax) * p[2], std::sin(ax) * p[1] + std::cos(ax) * p[2]};
        V3 r = {std::cos(ay) * q[0] + std::sin(ay) * q[2], q[1], -std::sin(ay) * q[0] + std::cos(ay) * q[2]};
        return V3{std::cos(az) * r[0] - std::sin(az) * r[1], std::sin(az) * r[0] + std::cos(az) * r[1], r[2]};
    };
    for (auto& p : ico) p = rot(p);
    // faces: vertex triples at edge length 2, wound outward
    std::vector<std::array<int, 3>> f;
    for (int i = 0; i < 12; ++i)
        for (int j = i + 1; j < 12; ++j)
            for (int k = j + 1; k < 12; ++k) {
                auto near = [&](int a, int b) {
                    V3 d = sub(ico[a], ico[b]);
                    return std::fabs(dt(d, d) - 4.0) < 1e-9;
                };
                if (!(near(i, j) && near(j, k) && near(i, k))) continue;
                V3 c = {ico[i][0] + ico[j][0] + ico[k][0], ico[i][1] + ico[j][1] + ico[k][1],
                        ico[i][2] + ico[j][2] + ico[k][2]};
                if (dt(crs(sub(ico[j], ico[i]), sub(ico[k], ico[i])), c) > 0)
                    f.push_back({i, j, k});
                else
                    f.push_back({i, k, j});
            }
    Geodesic g;
    for (auto& p : ico) g.pts.push_back(unit(p));
    std::map<std::pair<int, int>, int32_t> edge_base;  // (u<v) -> index of point k=1
    auto edge_pt = [&](int u, int v, int k) -> int32_t {
        if (k == 0) return u;
        if (k == n) return v;
        int a = std::min(u, v), b = std::max(u, v);
        int kk = (u == a) ? k : n - k;
        auto it = edge_base.find({a, b});
        if (it == edge_base.end()) {
            int32_t base = (int32_t)g.pts.size();
            for (int s = 1; s < n; ++s) {
                double t = (double)s / n;
                V3 p = {ico[a][0] + (ico[b][0] - ico[a][0]) * t, ico[a][1] + (ico[b][1] - ico[a][1]) * t,
                        ico[a][2] + (ico[b][2] - ico[a][2]) * t};
                g.pts.push_back(unit(p));
            }
            it = edge_base.emplace(std::make_pair(a, b), base).first;
        }
        return it->second + kk - 1;
    };
    for (auto& F : f) {
        int A = F[0], B = F[1], Cc = F[2];
        std::vector<int32_t> idx((n + 1) * (n + 1), -1);
        auto at = [&](int i, int j) -> int32_t& { return idx[i * (n + 1) + j]; };
        for (int i = 0; i <= n; ++i)
            for (int j = 0; i + j <= n; ++j) {
                if (j == 0) at(i, j) = edge_pt(A, B, i);
                else if (i == 0) at(i, j) = edge_pt(A, Cc, j);
                else if (i + j == n) at(i, j) = edge_pt(B, Cc, j);
                else {
                    double s = (double)i / n, t = (double)j / n;
                    V3 p;
                    for (int x = 0; x < 3; ++x)
                        p[x] = ico[A][x] + (ico[B][x] - ico[A][x]) * s + (ico[Cc][x] - ico[A][x]) * t;
                    at(i, j) = (int32_t)g.pts.size();
                    g.pts.push_back(unit(p));
                }
            }
        for (int i = 0; i < n; ++i)
            for (int j = 0; i + j < n; ++j) {
                g.faces.push_back({at(i, j), at(i + 1, j), at(i, j + 1)});
                if (i + j + 1 < n) g.faces.push_back({at(i + 1, j), at(i + 1, j + 1), at(i, j + 1)});
            }
    }
    return g;
}

int64_t emit_sphere(const Geodesic& g, const V3& c, double r, int32_t inner, int32_t outer, float* tri, float* nrm,
                    int32_t* med, int64_t k)
{
    for (auto& F : g.faces) {
        V3 p[3];
        for (int v = 0; v < 3; ++v)
            for (int x = 0; x < 3; ++x) p[v][x] = c[x] + r * g.pts[F[v]][x];
        put_tri(tri, nrm, med, k++, p[0], p[1], p[2], inner, outer);
    }
    return k;
}

// Prahl's tabulation of the molar extinction of hemoglobin (cm^-1 / M),
// S. Prahl, "Tabulated molar extinction coefficient for hemoglobin in water",
// Oregon Medical Laser Center (1998), from Gratzer and Kollias.
struct HbRow {
    double nm, hbo2, hb;
};
const HbRow kHb[] = {{660.0, 319.6, 3226.56}, {940.0, 1214.0, 693.44}};
// Absorption of water (cm^-1), Hale & Querry, Appl. Opt. 12:555 (1973).
const double kWater660 = 0.0036, kWater940 = 0.27;

}  // namespace

extern "C" {

void lpm_mf_default(lpm_mf_params* p)
{
    p->r_maternal = 15.0;
    p->t_skin = 0.2;
    p->t_fat = 1.0;
    p->t_uterus = 0.8;
    p->fetal_depth = 2.5;
    p->r_head = 4.5;
    p->t_scalp = 0.1;
    p->t_skull = 0.3;
    p->freq_maternal = 38;
    p->freq_fetal = 38;
    p->freq_torso = 0;
    p->torso_axes[0] = 4.5;
    p->torso_axes[1] = 4.0;
    p->torso_axes[2] = 6.0;
    p->torso_center[0] = 0.0;
    p->torso_center[1] = 0.0;
    p->torso_center[2] = 18.0;
    p->regions = 10;
    p->t_csf = 0.2;
    p->t_grey = 0.4;
    p->freq_fetal = 27;
    p->freq_torso = 20;
}

int64_t lpm_mf_num_triangles(const lpm_mf_params* p)
{
    if (!p || p->freq_maternal < 1 || p->freq_fetal < 1 || p->freq_torso < 0) return -1;
    if (p->regions != 7 && p->regions != 10) return -1;
    const int64_t head = p->regions == 10 ? 5 : 3;  // fetal spheres
    return 4 * 20LL * p->freq_maternal * p->freq_maternal + head * 20LL * p->freq_fetal * p->freq_fetal +
           20LL * p->freq_torso * p->freq_torso;
}

int lpm_mf_build(const lpm_mf_params* p, float* tri, float* nrm, int32_t* med, int64_t ntri)
{
    if (!p || !tri || !nrm || !med || ntri != lpm_mf_num_triangles(p)) return 3;
    const double R = p->r_maternal;
    const double r_cav = R - p->t_skin - p->t_fat - p->t_uterus;
    const V3 cm = {0.0, 0.0, R};
    const V3 cf = {0.0, 0.0, p->fetal_depth + p->r_head};
    // the fetal head must sit inside the amniotic cavity
    double dz = std::fabs(cf[2] - cm[2]);
    const bool ten = p->regions == 10;
    const double inner = p->t_scalp + p->t_skull + (ten ? p->t_csf + p->t_grey : 0.0);
    if (!(r_cav > 0 && p->r_head > inner && dz + p->r_head < r_cav)) return 3;
    if (ten && !(p->t_csf > 0 && p->t_grey > 0)) return 3;
    Geodesic gm = make_geodesic(p->freq_maternal);
    Geodesic gf = (p->freq_fetal == p->freq_maternal) ? gm : make_geodesic(p->freq_fetal);
    int64_t k = 0;
    k = emit_sphere(gm, cm, R, 1, 0, tri, nrm, med, k);
    k = emit_sphere(gm, cm, R - p->t_skin, 2, 1, tri, nrm, med, k);
    k = emit_sphere(gm, cm, R - p->t_skin - p->t_fat, 3, 2, tri, nrm, med, k);
    k = emit_sphere(gm, cm, r_cav, 4, 3, tri, nrm, med, k);
    k = emit_sphere(gf, cf, p->r_head, 5, 4, tri, nrm, med, k);
    k = emit_sphere(gf, cf, p->r_head - p->t_scalp, 6, 5, tri, nrm, med, k);
    k = emit_sphere(gf, cf, p->r_head - p->t_scalp - p->t_skull, 7, 6, tri, nrm, med, k);
    if (ten) {
        k = emit_sphere(gf, cf, p->r_head - p->t_scalp - p->t_skull - p->t_csf, 8, 7, tri, nrm, med, k);
        k = emit_sphere(gf, cf, p->r_head - inner, 9, 8, tri, nrm, med, k);
    }
    if (p->freq_torso > 0) {
        // the torso must lie in the cavity and clear of the head (checked
        // with its largest semi-axis, a bounding sphere)
        const V3 ct = {p->torso_center[0], p->torso_center[1], p->torso_center[2]};
        const double amax = std::max(p->torso_axes[0], std::max(p->torso_axes[1], p->torso_axes[2]));
        const double dm = std::sqrt(dt(sub(ct, cm), sub(ct, cm))), dh = std::sqrt(dt(sub(ct, cf), sub(ct, cf)));
        if (!(p->torso_axes[0] > 0 && p->torso_axes[1] > 0 && p->torso_axes[2] > 0 && dm + amax < r_cav &&
              dh > p->r_head + amax))
            return 3;
        Geodesic gt = make_geodesic(p->freq_torso);
        for (auto& F : gt.faces) {  // the unit sphere scaled per axis: an ellipsoid, still wound outward
            V3 q[3];
            for (int v = 0; v < 3; ++v)
                for (int x = 0; x < 3; ++x) q[v][x] = ct[x] + p->torso_axes[x] * gt.pts[F[v]][x];
            put_tri(tri, nrm, med, k++, q[0], q[1], q[2], ten ? 10 : 5, 4);
        }
    }
    return k == ntri ? 0 : 1;
}

// Tissue model: mua = B [S mua_HbO2 + (1 - S) mua_Hb] + W mua_water with
// whole-blood hemoglobin 150 g/L (64 500 g/mol); mus' = a (lambda/500)^-b,
// mus = mus' / (1 - g).  Blood (B) and water (W) fractions and the (a, b)
// scattering parameters are representative values of the ranges reviewed by
// S. L. Jacques, Phys. Med. Biol. 58:R37 (2013).  All internal tissues share
// n = 1.40 (the reference's fetal fixture likewise matches its soft tissues,
// models/model4/config.xml): under the reference's signed-cosine Fresnel
// (MCLightPropagation.cl:921-1018) a crossing against a face normal between
// unequal indices is almost always a reflection, so mismatched layers would
// confine every photon to the outermost one.
int lpm_mf_optics(double nm, double spo2_m, double spo2_f, float* out, int32_t ntissue)
{
    if (!out || (ntissue != 8 && ntissue != 11) || spo2_m < 0 || spo2_m > 1 || spo2_f < 0 || spo2_f > 1) return 3;
    const HbRow* row = nullptr;
    for (const HbRow& r : kHb)
        if (std::fabs(r.nm - nm) < 1e-9) row = &r;
    if (!row) return 3;
    const double water = (nm < 800) ? kWater660 : kWater940;
    const double c_hb = 150.0 / 64500.0;  // mol/L
    const double ln10 = std::log(10.0);
    const double mu_oxy = ln10 * row->hbo2 * c_hb, mu_deoxy = ln10 * row->hb * c_hb;
    struct T {
        double n, B, W, a, b, g;
        int fetal;
    };
    // the 8-tissue table (regions 7) is rows 0-7; the 11-tissue table
    // (regions 10) is rows 0-6, then 8-11 (the brain row 7 becomes grey
    // matter, 9)
    const T t8[12] = {
        {1.00, 0.00, 0.00, 0.0, 0.0, 0.0, 0},     // ambient
        {1.40, 0.02, 0.65, 46.0, 1.42, 0.90, 0},  // maternal skin
        {1.40, 0.01, 0.15, 20.0, 0.70, 0.90, 0},  // subcutaneous fat
        {1.40, 0.10, 0.80, 13.0, 1.20, 0.90, 0},  // uterine wall
        {1.40, 0.00, 0.99, 0.1, 0.0, 0.90, 0},    // amniotic fluid (mus' 0.1/cm)
        {1.40, 0.03, 0.80, 46.0, 1.42, 0.90, 1},  // fetal scalp
        {1.40, 0.03, 0.30, 22.9, 0.72, 0.90, 1},  // fetal skull
        {1.40, 0.04, 0.85, 24.2, 1.61, 0.90, 1},  // fetal brain (regions 7)
        {1.40, 0.00, 0.99, 0.2, 0.0, 0.90, 1},    // CSF (mus' 0.2/cm)
        {1.40, 0.04, 0.85, 24.2, 1.61, 0.90, 1},  // grey matter
        {1.40, 0.03, 0.75, 40.8, 3.09, 0.90, 1},  // white matter
        {1.40, 0.04, 0.75, 18.9, 1.29, 0.90, 1},  // fetal body (generic soft tissue)
    };
    T t[11];
    for (int i = 0; i < ntissue; ++i) t[i] = (ntissue == 8 || i < 7) ? t8[i] : t8[i + 1];
    for (int i = 0; i < ntissue; ++i) {
        double S = t[i].fetal ? spo2_f : spo2_m;
        double mua = t[i].B * (S * mu_oxy + (1 - S) * mu_deoxy) + t[i].W * water;
        double musp = t[i].a * std::pow(nm / 500.0, -t[i].b);
        double mus = (i == 0) ? 0.0 : musp / (1 - t[i].g);
        out[4 * i] = (float)t[i].n;
        out[4 * i + 1] = (float)(i == 0 ? 0.0 : mua);
        out[4 * i + 2] = (float)mus;
        out[4 * i + 3] = (float)t[i].g;
    }
    return 0;
}

int64_t lpm_slab_num_triangles(int32_t nlayer) { return nlayer < 1 ? -1 : 2 + 2LL * nlayer + 8LL * nlayer; }

int lpm_slab_build(int32_t nl, const double* th, double hx, double hy, float* tri, float* nrm, int32_t* med,
                   int64_t ntri)
{
    if (nl < 1 || !th || !tri || !nrm || !med || ntri != lpm_slab_num_triangles(nl) || !(hx > 0) || !(hy > 0))
        return 3;
    std::vector<double> z(nl + 1, 0.0);
    for (int i = 0; i < nl; ++i) {
        if (!(th[i] > 0)) return 3;
        z[i + 1] = z[i] + th[i];
    }
    int64_t k = 0;
    auto quad = [&](V3 a, V3 b, V3 c, V3 d, int32_t in, int32_t out) {
        put_tri(tri, nrm, med, k++, a, b, c, in, out);
        put_tri(tri, nrm, med, k++, a, c, d, in, out);
    };
    // top z = 0 (normal -z), interfaces (normal -z, inner = deeper layer), bottom (+z)
    for (int i = 0; i < nl; ++i) {
        double zz = z[i];
        quad({-hx, -hy, zz}, {-hx, hy, zz}, {hx, hy, zz}, {hx, -hy, zz}, i + 1, i);
    }
    quad({-hx, -hy, z[nl]}, {hx, -hy, z[nl]}, {hx, hy, z[nl]}, {-hx, hy, z[nl]}, nl, 0);
    for (int i = 0; i < nl; ++i) {
        double za = z[i], zb = z[i + 1];
        quad({hx, -hy, za}, {hx, hy, za}, {hx, hy, zb}, {hx, -hy, zb}, i + 1, 0);
        quad({-hx, -hy, za}, {-hx, -hy, zb}, {-hx, hy, zb}, {-hx, hy, za}, i + 1, 0);
        quad({-hx, hy, za}, {-hx, hy, zb}, {hx, hy, zb}, {hx, hy, za}, i + 1, 0);
        quad({-hx, -hy, za}, {hx, -hy, za}, {hx, -hy, zb}, {-hx, -hy, zb}, i + 1, 0);
    }
    return k == ntri ? 0 : 1;
}

int lpm_box(const double lo[3], const double hi[3], int32_t inner, int32_t outer, float* tri, float* nrm,
            int32_t* med)
{
    if (!lo || !hi || !tri || !nrm || !med) return 3;
    int64_t k = 0;
    auto quad = [&](V3 a, V3 b, V3 c, V3 d) {
        put_tri(tri, nrm, med, k++, a, b, c, inner, outer);
        put_tri(tri, nrm, med, k++, a, c, d, inner, outer);
    };
    double x0 = lo[0], y0 = lo[1], z0 = lo[2], x1 = hi[0], y1 = hi[1], z1 = hi[2];
    quad({x0, y0, z0}, {x0, y1, z0}, {x1, y1, z0}, {x1, y0, z0});  // -z
    quad({x0, y0, z1}, {x1, y0, z1}, {x1, y1, z1}, {x0, y1, z1});  // +z
    quad({x1, y0, z0}, {x1, y1, z0}, {x1, y1, z1}, {x1, y0, z1});  // +x
    quad({x0, y0, z0}, {x0, y0, z1}, {x0, y1, z1}, {x0, y1, z0});  // -x
    quad({x0, y1, z0}, {x0, y1, z1}, {x1, y1, z1}, {x1, y1, z0});  // +y
    quad({x0, y0, z0}, {x1, y0, z0}, {x1, y0, z1}, {x0, y0, z1});  // -y
    return 0;
}

int64_t lpm_merge_coincident(float* tri, float* nrm, int32_t* med, int64_t ntri)
{
    if (!tri || !nrm || !med || ntri < 0) return -1;
    // key: sorted vertex triple (exact float equality)
    using Key = std::array<float, 9>;
    auto key = [&](int64_t t) {
        std::array<std::array<float, 3>, 3> v;
        for (int a = 0; a < 3; ++a)
            for (int x = 0; x < 3; ++x) v[a][x] = tri[9 * t + 3 * a + x];
        std::sort(v.begin(), v.end());
        Key k;
        for (int a = 0; a < 3; ++a)
            for (int x = 0; x < 3; ++x) k[3 * a + x] = v[a][x];
        return k;
    };
    std::map<Key, int64_t> first;
    std::vector<char> drop((size_t)ntri, 0);
    for (int64_t t = 0; t < ntri; ++t) {
        Key k = key(t);
        auto it = first.find(k);
        if (it == first.end()) {
            first.emplace(k, t);
            continue;
        }
        int64_t s = it->second;
        double d = (double)nrm[3 * s] * nrm[3 * t] + (double)nrm[3 * s + 1] * nrm[3 * t + 1] +
                   (double)nrm[3 * s + 2] * nrm[3 * t + 2];
        if (d < 0 && !drop[s]) {  // opposite winding: one interface
            med[2 * s + 1] = med[2 * t];
            drop[t] = 1;
        }
    }
    int64_t w = 0;
    for (int64_t t = 0; t < ntri; ++t) {
        if (drop[t]) continue;
        if (w != t) {
            std::memmove(tri + 9 * w, tri + 9 * t, 9 * sizeof(float));
            std::memmove(nrm + 3 * w, nrm + 3 * t, 3 * sizeof(float));
            std::memmove(med + 2 * w, med + 2 * t, 2 * sizeof(int32_t));
        }
        ++w;
    }
    return w;
}

// Emitter::CalculatePointOfOrigin + CalculateEmitterModel (Emitter.cpp:186-205,
// 299-339), in float as the reference's std::vector<float> arithmetic.
int lpm_emitter(const float p0[3], const float h[3], float distance, float pstart[3], float pb[12])
{
    if (!p0 || !h || !pstart || !pb) return 3;
    for (int x = 0; x < 3; ++x) pstart[x] = -h[x] * distance + p0[x];
    float ez[3] = {h[0], h[1], h[2]}, a[3] = {0, 0, 0};
    if (ez[2] != 0) {
        a[0] = 1;
        a[1] = 1;
        a[2] = -(ez[0] + ez[1]) / ez[2];
    } else if (ez[1] != 0) {
        a[0] = 1;
        a[1] = -ez[0] / ez[1];
        a[2] = 0;
    } else if (ez[0] != 0) {
        a[0] = 0;
        a[1] = 1;
        a[2] = 0;
    }
    auto normalize = [](float* v) {
        float f = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int x = 0; x < 3; ++x) v[x] = v[x] / f;
    };
    auto cross = [](const float* u, const float* w, float* r) {
        r[0] = u[1] * w[2] - u[2] * w[1];
        r[1] = u[2] * w[0] - u[0] * w[2];
        r[2] = u[0] * w[1] - u[1] * w[0];
    };
    normalize(a);
    normalize(ez);
    float ey[3], ex[3];
    cross(ez, a, ey);
    normalize(ey);
    cross(ey, ez, ex);
    normalize(ex);
    for (int x = 0; x < 3; ++x) {
        pb[x] = p0[x];
        pb[3 + x] = ex[x];
        pb[6 + x] = ey[x];
        pb[9 + x] = ez[x];
    }
    return 0;
}

int lpm_cell_normals(const float* tri, int64_t ntri, float* nrm)
{
    if (!tri || !nrm || ntri < 0) return 3;
    for (int64_t t = 0; t < ntri; ++t) {
        V3 a, b, c;
        for (int x = 0; x < 3; ++x) {
            a[x] = tri[9 * t + x];
            b[x] = tri[9 * t + 3 + x];
            c[x] = tri[9 * t + 6 + x];
        }
        V3 n = crs(sub(b, a), sub(c, a));
        double l = std::sqrt(dt(n, n));
        for (int x = 0; x < 3; ++x) nrm[3 * t + x] = l > 0 ? (float)(n[x] / l) : 0.0f;
    }
    return 0;
}

}  // extern "C"
