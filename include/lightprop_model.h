/*
 * lightprop_model.h -- scene construction for liblp.so (model layer).
 *
 * The reference builds its scenes from OBJ meshes through VTK
 * (src/IO/ImportExport.cpp:341-433, 952-1053) and its source model in
 * src/Model/Emitter.cpp:186-339.  This header exposes the same model
 * interfaces without VTK, plus generators for the synthetic benchmark scenes
 * that the reference does not ship.  Two-call pattern: query the size, then
 * fill caller-owned buffers.  Mesh convention (the reference's): one surface
 * per interface, med = (inner, outer), normals point from inner to outer.
 */
#ifndef LIGHTPROP_MODEL_H
#define LIGHTPROP_MODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Maternal-fetal abdomen (synthetic benchmark scene, no reference
 * counterpart): concentric maternal spheres (skin, fat, uterine
 * wall, then the amniotic cavity) tangent to z = 0 at the origin, centred at
 * (0, 0, r_maternal), over a fetal head of concentric spheres whose top lies
 * at depth `fetal_depth` on the z axis, and optionally a fetal torso.
 * regions = 10 (default): tissue ids 0 ambient, 1 maternal skin, 2 fat,
 * 3 uterine wall, 4 amniotic fluid, 5 fetal scalp, 6 fetal skull, 7 CSF,
 * 8 grey matter, 9 white matter, 10 fetal body (the torso).
 * regions = 7: 0-6 as above, 7 fetal brain (no CSF / grey / white shells),
 * the torso (if any) of the scalp's tissue 5.
 * Each sphere is a geodesic icosphere of frequency `freq` (20 freq^2
 * faces), watertight. */
typedef struct lpm_mf_params {
    double r_maternal;   /* outer abdomen radius [cm] (15)                      */
    double t_skin, t_fat, t_uterus; /* maternal layer thicknesses (0.2, 1.0, 0.8) */
    double fetal_depth;  /* depth of the fetal scalp top below the skin (2.5)    */
    double r_head;       /* fetal head radius (4.5)                             */
    double t_scalp, t_skull;        /* fetal layer thicknesses (0.1, 0.3)      */
    int32_t freq_maternal;          /* icosphere frequency of maternal spheres */
    int32_t freq_fetal;             /* icosphere frequency of fetal spheres    */
    /* optional fetal torso: an ellipsoid of fetal soft tissue (tissue 5, the
     * scalp's optics) in the amniotic fluid; freq_torso 0 = none */
    int32_t freq_torso;
    double torso_axes[3];           /* semi-axes along x, y, z (4.5, 4.0, 6.0)  */
    double torso_center[3];         /* centre (0, 0, 18.0): below the head      */
    int32_t regions;                /* 10 (default) or 7, see above             */
    double t_csf, t_grey;           /* CSF and grey matter thicknesses (0.2, 0.4) */
} lpm_mf_params;

/* Fills *p with the defaults above: 10 regions, maternal spheres of
 * frequency 38, the five fetal head spheres of frequency 27, a torso of
 * frequency 20 (196 420 faces). */
void lpm_mf_default(lpm_mf_params *p);
int64_t lpm_mf_num_triangles(const lpm_mf_params *p);
int lpm_mf_build(const lpm_mf_params *p, float *tri, float *nrm, int32_t *med, int64_t ntri);

/* Optical properties (n, mua, mus, g) of the maternal-fetal tissues at
 * 660 or 940 nm (the two rows of the haemoglobin table; other wavelengths
 * return 3) for maternal and fetal arterial oxygen saturations; ntissue 11
 * (regions 10) or 8 (regions 7). */
int lpm_mf_optics(double wavelength_nm, double spo2_maternal, double spo2_fetal, float *nmuasg, int32_t ntissue);

/* Planar layered slab in the reference convention: x, y in [-hx, hx] x
 * [-hy, hy], layer i between z_i and z_i + thickness[i] (z_0 = 0), one
 * surface per interface: 2 + 2 nlayer + 8 nlayer faces.  Tissue i + 1 is
 * layer i; 0 is the ambient. */
int64_t lpm_slab_num_triangles(int32_t nlayer);
int lpm_slab_build(int32_t nlayer, const double *thickness, double hx, double hy, float *tri, float *nrm,
                   int32_t *med, int64_t ntri);

/* Closed boxes sharing faces (e.g. three stacked 12-face boxes) to the
 * reference convention: a pair of coincident faces with opposite winding is
 * replaced by the lower-index face, its outer id set to the inner id of the
 * other.  Returns the new face count (<= ntri); arrays are compacted in place. */
int64_t lpm_merge_coincident(float *tri, float *nrm, int32_t *med, int64_t ntri);

/* Box of 12 faces, outward normals, (inner, outer) = (inner_id, outer_id). */
int lpm_box(const double lo[3], const double hi[3], int32_t inner_id, int32_t outer_id, float *tri, float *nrm,
            int32_t *med);

/* Emitter model (Emitter.cpp:186-339): start point p0 - heading * distance,
 * start direction = heading, and the detection basis eX, eY, eZ derived from
 * the heading.  poi_basis = (p0, eX, eY, eZ). */
int lpm_emitter(const float p0[3], const float heading[3], float distance, float pstart[3], float poi_basis[12]);

/* Cell normals from face winding: normalize(cross(p1 - p0, p2 - p0)) in
 * double, stored as float (the role of vtkPolyDataNormals in
 * ImportExport.cpp:372-379). */
int lpm_cell_normals(const float *tri, int64_t ntri, float *nrm);

#ifdef __cplusplus
}
#endif
#endif
