/*
 * lightprop_host.h -- C ABI of the host layer around the transport library:
 * scene ingestion without VTK or Boost, the emitter model, the tally images
 * and the reference's post-processing and text outputs.
 *
 * The C++ classes behind it (lightpropagation_amd/host/lp_host.hpp) keep the
 * reference's class and method names -- ImportExport, Emitter,
 * CalculatePropagation, Analysis, Logs -- so the reference's main flow
 * (src/main/main.cpp:101-320) maps onto them call for call.  This header is
 * the same layer for FFI callers (ctypes, cgo, JNI) and for the tests.
 *
 * Conventions as in lightprop.h: plain types, caller-owned buffers, int
 * status (0 = ok), no exceptions across the boundary; lph_last_error() gives
 * the text of the last failure of this thread.
 */
#ifndef LIGHTPROP_HOST_H
#define LIGHTPROP_HOST_H

#include <stdint.h>

#include "lightprop.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lph_scene lph_scene;

/* Text of the last failed lph_* call on this thread (never NULL); the
 * reference reports errors on std::cout and through its return codes. */
const char *lph_last_error(void);

/* Reads a configuration XML (ReadXML::ParseFile, src/IO/ReadXML.cpp:93-168
 * with the element handlers of :598-622, 663-871, 950-1010), the OBJ model it
 * names (ImportExport::ImportObjModel + GenMediumIDsFromObjModel,
 * src/IO/ImportExport.cpp:341-433, 952-1053) and its emitter file
 * (ImportEmitterModel, :381-414), then runs the emitter model
 * (Emitter::GetEmittermodel / GetBasisVectors, src/Model/Emitter.cpp:186-339).
 * A model or emitter path that does not exist is looked up by its file name
 * in the configuration's directory. */
int lph_scene_load(const char *config_path, lph_scene **out);
void lph_scene_free(lph_scene *s);

/* Sizes: faces, tissues, start directions. */
int lph_scene_sizes(const lph_scene *s, int64_t *ntri, int32_t *ntissue, int32_t *nstart);
/* tri[9T] (vertex coordinates), nrm[3T] (unit cell normals from the face
 * winding), med[2T] (inner, outer tissue ids). */
int lph_scene_mesh(const lph_scene *s, float *tri, float *nrm, int32_t *med);
/* nmuasg[4R] = (n, mua, mus, g) per tissue id. */
int lph_scene_materials(const lph_scene *s, float *nmuasg, int32_t *fetal_id);
/* Start point, start directions vstart[3A] with counts vocc[A], and
 * (POI, eX, eY, eZ) = poi_basis[12]. */
int lph_scene_source(const lph_scene *s, float pstart[3], float *vstart, int32_t *vocc, float poi_basis[12]);
/* model extent, detection grid (x, y, z, r, phi, rho), PixelPerUnit factor,
 * roulette chance, weight threshold, number of photons (the values
 * main.cpp:101-320 reads from the configuration through ReadXML). */
int lph_scene_params(const lph_scene *s, float model_dims[3], float det_grid[6], float *factor, float *chance,
                     float *w_th, int64_t *nphotons);
/* Sets mesh, materials, source, grid and (chance, w_th, factor, seed,
 * rng_mode, max_steps) on a transport context: the setter sequence of
 * main.cpp:174-222 in one call. */
int lph_scene_apply(const lph_scene *s, lp_context *ctx, uint64_t seed, int32_t rng_mode, int64_t max_steps);

/* Emitter model (Emitter::CalculateEmitterModel, Emitter.cpp:186-339):
 * start point p0 - heading * distance, one start direction per emitter
 * angle (the heading rotated by the angle, Emitter::RotateVector :389-473),
 * photon counts per direction from the relative radiant intensities, and the
 * basis (eX, eY, eZ) of the emitter frame. */
int lph_emitter_model(const float p0[3], const float heading[3], float distance, const float *angles_deg,
                      const float *rri, int32_t nangles, int64_t nphotons, float pstart[3], float *vstart,
                      int32_t *vocc, float ex[3], float ey[3], float ez[3]);

/* ---- tally images and post-processing (src/IO/Logs.cpp, src/Analysis) ---- */

/* A tally as the reference's float image: extent {x0, x1, y0, y1} and
 * ncomp values per pixel, x fastest.  which: LP_TALLY_SFT (3 components,
 * x in [-dimx/2, dimx/2], y in [-dimy/2, dimy/2]), LP_TALLY_RTD (3, r in
 * [0, dimR - 1], alpha in [0, dimAlpha - 1]), LP_TALLY_ARZ and
 * LP_TALLY_PHIRZ (1, r in [-dimR, dimR], z in [-dimz, 0]); component -1 =
 * all, else that component only.  Values are the fixed-point sums as float
 * (GetSFTImage / GetRTdImage / GetArzImage / GetPhirzImage,
 * CalculatePropagation.cpp:897-1056).  With out == NULL only extent and
 * ncomp are returned. */
int lph_image(lp_context *ctx, int32_t which, int32_t component, int32_t extent[4], int32_t *ncomp, float *out);

/* MCML scale factors (Logs::ScalePerRadius1D, ScalePerAngle1D, Scale2D,
 * Logs.cpp:507-516, 651-672); det_grid as above. */
float lph_scale_per_radius_1d(const float det_grid[6], int64_t nphotons, int32_t ir);
float lph_scale_per_angle_1d(const float det_grid[6], int64_t nphotons, int32_t ia);
float lph_scale_2d(const float det_grid[6], int64_t nphotons, int32_t ir, int32_t ia);

/* Profiles of a one-component image as the reference's text files write them
 * (two columns per row): reflectance per radius (CalcReflectancePerRadius,
 * Logs.cpp:520-561), per angle (CalcReflectancePerAngle, :674-710) and
 * fluence per depth (FluenceZ, :1075-1110).  rows receives the row count;
 * col0/col1 receive the columns (NULL to query the count). */
int lph_profile(int32_t kind, const float *image, const int32_t extent[4], const float det_grid[6],
                int64_t nphotons, int32_t *rows, float *col0, float *col1);
#define LPH_PER_RADIUS 0
#define LPH_PER_ANGLE 1
#define LPH_PER_Z 2

/* Analysis::CalculateSurfaceSums (Analysis.cpp:144-156) over a 3-component
 * SFT image: float sums of R, F, T in the reference's pixel order. */
int lph_surface_sums(const float *sft_image, const int32_t extent[4], float sums[3]);

/* The reference's main flow (src/main/main.cpp:101-320) without pictures:
 * import the configuration, run nphotons photons (<= 0: the configuration's
 * count) on device `device`, print the Analysis totals, and when out_dir is
 * not NULL write the Logs text files there (SaveSurfaceArr, SaveRecordRd,
 * the _perRadius / _perAngle files of the surface, transmission and fetal
 * images, and FluenceZ).  sums receives (R, F, T) as fractions of nphotons. */
int lph_run_config(const char *config_path, int64_t nphotons, uint64_t seed, int32_t device, const char *out_dir,
                   double sums[3]);

#ifdef __cplusplus
}
#endif
#endif /* LIGHTPROP_HOST_H */
