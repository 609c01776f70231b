/*
 * piko_oracle.h -- CPU restatement of the Piko reference CPU backend for the
 * fixed-point raster pipeline (VertexShade -> AssignBin -> Raster -> DepthTest
 * -> GouraudShade -> PikoScreen).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke
 * entry and the CPU-baseline leg of the benchmark may load it; the product
 * (the HIP library) never links, loads or falls back to it.
 *
 * Every function cites the reference file:line it restates.  The execution
 * model is the reference CPU backend run single-threaded in primitive
 * submission order (pikoc/src/Backend/CPUBackend.cpp:428-571 with one host
 * thread), which is the deterministic reading of that backend.
 *
 * Parameterisation beyond the reference (documented deviations):
 *   S  subpixel bits (reference hard-codes 4: vertexShader.pikostage:46-53;
 *      the benchmark configurations use 8).  Every literal 16 / 0x8 / <<4 of
 *      the reference becomes 1<<S / 1<<(S-1) / <<S.  The RoundDown/RoundUp
 *      masks of rasterMacros.h:48-49 are kept literally (0xff).
 *   T  raster bin edge in pixels (reference: 8, raster.pikostage:14-15).
 *   bins per axis are ceil(W/T) with pixels clipped to the screen (reference
 *      floors, stage.h:273-279, which drops the right/top partial bins).
 *   edge functions are evaluated in int64 (exact); the reference uses int32,
 *      identical wherever int32 does not overflow (always at S=4 for screens
 *      up to 2k; never enough at S=8).
 *   float->fixed conversion clamps to +-2^29 before truncation (the reference
 *      conversion is undefined behaviour outside int range).
 */
#ifndef PIKO_ORACLE_H
#define PIKO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Setup triangle (raster_stri, rasterTypes.h:36-71) in SoA form. */
typedef struct {
  int32_t *xy;    /* [N*6] x0 y0 x1 y1 x2 y2 (fixed point, S fractional bits) */
  float   *z;     /* [N*3] */
  uint32_t *col;  /* [N*3] packed ABGR8 */
  uint8_t *keep;  /* [N]   1 if the vertex stage emitted the triangle */
} oracle_stri;

/* Vertex stage, VTX_PRETRANSFORM + GORAUD variant
 * (vertexShader.pikostage:36-109, rasterPipe.h:6-7).
 * pos: [N*9] screen-space x,y,z per vertex; col: [N*3] ABGR8.  Returns the
 * number of emitted triangles. */
int oracle_vs_pretransformed(const float *pos, const uint32_t *col, int n,
                             int W, int H, int S, const oracle_stri *out);

/* Vertex stage for world-space input (config 2): vtxTransform
 * (rasterMacros.h:14-21, math.h:365-373) followed by per-vertex directional
 * Lambert lighting computeLightingHost + toABGRHost (main.cpp:317-334).
 * pos: [N*9] world xyz, nrm: [N*9] unit normals, vp: 16 floats column-major.
 * light/matcol/ambient: 3 floats each. */
int oracle_vs_world(const float *pos, const float *nrm, int n, int W, int H,
                    int S, const float *vp, const float *light,
                    const float *matcol, const float *ambient,
                    const oracle_stri *out);

/* Raster stage assignBin (raster.pikostage:40-117): conservative
 * trivial-reject bin walk.  Writes CSR bin lists in submission order.
 * bin_start: [nbins+1]; bin_tris: capacity `cap` entries.  Returns the total
 * number of (triangle, bin) pairs, or -1 if it exceeds cap (bin_tris is then
 * only partially written; call with cap=0 to size). */
long long oracle_bin(const oracle_stri *in, int n, int W, int H, int S, int T,
                     int64_t *bin_start, int32_t *bin_tris, long long cap);

/* Raster stage process + PikoScreen store (raster.pikostage:124-321,
 * stage.h:57-61) over bins in bin order, triangles in bin-list order.
 * color/depth: [W*H], initialised here to the PikoScreen clear colour
 * 0xff663313 (stage.h:83-85) and depth 1.0f (main.cpp:402-409). */
void oracle_raster(const oracle_stri *in, int W, int H, int S, int T,
                   const int64_t *bin_start, const int32_t *bin_tris,
                   uint32_t *color, float *depth);

/* Same fragment arithmetic without binning: every emitted triangle in
 * submission order against every pixel of the screen.  Independent check of
 * the bin walk's conservativeness and of bin-size invariance. */
void oracle_raster_bruteforce(const oracle_stri *in, int n, int W, int H, int S,
                              uint32_t *color, float *depth);

/* Whole frame for pretransformed input with a given thread count for the
 * process kernel (bins block-partitioned over threads as CPUBackend.cpp:
 * 534-557 does).  threads<=1 is the deterministic oracle.  Scratch is
 * allocated internally.  Returns number of (tri,bin) pairs, -1 on failure. */
long long oracle_frame_pretransformed(const float *pos, const uint32_t *col,
                                      int n, int W, int H, int S, int T,
                                      int threads, uint32_t *color,
                                      float *depth);

long long oracle_frame_world(const float *pos, const float *nrm, int n, int W,
                             int H, int S, int T, const float *vp,
                             const float *light, const float *matcol,
                             const float *ambient, int threads,
                             uint32_t *color, float *depth);

/* Fragment count and covered-pixel count of a finished frame's inputs
 * (depth-complexity statistics for scene design). */
void oracle_fragment_stats(const oracle_stri *in, int n, int W, int H, int S,
                           long long *fragments, long long *pixels_covered);

/* Helpers exposed for known-answer tests. */
uint32_t oracle_to_abgr(float r, float g, float b);       /* math.h:286-289 */
void oracle_from_abgr(uint32_t c, float *rgb);             /* math.h:323-330 */
int32_t oracle_fixpt(float v, int S);                      /* vertexShader:46 */

/* ---- Reyes (reyes_oracle.c): Split -> Dice -> Shade ------------------- */

/* Micropolygon after dice + shade set-up: screen corners in the order
 * p0 (u,v), p1 (u+du,v), p2 (u+du,v+dv), p3 (u,v+dv) (dice.pikostage:
 * 206-295), flat depth = mean corner z (shade.pikostage:98) and flat colour
 * = clamped mean of the four Phong colours (shade.pikostage:108-138). */
typedef struct {
  float x[4], y[4], z;
  float rgb[3];
} oracle_upoly;

/* Split stage (split.pikostage:240-342) run level by level.  patches: [np*64]
 * control points (16 x xyzw).  Leaves in (level, position) order are written
 * to leaves [cap*64] with the reference's nu/nv when the pointers are non-null.
 * Returns the leaf count (also when it exceeds cap: call with cap=0 to size). */
long long oracle_reyes_split(const float *patches, int np, int W, int H, const float *vp, float threshold,
                             int max_splits, float *leaves, int *leaf_nu, int *leaf_nv, long long cap);

/* Dice stage (dice.pikostage:159-312).  dice=0: the reference's (nu+1)(nv+1)
 * grid from leaf_nu/leaf_nv; dice=N>0: N x N quads covering the leaf.
 * disp: displacement disp*(x^2-y^2) along the unit normal.  Returns the
 * micropolygon count (writes at most cap). */
long long oracle_reyes_dice(const float *leaves, const int *leaf_nu, const int *leaf_nv, long long nleaf, int dice,
                            int W, int H, const float *vp, float disp, oracle_upoly *out, long long cap);

/* Shade stage sampling (shade.pikostage:88-257) in micropolygon order, spp
 * in {1, 4} (pixel centre, or the centres of the 2x2 strata), resolved by a
 * box filter to rgba [W*H*4] FP32; sample_z [W*H*spp] (float bits) if
 * non-null. */
void oracle_reyes_shade(const oracle_upoly *up, long long nup, int W, int H, int spp, float *rgba,
                        int32_t *sample_z);

/* Split + dice + shade.  Returns the micropolygon count. */
long long oracle_reyes_frame(const float *patches, int np, int W, int H, const float *vp, float threshold,
                             int dice, int spp, float disp, float *rgba, int32_t *sample_z);

#ifdef __cplusplus
}
#endif
#endif
