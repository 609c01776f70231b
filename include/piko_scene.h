/*
 * piko_scene.h -- C-ABI of the synthetic scene generators (libpiko_scene.so).
 *
 * The reference loads scenes through its .scene parser and assimp and builds
 * the camera with gluPerspective/gluLookAt (pikoc/samples/rasterPipelineFixPt/
 * main.cpp:268-297, 336-400).  Neither exists here, so workloads are
 * generated.  Every generator is a pure function of its arguments.
 * Return value: 0 on success, -1 on invalid arguments.
 */
#ifndef PIKO_SCENE_H
#define PIKO_SCENE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pretransformed screen-space triangle soup (VTX_PRETRANSFORM input,
 * raster_wtri of rasterTypes.h:8-34): centroid uniform over [0,W)x[0,H),
 * area = area_median*exp(sigma*N(0,1)), equilateral, uniform rotation,
 * counter-clockwise; per-vertex z uniform in [z_lo,z_hi); per-vertex RGB
 * uniform [0,1) packed with the truncating toABGRHost (main.cpp:330-334).
 * pos: [n*9] (x,y,z per vertex), col: [n*3]. */
int piko_scene_tris_uniform(uint64_t seed, int n, int W, int H, float area_median, float sigma,
                            float z_lo, float z_hi, float *pos, uint32_t *col);

/* World-space height field over [-1,1]^2 (x,z), y = amplitude * value noise
 * with `octaves` octaves (lattice frequency base_freq per unit for the first
 * octave, doubling per octave; piko_scene_heightfield uses 4); grid x grid
 * quads, two triangles each, row-major.
 * pos/nrm: [2*grid*grid*9] triangle soup of positions and unit normals. */
long long piko_scene_heightfield_count(int grid);
int piko_scene_heightfield(uint64_t seed, int grid, float amplitude, int octaves, float *pos,
                           float *nrm);
int piko_scene_heightfield_freq(uint64_t seed, int grid, float amplitude, int octaves, float base_freq,
                                float *pos, float *nrm);

/* World-space triangle soup (the uniform-soup workload placed in a cube so a
 * camera can view it): centroids uniform in [-extent, extent]^3, uniform
 * random facing, equilateral with circumradius size_median*exp(sigma*N(0,1)),
 * counter-clockwise about its unit normal (stored per vertex). */
int piko_scene_soup_world(uint64_t seed, int n, float extent, float size_median, float sigma, float *pos,
                          float *nrm);

/* Right-handed look-at times a perspective projection that maps eye-space
 * depth [znear, zfar] to NDC z [0,1]; column-major 16 floats as
 * matmultfloat4 (math.h:365-373) indexes them. */
int piko_camera_view_proj(const float *eye, const float *target, const float *up,
                          float fovy_deg, float aspect, float znear, float zfar, float *m);

/* Per-scene sizes of a batch: triangle count log-uniform in [n_lo, n_hi],
 * pixel count log-uniform in [px_lo_edge^2, px_hi_edge^2], aspect drawn from
 * {1, 4:3, 16:9}; scene s uses seed base_seed + s. */
int piko_scene_batch_spec(uint64_t base_seed, int nscenes, int n_lo, int n_hi, int px_lo_edge,
                          int px_hi_edge, int *n_tris, int *W, int *H);

/* Reyes input: nlat x nlon bicubic Bezier patches tiling the unit sphere
 * (lat-long), every control point of a shared (3*nlat+1) x 3*nlon lattice
 * pushed radially by U[-noise, noise] (longitude periodic), so patches stay
 * connected.  The pole rows lie on rings of angular radius 1e-4 rad rather
 * than on one point, so no patch has a vanishing derivative.  cp:
 * [nlat*nlon*64] floats, 16 control points x (x,y,z,w=1) per patch as
 * piko_patch.CP (reyesPipe/pikoTypes.h:65-73); patch (a,b) at index
 * a*nlon+b; CP[4*r+c] has r along longitude (v) and c along latitude (u). */
int piko_scene_sphere_patches(uint64_t seed, int nlat, int nlon, float noise, float *cp);

#ifdef __cplusplus
}
#endif
#endif
