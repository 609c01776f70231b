// vertex_module.hip -- a loadable gfx950 code object holding the world-space
// vertex stage and the PikoScreen clear, laid out the way the reference's
// generated host code expects of a pipeline module: the per-frame uniforms
// in a module global named constState (fetched with cuModuleGetGlobal and
// filled with cuMemcpyHtoD before the first launch, pikoc/src/Backend/
// PTXBackend.cpp:242-246) and one extern "C" entry per kernel, named
// kernel<N> (PTXBackend.cpp:860-870).  The library's own raster path keeps
// its uniforms in kernel arguments instead; this object is what the runtime
// shim (piko_rt_module_load / get_function / get_global / launch) loads.
//
// Built with hipcc --genco --offload-arch=gfx950 -ffp-contract=off into
// piko_amd/vertex_module.hsaco.
#include "piko_hip.h"
#include "raster_kernels.h"

__device__ piko_constant_state constState;

namespace {

__device__ piko::FrameArgs frame_args(int S) {
  piko::FrameArgs fa{};
  fa.W = constState.screenSizeX;
  fa.H = constState.screenSizeY;
  fa.S = S;
  fa.halfW = constState.halfW;
  fa.halfH = constState.halfH;
  for (int k = 0; k < 16; ++k) fa.vp[k] = constState.viewProjMatrix[k];
  for (int c = 0; c < 3; ++c) {
    fa.light[c] = constState.light[c];
    fa.matcol[c] = constState.matcol[c];
    fa.ambient[c] = constState.ambient[c];
  }
  return fa;
}

}  // namespace

// kernel0: VertexShade for world-space input (vertexShader.pikostage with
// the lighting of main.cpp:310-334): per triangle the three fixed-point
// screen vertices xy[6], depths z[3], lit colours col[3] and the keep flag
// (front facing, inside the guard band, not between sample rows).  One
// thread per triangle; pos and nrm are 9 floats per triangle.
extern "C" __global__ void __launch_bounds__(256) kernel0(const float *pos, const float *nrm, int n, int S,
                                                          int32_t *xy, float *z, uint32_t *col,
                                                          uint8_t *keep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const piko::FrameArgs fa = frame_args(S);
  int p[6];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float *v = pos + 9 * (size_t)i + 3 * k;
    const float *nv = nrm + 9 * (size_t)i + 3 * k;
    float zk;
    piko::transform_vertex(v, fa, p[2 * k], p[2 * k + 1], zk);
    z[3 * (size_t)i + k] = zk;
    col[3 * (size_t)i + k] = piko::light_vertex(nv[0], nv[1], nv[2], fa);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) xy[6 * (size_t)i + k] = p[k];
  keep[i] = piko::vs_keep(p, fa.W, fa.H, S) ? 1 : 0;
}

// kernel1: PikoScreen clear (stage.h:83-85) and depth clear to 1.0
// (main.cpp:402-409) over W*H pixels taken from constState.
extern "C" __global__ void __launch_bounds__(256) kernel1(uint32_t *color, float *depth) {
  const long long npix = (long long)constState.screenSizeX * constState.screenSizeY;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  color[i] = piko::kClearColor;
  depth[i] = __int_as_float(piko::kDepthClearBits);
}
