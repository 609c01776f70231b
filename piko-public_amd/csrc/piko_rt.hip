// piko_rt.hip -- runtime shim for code generated against a HIP target.
//
// The PTX backend's generated host code drives the CUDA driver API directly
// (pikoc/src/Backend/PTXBackend.cpp: cuInit/cuCtxCreate 112-130,
// cuModuleLoad 127, cuMemAlloc/cuMemcpyHtoD 143-202 and 709-759,
// cuModuleGetGlobal 244-247, cuModuleGetFunction 870, cuLaunchKernel 996,
// cuCtxSynchronize 251/278/824, cuMemcpyDtoH 971, cuMemFree/cuCtxDestroy
// 294-320 and 840-854).  A HIP target emits calls to these C functions
// instead; each maps onto one hipModule*/hip* call and returns a piko_status
// rather than aborting.
#include <hip/hip_runtime.h>

#include "piko_hip.h"

namespace {
// A failed HIP call also sets the calling thread's last-error state, which a
// later, unrelated launch check (hipGetLastError after a kernel launch) would
// report as its own failure.  The CUDA driver API has no such state, and the
// status returned here already carries the failure, so it is consumed.
int st(hipError_t e) {
  if (e == hipSuccess) return PIKO_OK;
  (void)hipGetLastError();
  if (e == hipErrorOutOfMemory) return PIKO_ERR_OOM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return PIKO_ERR_NO_DEVICE;
  if (e == hipErrorInvalidValue) return PIKO_ERR_INVALID;
  return PIKO_ERR_HIP;
}
}  // namespace

extern "C" {

int piko_rt_init(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return PIKO_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return PIKO_ERR_INVALID;
  return st(hipSetDevice(device));
}

int piko_rt_module_load(const char *path, piko_module_t *mod) {
  if (!path || !mod) return PIKO_ERR_INVALID;
  hipModule_t m = nullptr;
  const int rc = st(hipModuleLoad(&m, path));
  *mod = rc == PIKO_OK ? (piko_module_t)m : nullptr;
  return rc;
}

int piko_rt_module_unload(piko_module_t mod) {
  if (!mod) return PIKO_ERR_INVALID;
  return st(hipModuleUnload((hipModule_t)mod));
}

int piko_rt_get_function(piko_module_t mod, const char *name, piko_fn_t *fn) {
  if (!mod || !name || !fn) return PIKO_ERR_INVALID;
  hipFunction_t f = nullptr;
  const int rc = st(hipModuleGetFunction(&f, (hipModule_t)mod, name));
  *fn = rc == PIKO_OK ? (piko_fn_t)f : nullptr;
  return rc;
}

int piko_rt_get_global(piko_module_t mod, const char *name, void **dptr, size_t *bytes) {
  if (!mod || !name || !dptr) return PIKO_ERR_INVALID;
  hipDeviceptr_t p = nullptr;
  size_t b = 0;
  const int rc = st(hipModuleGetGlobal(&p, &b, (hipModule_t)mod, name));
  *dptr = rc == PIKO_OK ? (void *)p : nullptr;
  if (bytes) *bytes = b;
  return rc;
}

int piko_rt_malloc(void **dptr, size_t bytes) {
  if (!dptr) return PIKO_ERR_INVALID;
  return st(hipMalloc(dptr, bytes ? bytes : 1));
}

int piko_rt_free(void *dptr) { return dptr ? st(hipFree(dptr)) : PIKO_OK; }

int piko_rt_memcpy_htod(void *dst, const void *src, size_t bytes) {
  if ((!dst || !src) && bytes) return PIKO_ERR_INVALID;
  return bytes ? st(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)) : PIKO_OK;
}

int piko_rt_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  if ((!dst || !src) && bytes) return PIKO_ERR_INVALID;
  return bytes ? st(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)) : PIKO_OK;
}

int piko_rt_memcpy_dtod(void *dst, const void *src, size_t bytes) {
  if ((!dst || !src) && bytes) return PIKO_ERR_INVALID;
  return bytes ? st(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice)) : PIKO_OK;
}

int piko_rt_memset_d32(void *dst, uint32_t value, size_t count) {
  if (!dst && count) return PIKO_ERR_INVALID;
  return count ? st(hipMemsetD32((hipDeviceptr_t)dst, (int)value, count)) : PIKO_OK;
}

int piko_rt_launch(piko_fn_t fn, unsigned gx, unsigned bx, unsigned shmem, piko_stream_t s, void **args) {
  if (!fn || gx == 0 || bx == 0) return PIKO_ERR_INVALID;
  return st(hipModuleLaunchKernel((hipFunction_t)fn, gx, 1, 1, bx, 1, 1, shmem, (hipStream_t)s, args, nullptr));
}

int piko_rt_sync(piko_stream_t s) {
  return s ? st(hipStreamSynchronize((hipStream_t)s)) : st(hipDeviceSynchronize());
}

int piko_rt_stream_create(piko_stream_t *s) {
  if (!s) return PIKO_ERR_INVALID;
  hipStream_t h = nullptr;
  const int rc = st(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
  *s = rc == PIKO_OK ? (piko_stream_t)h : nullptr;
  return rc;
}

int piko_rt_stream_destroy(piko_stream_t s) { return s ? st(hipStreamDestroy((hipStream_t)s)) : PIKO_ERR_INVALID; }

int piko_rt_event_create(piko_event_t *e) {
  if (!e) return PIKO_ERR_INVALID;
  hipEvent_t h = nullptr;
  const int rc = st(hipEventCreate(&h));
  *e = rc == PIKO_OK ? (piko_event_t)h : nullptr;
  return rc;
}

int piko_rt_event_destroy(piko_event_t e) { return e ? st(hipEventDestroy((hipEvent_t)e)) : PIKO_ERR_INVALID; }

int piko_rt_event_record(piko_event_t e, piko_stream_t s) {
  return e ? st(hipEventRecord((hipEvent_t)e, (hipStream_t)s)) : PIKO_ERR_INVALID;
}

int piko_rt_event_elapsed(piko_event_t a, piko_event_t b, float *ms) {
  if (!a || !b || !ms) return PIKO_ERR_INVALID;
  const int rc = st(hipEventSynchronize((hipEvent_t)b));
  return rc != PIKO_OK ? rc : st(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
}

int piko_rt_mem_info(size_t *free_bytes, size_t *total_bytes) {
  if (!free_bytes || !total_bytes) return PIKO_ERR_INVALID;
  return st(hipMemGetInfo(free_bytes, total_bytes));
}

int piko_rt_shutdown(void) { return st(hipDeviceSynchronize()); }

const char *piko_rt_error_string(int status) { return piko_status_string(status); }

}  // extern "C"
