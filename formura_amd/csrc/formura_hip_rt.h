// Runtime support for generated formura HIP libraries: device buffers, the
// stream, timing, and the error convention of the reference's generated C
// (message on stderr, then exit(1); Templates.hs:243).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>

namespace fmr_rt {
[[noreturn]] inline void fatal(const char *msg) {
  fprintf(stderr, "formura_hip: %s\n", msg);
  fflush(stderr);
  exit(1);
}
[[noreturn]] inline void fatal_hip(const char *what, hipError_t e, const char *file, int line) {
  fprintf(stderr, "formura_hip: %s failed at %s:%d: %s\n", what, file, line, hipGetErrorString(e));
  fflush(stderr);
  exit(1);
}
}  // namespace fmr_rt

#define FMR_CHECK(x)                                                        \
  do {                                                                      \
    hipError_t fmr_e_ = (x);                                                \
    if (fmr_e_ != hipSuccess) fmr_rt::fatal_hip(#x, fmr_e_, __FILE__, __LINE__); \
  } while (0)

// Periodic index wrap for |i - [0,N)| < N (the halo of the reference is never
// wider than the grid), and a general modulo for tiny grids.
template <int N>
__device__ __forceinline__ int fmr_wrap(int i) {
  return i < 0 ? i + N : (i >= N ? i - N : i);
}
template <int N>
__device__ __forceinline__ int fmr_mod(int i) {
  int r = i % N;
  return r < 0 ? r + N : r;
}

namespace fmr_rt {

struct Runtime {
  int ready = 0;
  int resident = 0;  // 0: Formura_Forward syncs formura_data both ways
  size_t bytes = 0;
  char *cur = nullptr;  // current state image (layout of the grid struct)
  char *nxt = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  long long launches = 0;  // kernel launches in the last Formura_Forward_Steps
  void swap() {
    char *t = cur;
    cur = nxt;
    nxt = t;
  }
};

inline void *device_alloc(size_t b) {
  void *p = nullptr;
  if (b == 0) b = 1;
  FMR_CHECK(hipMalloc(&p, b));
  return p;
}

inline void device_free(void *p) {
  if (p) FMR_CHECK(hipFree(p));
}

inline void init(Runtime &rt, size_t bytes) {
  if (rt.ready) return;  // a second Formura_Init re-uploads only
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) fatal("no HIP device is visible");
  rt.bytes = bytes;
  rt.cur = (char *)device_alloc(bytes);
  rt.nxt = (char *)device_alloc(bytes);
  FMR_CHECK(hipStreamCreateWithFlags(&rt.stream, hipStreamNonBlocking));
  FMR_CHECK(hipEventCreate(&rt.ev0));
  FMR_CHECK(hipEventCreate(&rt.ev1));
  rt.ready = 1;
}

inline void upload(Runtime &rt, const void *host) {
  FMR_CHECK(hipMemcpyAsync(rt.cur, host, rt.bytes, hipMemcpyHostToDevice, rt.stream));
  FMR_CHECK(hipStreamSynchronize(rt.stream));
}

inline void download(Runtime &rt, void *host) {
  FMR_CHECK(hipMemcpyAsync(host, rt.cur, rt.bytes, hipMemcpyDeviceToHost, rt.stream));
  FMR_CHECK(hipStreamSynchronize(rt.stream));
}

inline void timer_start(Runtime &rt) { FMR_CHECK(hipEventRecord(rt.ev0, rt.stream)); }

inline void timer_stop(Runtime &rt) {
  FMR_CHECK(hipEventRecord(rt.ev1, rt.stream));
  FMR_CHECK(hipEventSynchronize(rt.ev1));
  float ms = 0.f;
  FMR_CHECK(hipEventElapsedTime(&ms, rt.ev0, rt.ev1));
  rt.last_ms = ms;
}

inline void shutdown(Runtime &rt) {
  if (!rt.ready) return;
  FMR_CHECK(hipStreamSynchronize(rt.stream));
  device_free(rt.cur);
  device_free(rt.nxt);
  FMR_CHECK(hipEventDestroy(rt.ev0));
  FMR_CHECK(hipEventDestroy(rt.ev1));
  FMR_CHECK(hipStreamDestroy(rt.stream));
  rt = Runtime();
}

}  // namespace fmr_rt
