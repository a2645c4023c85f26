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

// a / b for a divisor b fixed at generation time, correctly rounded (the value
// of IEEE division) without the device's division sequence.  The generator
// passes y = RN(1/b), computed on the host, and b > 0 with 2^-100 <= b <= 2^100.
//   q0 = RN(a*y) is within 2 ulp of a/b;  q1 = q0 - (b*q0 - a)*y (fused) is
//   faithful;  q2 = q1 - (b*q1 - a)*y (fused) is RN(a/b) by Markstein's theorem
//   (IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of Floating-Point
//   Arithmetic, 2nd ed., Thm. 4.10): b*q1 - a is exact for a faithful q1, and
//   y has relative error below 2^-53.
// The theorem needs no underflow or overflow: with 2^-900 <= |a| <= 2^900 the
// quotient and the remainders are normal (fmr_divc_in_range).  Other a (tiny,
// huge, Inf, NaN) take the division.  The remainder is formed as b*q - a and
// subtracted, so a zero dividend keeps its sign.
__host__ __device__ __forceinline__ bool fmr_divc_in_range(double a) {
  const double aa = fabs(a);
  return (aa >= 0x1p-900 && aa <= 0x1p+900) || aa == 0.0;
}

__host__ __device__ __forceinline__ double fmr_divc_fast(double a, double b, double y) {
  double q = a * y;
  double r = fma(b, q, -a);
  q = fma(-r, y, q);
  r = fma(b, q, -a);
  return fma(-r, y, q);
}

// One correction, for divisors whose host-rounded reciprocal is close enough:
// if |b*y - 1| <= 2^-54 then q0 = RN(a*y) is already faithful (|a*y - a/b| <=
// |a/b| 2^-54 < ulp(a/b)/2, so q0 is one of the two doubles around a/b, also
// where a*y and a/b lie in different binades), and one Markstein step
// q1 = q0 - (b*q0 - a)*y gives RN(a/b).  The generator checks the condition in
// exact rational arithmetic (b = RN(1e-4) with y = 1e4: |b*y - 1| = 0.043 * 2^-53).
//
// The remainder is formed at scale 2^120: r = (b 2^120) q0 - a 2^120, then
// q1 = q0 - r (y 2^-120) (the scaled constants fold at compile time; scaling by
// a power of two is exact here).  The fused operation sees the same real
// number q0 - (b q0 - a) y as unscaled, so wherever the unscaled remainder is
// exact the result is the same double; the scaled remainder is a multiple of
// 2^(e_b + 120 - 52) * ulp(q0) >= 2^-1020 and below 2^53 of those, so it is
// exact for every q0 -- subnormal dividends and quotients included, where the
// unscaled one would underflow.  For a subnormal quotient Q = a/b: q0 is within
// 3/4 ulp of Q, and q0 - r y = Q + e (Q - q0) with |e| = |b y - 1| <= 2^-54,
// while Q lies at least 1/(2M) > 2^-54 ulp from every rounding midpoint (b =
// M 2^k, M odd and above every subnormal's integer significand), so the fused
// rounding gives RN(Q) as in the normal range.  Valid for |a| <= 2^900 (a 2^120
// and a y stay finite): fmr_divc1_in_range.
__host__ __device__ __forceinline__ bool fmr_divc1_in_range(double a) { return fabs(a) <= 0x1p+900; }

__host__ __device__ __forceinline__ double fmr_divc1_fast(double a, double b, double y) {
  const double q = a * y;
  const double r = fma(b * 0x1p120, q, -(a * 0x1p120));
  return fma(-r, y * 0x1p-120, q);
}

__host__ __device__ __forceinline__ double fmr_divc(double a, double b, double y) {
  if (__builtin_expect(!fmr_divc_in_range(a), 0)) return a / b;
  return fmr_divc_fast(a, b, y);
}

__host__ __device__ __forceinline__ double fmr_divc1(double a, double b, double y) {
  if (__builtin_expect(!fmr_divc1_in_range(a), 0)) return a / b;
  return fmr_divc1_fast(a, b, y);
}

// The division again for a wave that holds a dividend outside the fused
// sequence's range.  A dividend with lo <= |a| < 2^-900 is scaled by 2^600
// (exact: the result is normal, below 2^-300), divided with the fused sequence
// (its range holds it), and the quotient scaled back by 2^-600 (exact: the
// generator passes lo = 2^ceil(log2 b) * 2^-1021, so |a/b| >= 2^-1021 and the
// quotient stays normal, where scaling by a power of two commutes with
// rounding); every other out-of-range dividend (smaller, huge, Inf, NaN) takes
// the division.  Lanes already in range get the fused sequence unscaled.
__host__ __device__ __forceinline__ double fmr_divc_redo(double a, double b, double y, double lo) {
  const double aa = fabs(a);
  const bool tiny = aa < 0x1p-900 && aa >= lo;
  if (__builtin_expect(!(tiny || fmr_divc_in_range(a)), 0)) return a / b;
  return fmr_divc_fast(a * (tiny ? 0x1p600 : 1.0), b, y) * (tiny ? 0x1p-600 : 1.0);
}

// (fmr_divc1_fast covers tiny dividends itself: only huge, Inf and NaN ones
// are redone, with the division)
__host__ __device__ __forceinline__ double fmr_divc1_redo(double a, double b, double y, double lo) {
  (void)lo;
  return fmr_divc1(a, b, y);
}

// Speculative forms for point functions: the fused sequences without the
// branch; the lanes whose dividend is not fmr_divc_in_range are or-ed into the
// wave mask `bad` (one ballot per comparison, each a single compare into a lane
// mask: !in_range = (!(|a| >= 2^-900) || !(|a| <= 2^900)) && a != 0).  A caller
// evaluates the whole point with these and, if its wave's mask came out
// nonzero, evaluates it again with fmr_divc / fmr_divc1 (a branch on a
// wave-uniform value).  Without a branch per division the compiler can issue
// all of a point's loads before its arithmetic.
__device__ __forceinline__ unsigned long long fmr_divc_out_of_range(double a) {
  const double aa = fabs(a);
  return (__builtin_amdgcn_ballot_w64(!(aa >= 0x1p-900)) | __builtin_amdgcn_ballot_w64(!(aa <= 0x1p+900))) &
         __builtin_amdgcn_ballot_w64(aa != 0.0);
}

// The lanes whose dividend is outside fmr_divc1_in_range: one comparison
// (false for NaN, so NaN lanes are counted).
__device__ __forceinline__ unsigned long long fmr_divc1_out_of_range(double a) {
  return __builtin_amdgcn_ballot_w64(!(fabs(a) <= 0x1p+900));
}

__device__ __forceinline__ double fmr_divcs(double a, double b, double y, unsigned long long &bad) {
  bad |= fmr_divc_out_of_range(a);
  return fmr_divc_fast(a, b, y);
}

__device__ __forceinline__ double fmr_divc1s(double a, double b, double y, unsigned long long &bad) {
  bad |= fmr_divc1_out_of_range(a);
  return fmr_divc1_fast(a, b, y);
}

// Division by a computed value (not a literal) in the opt-in contract form:
// the hardware reciprocal approximation refined by two Newton steps, the
// quotient, and one correction of it.  For normal operands and a normal
// quotient the result is within a few units in the last place of IEEE
// division, without the device's scaling and special-case steps; zero,
// infinite, subnormal and NaN operands are not handled as IEEE division
// handles them.  Only the contract form, checked against the exact path to a
// stated tolerance, uses it.
__device__ __forceinline__ double fmr_div_contract(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  r = fma(fma(-b, r, 1.0), r, r);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}

// Per-level range test of the tile kernels: every value a level writes (and
// every value of the loaded tile) is 0 or a finite x with 2^-848 <= |x| <= xmax.
// Such values are integer multiples of 2^-900, so every dividend the generator
// proved to be an integer combination of them (and of literals) is 0 or in
// fmr_divc's range, and the next level divides with the fused sequences alone.
__device__ __forceinline__ bool fmr_level_ok(double x, double xmax) {
  const double ax = fabs(x);
  return (ax >= 0x1p-848 && ax <= xmax) || ax == 0.0;
}

// Positivity checks of the mixed-precision form: every FP64 square-root
// argument (e.g. gamma*p/rho of a wave speed) must be > 0; a non-positive or
// NaN argument is counted (vector atomic, taken only on a violation).
__device__ unsigned long long fmr_violations = 0;
#define FMR_POSITIVE(x) \
  do { \
    if (!((x) > 0)) atomicAdd(&fmr_violations, 1ull); \
  } while (0)

namespace fmr_rt {

// FP64 totals of a field, in a fixed order (deterministic): TOTALS_BLOCKS
// workgroups sum fixed grid-stride subsets (shuffles within a wave, then the
// waves in order); one wave adds the partial sums in order.
constexpr int TOTALS_BLOCKS = 1024;

template <typename T>
__global__ void __launch_bounds__(256) k_totals_partial(const T *__restrict__ f, long long n, double *__restrict__ part) {
  double s = 0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) s += (double)f[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (w[0] + w[1]) + (w[2] + w[3]);
}

__global__ void __launch_bounds__(64) k_totals_final(const double *__restrict__ part, int nb, double *__restrict__ out) {
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += 64) s += part[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (threadIdx.x == 0) *out = s;
}

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
  double *scratch = nullptr;  // totals: partial sums and results
  size_t scratch_n = 0;
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

// Sum of n values of f into *dev_out through TOTALS_BLOCKS partial sums in
// `part` (every slot written before it is read), enqueued on `s`.
template <typename T>
inline void field_total_with(double *part, const T *f, long long n, double *dev_out, hipStream_t s) {
  hipLaunchKernelGGL(k_totals_partial<T>, dim3(TOTALS_BLOCKS), dim3(256), 0, s, f, n, part);
  hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(64), 0, s, part, TOTALS_BLOCKS, dev_out);
  FMR_CHECK(hipGetLastError());
}

template <typename T>
inline void field_total(Runtime &rt, const T *f, long long n, double *dev_out) {
  if (rt.scratch_n < (size_t)TOTALS_BLOCKS) fatal("field_total: scratch not allocated");
  field_total_with(rt.scratch, f, n, dev_out, rt.stream);
}

// Partial sums of the allocation-free totals (Formura_HIP_Totals_Static): a
// file-scope device array, one copy per GPU at its own address, held until the
// process ends (8 KiB).  Every call on a GPU shares that GPU's copy, so calls
// that overlap on one GPU (two streams, two host threads) race.
static __device__ __align__(256) double scratch[TOTALS_BLOCKS];
constexpr int MAX_DEVICES = 64;

// Device address of `scratch` on the current device, looked up once per device
// with that device current (the array's name in host code is a host-side
// stand-in, not a device address).
inline double *static_scratch() {
  static double *ptr[MAX_DEVICES] = {};
  int dev = 0;
  FMR_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) fatal("static_scratch: device index out of range");
  if (!ptr[dev]) {
    void *p = nullptr;
    FMR_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptr[dev] = (double *)p;
  }
  return ptr[dev];
}

inline double *lazy_scratch(Runtime &rt, size_t n) {
  if (rt.scratch_n < n) {
    FMR_CHECK(hipStreamSynchronize(rt.stream));
    device_free(rt.scratch);
    rt.scratch = (double *)device_alloc(n * sizeof(double));
    rt.scratch_n = n;
  }
  return rt.scratch;
}

inline long long violations(Runtime &rt) {
  if (rt.ready) FMR_CHECK(hipStreamSynchronize(rt.stream));
  unsigned long long v = 0;
  FMR_CHECK(hipMemcpyFromSymbol(&v, HIP_SYMBOL(fmr_violations), sizeof(v), 0, hipMemcpyDeviceToHost));
  return (long long)v;
}

inline void shutdown(Runtime &rt) {
  if (!rt.ready) return;
  FMR_CHECK(hipStreamSynchronize(rt.stream));
  device_free(rt.cur);
  device_free(rt.nxt);
  device_free(rt.scratch);
  FMR_CHECK(hipEventDestroy(rt.ev0));
  FMR_CHECK(hipEventDestroy(rt.ev1));
  FMR_CHECK(hipStreamDestroy(rt.stream));
  rt = Runtime();
}

}  // namespace fmr_rt
