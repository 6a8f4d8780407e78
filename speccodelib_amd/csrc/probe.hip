// Streaming HBM probes: the rate plain kernels reach on this device, the
// practical ceiling the solver's passes are measured against.  Each thread
// moves 16-byte vectors (dwordx4), four per loop trip, over a grid-stride
// loop of 64 workgroups of 1024 threads per CU, with nontemporal loads;
// buffers are far larger than the 256 MB last-level cache.  This shape was
// the fastest of a sweep on MI355X (2-64 workgroups per CU, 256-1024
// threads, unroll 1-8, plain or nontemporal loads and stores, grid-stride or
// contiguous chunks): read 7.1 TB/s, write 5.7 TB/s, copy 5.3-5.5 TB/s.
#include <algorithm>

#include "scl_core.hpp"

namespace {

constexpr int PROBE_BS = 1024;
constexpr int PROBE_WG_PER_CU = 64;
constexpr int PROBE_UNROLL = 4;

// KIND 0: read (a sum that is stored only if it equals a value it never
// takes, so the loads stay); 1: write; 2: copy.
template <int KIND>
__global__ __launch_bounds__(PROBE_BS) void k_stream(const float4* __restrict__ src, float4* __restrict__ dst,
                                                     int64_t n, float* __restrict__ sink, float never) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4* s4 = reinterpret_cast<const v4*>(src);
  const int64_t stride = (int64_t)gridDim.x * PROBE_BS;
  const int64_t t0 = (int64_t)blockIdx.x * PROBE_BS + threadIdx.x;
  float acc = 0.f;
  int64_t i = t0;
  for (; i + (PROBE_UNROLL - 1) * stride < n; i += PROBE_UNROLL * stride) {
    float4 v[PROBE_UNROLL];
    if constexpr (KIND != 1) {
#pragma unroll
      for (int u = 0; u < PROBE_UNROLL; ++u) {
        const v4 t = __builtin_nontemporal_load(s4 + i + u * stride);
        v[u] = make_float4(t.x, t.y, t.z, t.w);
      }
    }
#pragma unroll
    for (int u = 0; u < PROBE_UNROLL; ++u) {
      if constexpr (KIND == 0) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
      else if constexpr (KIND == 1) dst[i + u * stride] = make_float4(never, never, never, (float)u);
      else dst[i + u * stride] = v[u];
    }
  }
  for (; i < n; i += stride) {
    if constexpr (KIND == 0) {
      const float4 v = src[i];
      acc += (v.x + v.y) + (v.z + v.w);
    } else if constexpr (KIND == 1) {
      dst[i] = make_float4(never, never, never, 0.f);
    } else {
      dst[i] = src[i];
    }
  }
  if (KIND == 0 && acc == never) sink[t0] = acc;
}

}  // namespace

extern "C" int32_t scl_stream_probe(int32_t kind, int64_t bytes, int32_t reps, double* ms_per_rep) {
  using namespace scl;
  if (kind < 0 || kind > 2 || bytes < (1 << 20) || reps < 1 || !ms_per_rep)
    return set_error("stream probe: kind 0..2, bytes >= 1 MiB, reps >= 1, non-null output"), SCL_EINVAL;
  const int64_t n = bytes / 16;
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  void *a = nullptr, *b = nullptr, *s = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int32_t rc = SCL_OK;
  auto fail = [&](const char* what, hipError_t e) {
    (void)hipGetLastError();
    set_error("stream probe: %s: %s", what, hipGetErrorString(e));
    rc = SCL_EHIP;
  };
  hipError_t e;
  if ((e = hipMalloc(&a, (size_t)n * 16)) != hipSuccess || (e = hipMalloc(&b, (size_t)n * 16)) != hipSuccess ||
      (e = hipMalloc(&s, (size_t)ncu * PROBE_WG_PER_CU * PROBE_BS * sizeof(float))) != hipSuccess) {
    fail("allocation", e);
  } else if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess ||
             (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
    fail("stream or events", e);
  } else if ((e = hipMemsetAsync(a, 0, (size_t)n * 16, st)) != hipSuccess) {
    fail("fill", e);
  } else {
    const dim3 grid((unsigned)(ncu * PROBE_WG_PER_CU)), blk(PROBE_BS);
    auto launch = [&]() {
      const float4* src = (const float4*)a;
      float4* dst = (float4*)b;
      if (kind == 0) hipLaunchKernelGGL((k_stream<0>), grid, blk, 0, st, src, dst, n, (float*)s, -1.f);
      else if (kind == 1) hipLaunchKernelGGL((k_stream<1>), grid, blk, 0, st, src, dst, n, (float*)s, -1.f);
      else hipLaunchKernelGGL((k_stream<2>), grid, blk, 0, st, src, dst, n, (float*)s, -1.f);
    };
    launch();  // warm-up
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, st);
    if ((e = hipGetLastError()) != hipSuccess) {
      fail("launch", e);
    } else if ((e = hipEventSynchronize(e1)) != hipSuccess) {
      fail("run", e);
    } else {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      *ms_per_rep = ms / reps;
    }
  }
  if (st) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (s) (void)hipFree(s);
  return rc;
}
