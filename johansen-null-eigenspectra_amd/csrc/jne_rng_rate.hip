// jne_rng_rate.hip -- the normal generator's own rate on this GPU: the
// ceiling of the moment passes, which draw every normal they use.
//
// Each thread draws `calls` Philox4x32-10 calls (4 normals each, normals4 of
// jne_rng.h: the same generator and operation sequence as the walk kernels)
// of its own seed, three independent calls per iteration as the moment
// kernels interleave them, and stores one sum.  Nothing else is timed: the
// stream kernel behind jne_walk_normals_device writes 4 bytes per normal to
// HBM and is bound by those writes, so it is not this ceiling.
//
// Usage: jne_rng_rate [--blocks B] [--calls C] [--reps R]
// Prints one JSON line: normals per second (median over the timed launches,
// HIP events around each launch), the launch shape and the sum of all
// outputs (a check that the work was done, and reproducible for fixed args).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "jne_rng.h"

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void rng_rate_kernel(uint32_t calls, float* __restrict__ out) {
  const uint32_t tid = blockIdx.x * kBlock + threadIdx.x;
  float acc = 0.0f;
  for (uint32_t q = 0; q < calls; q += 3) {
    float v[3][4];
#pragma unroll
    for (int j = 0; j < 3; ++j) jne::normals4(q + j, tid + 1u, 0u, v[j]);
#pragma unroll
    for (int j = 0; j < 3; ++j) acc += (v[j][0] + v[j][1]) + (v[j][2] + v[j][3]);
  }
  out[tid] = acc;
}

int fail(const char* what, hipError_t e) {
  std::fprintf(stderr, "jne_rng_rate: %s: %s\n", what, hipGetErrorString(e));
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  long blocks = 8192, calls = 1200, reps = 5;
  for (int i = 1; i + 1 < argc; i += 2) {
    const long v = std::strtol(argv[i + 1], nullptr, 10);
    if (!std::strcmp(argv[i], "--blocks")) blocks = v;
    else if (!std::strcmp(argv[i], "--calls")) calls = v;
    else if (!std::strcmp(argv[i], "--reps")) reps = v;
    else {
      std::fprintf(stderr, "usage: jne_rng_rate [--blocks B] [--calls C] [--reps R]\n");
      return 2;
    }
  }
  if (blocks < 1 || blocks > 65536 || calls < 3 || calls > 300000 || calls % 3 != 0 || reps < 1 || reps > 100) {
    std::fprintf(stderr, "jne_rng_rate: need 1 <= blocks <= 65536, 3 <= calls <= 300000 (a multiple of 3), "
                         "1 <= reps <= 100\n");
    return 2;
  }
  const size_t threads = (size_t)blocks * kBlock;
  float* d = nullptr;
  hipError_t e = hipMalloc(&d, threads * sizeof(float));
  if (e != hipSuccess) return fail("hipMalloc", e);
  hipEvent_t a, b;
  if ((e = hipEventCreate(&a)) != hipSuccess || (e = hipEventCreate(&b)) != hipSuccess)
    return fail("hipEventCreate", e);
  std::vector<float> ms;
  for (long r = 0; r <= reps; ++r) {  // launch 0 is a warm-up
    if ((e = hipEventRecord(a, nullptr)) != hipSuccess) return fail("hipEventRecord", e);
    hipLaunchKernelGGL(rng_rate_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, nullptr, (uint32_t)calls, d);
    if ((e = hipGetLastError()) != hipSuccess) return fail("launch", e);
    if ((e = hipEventRecord(b, nullptr)) != hipSuccess) return fail("hipEventRecord", e);
    if ((e = hipEventSynchronize(b)) != hipSuccess) return fail("hipEventSynchronize", e);
    float t = 0.0f;
    if ((e = hipEventElapsedTime(&t, a, b)) != hipSuccess) return fail("hipEventElapsedTime", e);
    if (r > 0) ms.push_back(t);
  }
  std::vector<float> h(threads);
  if ((e = hipMemcpy(h.data(), d, threads * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)
    return fail("hipMemcpy", e);
  double sum = 0.0;
  for (float x : h) sum += x;
  if ((e = hipFree(d)) != hipSuccess || (e = hipEventDestroy(a)) != hipSuccess ||
      (e = hipEventDestroy(b)) != hipSuccess)
    return fail("release", e);
  std::sort(ms.begin(), ms.end());
  const double med = ms[ms.size() / 2];
  const double normals = 4.0 * (double)threads * (double)calls;
  std::printf("{\"normals_per_s\": %.6e, \"median_launch_ms\": %.4f, \"threads\": %zu, \"calls_per_thread\": %ld, "
              "\"normals_per_launch\": %.6e, \"timed_launches\": %ld, \"sum_of_outputs\": %.6f}\n",
              normals / (med * 1e-3), med, threads, calls, normals, reps, sum);
  return 0;
}
