// Bandwidth probes that bound what the fused step can reach on this GPU.
//
// The fused step streams a fixed set of pitched arrays through HBM: each
// wavefront marches the rows of a 64*CPT-column tile, loading the next row of
// every input array before it consumes the current one, and storing one row
// of every output array per row.  k_pattern does exactly that and no stencil
// arithmetic, so its bandwidth is the ceiling of the access pattern itself.
// k_copy, k_read and k_write are the plain streaming forms beside it.
//
// This library is a measurement tool only: bench.py loads it for
// `--pattern-probe`, and the product library does not depend on it.  It
// allocates its own buffers with hipMalloc and frees them before returning.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

constexpr int PROBE_WPB = 4;  // wavefronts per block, as in the fused step

template <int NR, int NW, int CPT>
__global__ __launch_bounds__(256) void k_pattern(const float* __restrict__ in, float* __restrict__ out,
                                                 int pitch, int nz, size_t stride, int tiles_x, int tz,
                                                 int ntiles) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int t = blockIdx.x * PROBE_WPB + wid; t < ntiles; t += gridDim.x * PROBE_WPB) {
    const int c0 = (t % tiles_x) * 64 * CPT + lane * CPT, r0 = (t / tiles_x) * tz;
    float cur[NR][CPT], nxt[NR][CPT];
#pragma unroll
    for (int a = 0; a < NR; ++a)
#pragma unroll
      for (int c = 0; c < CPT; ++c) cur[a][c] = in[a * stride + (size_t)r0 * pitch + c0 + c];
#pragma unroll 2
    for (int r = r0; r < r0 + tz; ++r) {
      const int rn = min(r + 1, nz - 1);  // one row of lookahead, clamped to the grid
#pragma unroll
      for (int a = 0; a < NR; ++a)
#pragma unroll
        for (int c = 0; c < CPT; ++c) nxt[a][c] = in[a * stride + (size_t)rn * pitch + c0 + c];
      float s[CPT];
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        s[c] = 0.f;
#pragma unroll
        for (int a = 0; a < NR; ++a) s[c] += cur[a][c];
      }
#pragma unroll
      for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int c = 0; c < CPT; ++c) out[w * stride + (size_t)r * pitch + c0 + c] = s[c] + (float)w;
#pragma unroll
      for (int a = 0; a < NR; ++a)
#pragma unroll
        for (int c = 0; c < CPT; ++c) cur[a][c] = nxt[a][c];
    }
  }
}

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ in, float4* __restrict__ out, size_t n4) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

// Reads everything; a lane stores only if its sum equals a value the zeroed
// input never produces, so the loads stay and the writes are nil.
__global__ __launch_bounds__(256) void k_read(const float4* __restrict__ in, float* __restrict__ out, size_t n4) {
  float s = 0.f;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
    const float4 v = in[i];
    s += v.x + v.y + v.z + v.w;
  }
  if (s == 1234.5f) out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_write(float4* __restrict__ out, size_t n4) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
    out[i] = make_float4(1.f, 2.f, 3.f, (float)(i & 7));
}

// lds: dynamic LDS reserved per block and never touched; it caps the blocks
// resident per CU (160 KiB of LDS per CU) and so the waves per SIMD, to run
// the pattern at the occupancy of a kernel with many more registers.
template <int NR, int NW, int CPT>
void launch_pattern(hipStream_t st, int grid, int lds, const float* in, float* out, int pitch, int nz,
                    size_t stride, int tiles_x, int tz, int ntiles) {
  hipLaunchKernelGGL((k_pattern<NR, NW, CPT>), dim3(grid), dim3(64 * PROBE_WPB), (size_t)lds, st, in, out, pitch,
                     nz, stride, tiles_x, tz, ntiles);
}

typedef void (*PatternFn)(hipStream_t, int, int, const float*, float*, int, int, size_t, int, int, int);

template <int NR, int NW>
PatternFn pick_cpt(int cpt) {
  switch (cpt) {
    case 1: return launch_pattern<NR, NW, 1>;
    case 2: return launch_pattern<NR, NW, 2>;
    case 4: return launch_pattern<NR, NW, 4>;
    default: return nullptr;
  }
}

PatternFn pick(int nread, int nwrite, int cpt) {
  if (nwrite != 5) return nullptr;
  switch (nread) {
    case 8: return pick_cpt<8, 5>(cpt);
    case 10: return pick_cpt<10, 5>(cpt);
    case 12: return pick_cpt<12, 5>(cpt);
    default: return nullptr;
  }
}

// Times reps back-to-back launches of fn after 3 untimed ones.
template <class F>
int time_launches(hipStream_t st, int reps, double* us, F fn) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -2;
  for (int i = 0; i < 3; ++i) fn();
  (void)hipEventRecord(e0, st);
  for (int i = 0; i < reps; ++i) fn();
  (void)hipEventRecord(e1, st);
  int rc = hipEventSynchronize(e1) == hipSuccess ? 0 : -2;
  float ms = 0.f;
  if (rc == 0 && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = -2;
  if (rc == 0 && hipGetLastError() != hipSuccess) rc = -2;
  *us = 1e3 * (double)ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

}  // namespace

extern "C" {

// The fused step's access pattern over an nx x nz grid of pitched rows:
// nread input arrays (8, 10 or 12) and nwrite output arrays (5), cpt columns
// per lane (1, 2 or 4), tiles of tz rows, `grid` blocks of 4 wavefronts,
// lds_bytes (0 to 65536) of idle LDS per block to cap the occupancy.
// nx must be a multiple of 64*cpt and nz of tz.  Writes the mean launch time
// in microseconds to *us.  Returns 0, -1 for arguments out of range, -2 for a
// HIP error.
int fomo_probe_pattern(int nread, int nwrite, int cpt, int nx, int nz, int tz, int grid, int lds_bytes, int reps,
                       double* us) {
  const PatternFn fn = pick(nread, nwrite, cpt);
  if (!fn || !us || nx <= 0 || nz <= 0 || tz < 2 || reps <= 0 || grid <= 0 || nx % (64 * cpt) || nz % tz ||
      lds_bytes < 0 || lds_bytes > 65536)
    return -1;
  const int pitch = ((nx + 272 + 255) / 256) * 256;  // the fused step's row pitch
  const size_t stride = (size_t)(nz + 36) * pitch + 512;
  float *in = nullptr, *out = nullptr;
  if (hipMalloc(&in, (size_t)nread * stride * sizeof(float)) != hipSuccess) return -2;
  if (hipMalloc(&out, (size_t)nwrite * stride * sizeof(float)) != hipSuccess) {
    (void)hipFree(in);
    return -2;
  }
  int rc = -2;
  if (hipMemset(in, 0, (size_t)nread * stride * sizeof(float)) == hipSuccess &&
      hipMemset(out, 0, (size_t)nwrite * stride * sizeof(float)) == hipSuccess) {
    const int tiles_x = nx / (64 * cpt), ntiles = tiles_x * (nz / tz);
    rc = time_launches(nullptr, reps, us, [&] { fn(nullptr, grid, lds_bytes, in, out, pitch, nz, stride, tiles_x, tz, ntiles); });
  }
  (void)hipFree(in);
  (void)hipFree(out);
  return rc;
}

// Plain streaming over `bytes` (a multiple of 16): mode 0 copies one buffer
// to another, 1 only reads, 2 only writes.  Same return codes as above.
int fomo_probe_stream(int mode, size_t bytes, int grid, int reps, double* us) {
  if (!us || mode < 0 || mode > 2 || bytes == 0 || bytes % 16 || reps <= 0 || grid <= 0) return -1;
  const size_t n4 = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return -2;
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return -2;
  }
  int rc = -2;
  if (hipMemset(a, 0, bytes) == hipSuccess && hipMemset(b, 0, bytes) == hipSuccess) {
    if (mode == 0)
      rc = time_launches(nullptr, reps, us, [&] { hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, nullptr, a, b, n4); });
    else if (mode == 1)  // out holds grid*256 floats: 4 KiB per block, far below bytes here
      rc = (size_t)grid * 256 * sizeof(float) > bytes
               ? -1
               : time_launches(nullptr, reps, us, [&] {
                   hipLaunchKernelGGL(k_read, dim3(grid), dim3(256), 0, nullptr, a, (float*)b, n4);
                 });
    else
      rc = time_launches(nullptr, reps, us, [&] { hipLaunchKernelGGL(k_write, dim3(grid), dim3(256), 0, nullptr, b, n4); });
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

// The runtime's own streaming over `bytes` (a multiple of 16), as independent
// points beside the kernels above: mode 0 hipMemcpyAsync device to device,
// 1 hipMemsetAsync (bytes), 2 hipMemsetD32Async (32-bit words), all on the
// null stream.  Same return codes as above.
int fomo_probe_runtime(int mode, size_t bytes, int reps, double* us) {
  if (!us || mode < 0 || mode > 2 || bytes == 0 || bytes % 16 || reps <= 0) return -1;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return -2;
  if (mode == 0 && hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return -2;
  }
  int rc = -2;
  if (hipMemset(a, 0, bytes) == hipSuccess && (!b || hipMemset(b, 0, bytes) == hipSuccess)) {
    if (mode == 0)
      rc = time_launches(nullptr, reps, us, [&] { (void)hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, nullptr); });
    else if (mode == 1)
      rc = time_launches(nullptr, reps, us, [&] { (void)hipMemsetAsync(a, 1, bytes, nullptr); });
    else
      rc = time_launches(nullptr, reps, us, [&] { (void)hipMemsetD32Async((hipDeviceptr_t)a, 0x3f800000u, bytes / 4, nullptr); });
  }
  (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

}  // extern "C"
