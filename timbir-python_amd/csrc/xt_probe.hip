// Measurement probe: the ceiling of the voxel-line update's write pattern.
//
// The line kernel subtracts A * delta from e with one no-return FP32 buffer
// atomic per footprint entry: per view of a line, three bins, sixteen lanes on
// the sixteen contiguous floats of a packed parity row.  This probe issues
// exactly that pattern (1024 views x 768 bins x 16-float rows, 3 bins per view,
// lines drawn over a 512 x 512 grid, the view walk rotated per line) and
// nothing else, so its rate is the most the line kernel's atomic phase can reach
// on this device: the L2 atomic units, not HBM, set it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/xt_icd.h"

int xt_report_error(int code, const char *fmt, ...);

namespace {

constexpr int kViews = 1024, kBins = 768, kRow = 16, kGrid = 512, kFootBins = 3;

__global__ void __launch_bounds__(256) k_probe_footprint_atomics(const int *__restrict__ lines,
                                                                 int n_lines,
                                                                 const float *__restrict__ cs,
                                                                 const float *__restrict__ sn,
                                                                 float *e) {
  constexpr int G = 256 / kRow;  // view groups of a workgroup
  const int grp = threadIdx.x / kRow, lane = threadIdx.x % kRow;
  const uint32_t bytes = (uint32_t)((size_t)kViews * kBins * kRow * sizeof(float));
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(e, 0, bytes, 0x00020000);
  for (int i = blockIdx.x; i < n_lines; i += gridDim.x) {
    const int line = lines[i];
    const int y = line / kGrid, x = line % kGrid;
    const int rot = (int)(((uint32_t)line * 2654435761u) % kViews);
    for (int q = grp; q < kViews; q += G) {
      int v = q + rot;
      if (v >= kViews) v -= kViews;
      // first bin of the voxel's footprint in view v (parallel beam, 1.5-bin voxels)
      const float r = ((float)x - 255.5f) * 1.5f * cs[v] - ((float)y - 255.5f) * 1.5f * sn[v] +
                      (float)(kBins / 2);
      int fb = (int)r - 1;
      fb = fb < 0 ? 0 : (fb > kBins - kFootBins ? kBins - kFootBins : fb);
      const uint32_t off = (uint32_t)(((v * kBins + fb) * kRow + lane) * sizeof(float));
#pragma unroll
      for (int b = 0; b < kFootBins; b++)
        __builtin_amdgcn_raw_ptr_buffer_atomic_fadd_f32(1e-9f, rs, off, b * kRow * (int)sizeof(float),
                                                        0);
    }
  }
}

}  // namespace

extern "C" int xt_probe_footprint_atomics(int32_t n_lines, double *lane_ops_per_s) {
  if (n_lines < 1 || n_lines > kGrid * kGrid || !lane_ops_per_s)
    return xt_report_error(XT_E_ARG, "probe: 1..%d lines and an output pointer", kGrid * kGrid);
  const size_t ns = (size_t)kViews * kBins * kRow;
  float *e = nullptr, *cs = nullptr, *sn = nullptr;
  int *lines = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  int rc = XT_E_HIP;
  std::vector<float> c(kViews), s(kViews);
  for (int v = 0; v < kViews; v++) {
    c[v] = cosf((float)v * 3.14159265f / kViews);
    s[v] = sinf((float)v * 3.14159265f / kViews);
  }
  std::vector<int> L(n_lines);
  uint32_t h = 761521u;
  for (int i = 0; i < n_lines; i++) {
    h = h * 1664525u + 1013904223u;
    L[i] = (int)(h % (uint32_t)(kGrid * kGrid));
  }
  float ms = 0.f;
  const int reps = 3;
  if (hipMalloc((void **)&e, ns * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&cs, kViews * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&sn, kViews * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&lines, n_lines * sizeof(int)) != hipSuccess)
    goto done;
  if (hipMemset(e, 0, ns * sizeof(float)) != hipSuccess ||
      hipMemcpy(cs, c.data(), kViews * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(sn, s.data(), kViews * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(lines, L.data(), n_lines * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
    goto done;
  // one warm launch, then the timed ones (1024 workgroups, as the line kernel)
  hipLaunchKernelGGL(k_probe_footprint_atomics, dim3(1024), dim3(256), 0, 0, lines, n_lines, cs, sn,
                     e);
  if (hipEventRecord(a, 0) != hipSuccess) goto done;
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_probe_footprint_atomics, dim3(1024), dim3(256), 0, 0, lines, n_lines, cs,
                       sn, e);
  if (hipGetLastError() != hipSuccess || hipEventRecord(b, 0) != hipSuccess ||
      hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess)
    goto done;
  *lane_ops_per_s = (double)n_lines * kViews * kFootBins * kRow / ((double)ms / reps * 1e-3);
  rc = XT_OK;
done:
  if (rc != XT_OK) (void)xt_report_error(XT_E_HIP, "footprint atomic probe failed");
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (e) (void)hipFree(e);
  if (cs) (void)hipFree(cs);
  if (sn) (void)hipFree(sn);
  if (lines) (void)hipFree(lines);
  return rc;
}
