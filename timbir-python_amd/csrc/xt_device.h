// xt_device.h -- gfx950 kernels of the ICD path.
//
// Layout in HBM (one GPU holds the whole slab):
//   e, w, y     float [N_p][N_r][N_t]     error sinogram, weights, measurements
//                                         (t fastest: the slices of a voxel column
//                                         are contiguous, one 128 B line per
//                                         (view, bin) at N_t = 32)
//   d           float [N_r][N_t]          ring offsets
//   x           float [N_time][N_y][N_x][N_z]   object, z fastest (a column is one
//                                         contiguous run; wave lanes map to z)
//   Hr          float [N_p][65]           detector response table
//   mag_blk     float [N_time][nb][N_y][N_x]  per-z-block |update| of each line
//   magsum      float [N_time][N_y][N_x]  the same summed over blocks
//
// Reference arithmetic is reproduced where it is integer-like (bin indices of a
// system-matrix column, the detector table, the NHICD filter): those functions
// disable FMA contraction and follow the reference's float/double steps.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define XT_MAXA 6          // max detector bins of one A column (2*delta_xy/delta_r + 2)
#define XT_MAXA_WIDE 14    // the same for voxels up to 6 detector bins wide (coarse stages)
#define XT_NBINS 64        // DETECTOR_RESPONSE_BINS
#define XT_HR_STRIDE 65
#define XT_FP_BINS 64      // detector bins per forward-projection workgroup
#define XT_PI 3.14159265358979323846
#define XT_PI_2 1.57079632679489661923132169163975144
#define XT_PI_4 0.785398163397448309615660845819875721

struct XtGeom {
  int N_p, N_r, N_t, N_x, N_y, N_z, N_time, zo;
  int nb;          // z blocks of the magnitude map
  int max_views;   // max views of a frame
  float delta_r, delta_t, R0, RMax, OffsetR;
  float x0, y0, delta_xy;
  // prior (F/XT_Init.c:199-256,391-394)
  float MRF_P, C_S, C_T, Sigma_S_Q, Sigma_S_Q_P, Sigma_T_Q, Sigma_T_Q_P;
  float filt[27];  // Spatial_Filter, [dz][dy][dx] flattened
  float tfilt;     // Time_Filter[0]
  float alpha, T, delta;  // over-relaxation, generalized-Huber T and delta
  int positivity;
  // sinogram layout in HBM (e, w, y): pm = 0: [N_p][N_r][N_t]; pm = 1: split by
  // z-block parity, [2][N_p][N_r][H] with H = N_t/2, detector row t in plane
  // ((t / bs) & 1) at position (t / (2 bs)) * bs + t % bs (bs slices per block)
  int pm, bs, H;
  int wide;  // A columns of more than XT_MAXA bins (voxels wider than 2 detector bins)
  int pc;    // phase contrast: footprints are the detector Laplacian of the absorption ones
  float lap[9];  // Sinogram.Lap_Kernel, [m + 1][n + 1] for bin offset m, row offset n
};

// Offset of sinogram entry (view p, bin r, row t) in the HBM layout.
__device__ __forceinline__ size_t xt_sidx(const XtGeom &g, int p, int r, int t) {
  if (!g.pm) return ((size_t)p * g.N_r + r) * g.N_t + t;
  const int par = (t / g.bs) & 1, tt = (t / (2 * g.bs)) * g.bs + t % g.bs;
  return (((size_t)par * g.N_p + p) * g.N_r + r) * g.H + tt;
}

// ---------------------------------------------------------------------------
// System-matrix column for (row, col, view): the exact float steps of
// F/XT_AMatrix.c:50-110.  Writes the first candidate bin and MA values
// (0 for bins that are not part of the column); returns the entry count.
template <int MA>
__device__ __forceinline__ int xt_amatrix_column(const XtGeom &g, const float *__restrict__ H,
                                                 float cosv, float sinv, int row, int col,
                                                 int *first, float (&vals)[MA]) {
#pragma clang fp contract(off)
  float x = (float)((double)g.x0 + ((double)(float)col + 0.5) * (double)g.delta_xy);
  float y = (float)((double)g.y0 + ((double)(float)row + 0.5) * (double)g.delta_xy);
  float r = x * cosv - y * sinv;
  float rmin = r - g.delta_xy;
  float rmax = r + g.delta_xy;
#pragma unroll
  for (int b = 0; b < MA; b++) vals[b] = 0.f;
  *first = 0;
  if (rmax < g.R0 || rmin > g.RMax) return 0;
  int imin = (int)floorf((rmin - g.R0) / g.delta_r);
  int imax = (int)floorf((rmax - g.R0) / g.delta_r);
  if (imax >= g.N_r) imax = g.N_r - 1;
  if (imin < 0) imin = 0;
  *first = imin;
  int count = 0;
#pragma unroll
  for (int b = 0; b < MA; b++) {
    int j = imin + b;
    if (j <= imax) {
      float Rc = (float)((double)g.R0 + ((double)(float)j + 0.5) * (double)g.delta_r);
      float dr = fabsf(r - Rc);
      int k = (int)floorf(dr / g.OffsetR);
      if (k >= 0 && k < XT_NBINS) {
        float w1 = dr - (float)k * g.OffsetR;
        float w2 = g.OffsetR - w1;
        float f1 = (w2 * H[k] + w1 * H[k + 1]) / g.OffsetR;
        if (f1 > 0) {
          vals[b] = f1;
          count++;
        }
      }
    }
  }
  return count;
}

// ---------------------------------------------------------------------------
// Detector response table, one workgroup per view: the trapezoidal voxel
// profile (F/XT_Profile.c:46-103) convolved with the unit-area Hamming beam
// (F/XT_Profile.c:107-123), sampled at i*OffsetR (F/XT_Profile.c:128-173).
__global__ void __launch_bounds__(256) k_detector_response(XtGeom g, const float *__restrict__ views,
                                                           float beam_width, float *__restrict__ Hr) {
#pragma clang fp contract(off)
  __shared__ float prof[1536];
  __shared__ float beam[512];
  __shared__ float beam_sum;
  const int k = blockIdx.x, tid = threadIdx.x;
  float angle = views[k];
  while ((double)angle > XT_PI_2) angle = (float)((double)angle - XT_PI_2);
  while (angle < 0) angle = (float)((double)angle + XT_PI_2);
  float MaxVal;
  if ((double)angle <= XT_PI_4)
    MaxVal = (float)((double)g.delta_xy / cos((double)angle));
  else
    MaxVal = (float)((double)g.delta_xy / cos(XT_PI_2 - (double)angle));
  float temp = (float)cos(XT_PI_4);
  float dist1 = (float)((double)temp * cos(XT_PI_4 - (double)angle));
  float dist2 = (float)((double)temp * fabs(cos(XT_PI_4 + (double)angle)));
  float LeftCorner = 1 - dist1, LeftNear = 1 - dist2, RightNear = 1 + dist2, RightCorner = 1 + dist1;
  for (int j = tid; j < 1536; j += blockDim.x) {
    float t = (float)(2.0 * j / 1536);
    float v;
    if (t <= LeftCorner || t >= RightCorner)
      v = 0;
    else if (t > RightNear)
      v = MaxVal * (RightCorner - t) / (RightCorner - RightNear);
    else if (t >= LeftNear)
      v = MaxVal;
    else
      v = MaxVal * (t - LeftCorner) / (LeftNear - LeftCorner);
    prof[j] = v;
  }
  for (int i = tid; i < 512; i += blockDim.x)
    beam[i] = (float)(0.54 - 0.46 * cos((2 * XT_PI / 512) * i));
  __syncthreads();
  if (tid == 0) {
    float s = 0;
    for (int i = 0; i < 512; i++) s = s + beam[i];
    beam_sum = s;
  }
  __syncthreads();
  for (int i = tid; i < 512; i += blockDim.x) beam[i] /= beam_sum;
  __syncthreads();
  if (tid < XT_NBINS) {
    float r0 = -(beam_width) / 2;
    float StepSize = beam_width / 512;
    float TempConst = 1536 / (2 * g.delta_xy);
    float ProfileCenterR = tid * g.OffsetR;
    float rmin = ProfileCenterR - g.delta_xy;
    float sum = 0;
    for (int p = 0; p < 512; p++) {
      float r = r0 + p * StepSize;
      int idx = (int)floorf((r - rmin) * TempConst);
      if (idx >= 0 && idx <= 1535) sum += prof[idx] * beam[p];
    }
    Hr[(size_t)k * XT_HR_STRIDE + tid] = sum;
  } else if (tid == XT_NBINS) {
    Hr[(size_t)k * XT_HR_STRIDE + XT_NBINS] = 0.f;
  }
}

// ---------------------------------------------------------------------------
// Forward projection (F/XT_ForwardProject.c:8-25 summed over all voxels):
// one workgroup per (view, chunk of XT_FP_BINS bins) with FP64 accumulators
// [bin][slice] in LDS.  Its view groups walk the rows (or columns, whichever
// axis is more perpendicular to the ray) and enumerate the voxels whose
// column can reach the chunk; each voxel's exact A column is built once, and
// its entries times the line's slices are added to the accumulators.
// mode 0: out64 = Ax.   mode 1: e = y - Ax - d (float).
// The A-column table, when the state keeps one (k_build_atable): a candidate
// voxel's column is then read (first non-zero bin and four values) instead of
// built; vpos[v] is view v's position in its frame's list.
struct XtFpTable {
  const float4 *vals;
  const int *first;
  const int *vpos;
};

template <int ZL, int MA, bool TAB = false>
__global__ void __launch_bounds__(256) k_forward_project(XtGeom g, const float *__restrict__ x,
                                                         const float *__restrict__ Hr,
                                                         const float *__restrict__ cosv,
                                                         const float *__restrict__ sinv,
                                                         const int *__restrict__ view_frame,
                                                         int mode, double *__restrict__ out64,
                                                         const float *__restrict__ y,
                                                         const float *__restrict__ d,
                                                         float *__restrict__ e,
                                                         XtFpTable tab = {nullptr, nullptr, nullptr}) {
  constexpr int G = 256 / ZL;
  // accumulators [bin][slice] plus one spare row that takes the entries of a
  // column falling outside the chunk (added and never read: no branches)
  __shared__ double acc[(XT_FP_BINS + 1) * ZL];
  __shared__ float cval[G][ZL][MA];
  __shared__ int cfirst[G][ZL];
  const int v = blockIdx.y, j0 = blockIdx.x * XT_FP_BINS;
  const int tid = threadIdx.x, zl = tid % ZL, grp = tid / ZL;
  for (int q = tid; q < (XT_FP_BINS + 1) * ZL; q += 256) acc[q] = 0.0;
  __syncthreads();
  const int f = view_frame[v];
  if (f >= 0) {
    const float c = cosv[v], s = sinv[v];
    const float *H = Hr + (size_t)v * XT_HR_STRIDE;
    const float *xf = x + (size_t)f * g.N_y * g.N_x * g.N_z;
    // voxels whose column can reach bins j0..j0+XT_FP_BINS-1: projected centre within
    // delta_xy (+ half a bin, + a margin for float rounding) of the chunk's bin centres
    const int j1 = min(g.N_r, j0 + XT_FP_BINS) - 1;
    const double W = (double)g.delta_xy + 0.5 * (double)g.delta_r + 1e-3 * (double)g.delta_r;
    const double lo = (double)g.R0 + ((double)j0 + 0.5) * (double)g.delta_r - W;
    const double hi = (double)g.R0 + ((double)j1 + 0.5) * (double)g.delta_r + W;
    const double cd = c, sd = s, dxy = g.delta_xy;
    const bool by_row = fabs(cd) >= fabs(sd);
    const int n_outer = by_row ? g.N_y : g.N_x;
    // each view group walks its own rows (or columns); its lanes build the
    // columns of up to ZL candidate voxels at once, then add their products with
    // the line's slices into the block's FP64 accumulators
    for (int o = grp; o < n_outer; o += G) {
      int k0, k1;
      if (by_row) {
        const double yc = (double)g.y0 + (o + 0.5) * dxy;
        double xa = (lo + yc * sd) / cd, xb = (hi + yc * sd) / cd;
        if (xa > xb) { const double t = xa; xa = xb; xb = t; }
        k0 = (int)floor((xa - (double)g.x0) / dxy - 0.5);
        k1 = (int)ceil((xb - (double)g.x0) / dxy - 0.5);
        if (k0 < 0) k0 = 0;
        if (k1 > g.N_x - 1) k1 = g.N_x - 1;
      } else {
        const double xc = (double)g.x0 + (o + 0.5) * dxy;
        double ya = (xc * cd - hi) / sd, yb = (xc * cd - lo) / sd;
        if (ya > yb) { const double t = ya; ya = yb; yb = t; }
        k0 = (int)floor((ya - (double)g.y0) / dxy - 0.5);
        k1 = (int)ceil((yb - (double)g.y0) / dxy - 0.5);
        if (k0 < 0) k0 = 0;
        if (k1 > g.N_y - 1) k1 = g.N_y - 1;
      }
      for (int kb = k0; kb <= k1; kb += ZL) {
        {
          const int k = kb + zl;
          int first = 0;
          float vals[MA];
#pragma unroll
          for (int b = 0; b < MA; b++) vals[b] = 0.f;
          if (k <= k1) {
            const int row = by_row ? o : k, col = by_row ? k : o;
            if constexpr (TAB) {
              static_assert(MA == 4, "table columns hold four bins");
              const size_t ti = ((size_t)f * g.N_y * g.N_x + (size_t)row * g.N_x + col) *
                                    (size_t)g.max_views + (size_t)tab.vpos[v];
              const float4 t4 = tab.vals[ti];
              first = tab.first[ti];
              vals[0] = t4.x;
              vals[1] = t4.y;
              vals[2] = t4.z;
              vals[3] = t4.w;
            } else {
              xt_amatrix_column(g, H, c, s, row, col, &first, vals);
            }
          }
          cfirst[grp][zl] = first;
#pragma unroll
          for (int b = 0; b < MA; b++) cval[grp][zl][b] = vals[b];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const int nk = min(ZL, k1 - kb + 1);
        if (zl < g.N_z) {
          // this lane's slice of the batch's voxels (consecutive along the row or
          // column o), eight loads in flight ahead of their sums
          const int stride = by_row ? g.N_z : g.N_x * g.N_z;
          const float *xb = xf + (by_row ? (o * g.N_x + kb) : (kb * g.N_x + o)) * g.N_z + zl;
          for (int m0 = 0; m0 < nk; m0 += 8) {
            float xv[8];
#pragma unroll
            for (int u = 0; u < 8; u++) xv[u] = xb[min(m0 + u, nk - 1) * stride];  // in range
#pragma unroll
            for (int u = 0; u < 8; u++) {
              const int m = m0 + u;
              if (m < nk) {
                const double xd = (double)xv[u];
                const int first = cfirst[grp][m] - j0;
#pragma unroll
                for (int b = 0; b < MA; b++) {
                  const int jj = first + b;
                  const int slot = (jj >= 0 && jj < XT_FP_BINS) ? jj : XT_FP_BINS;
                  atomicAdd(&acc[slot * ZL + zl], xd * (double)cval[grp][m][b]);
                }
              }
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      }
    }
  }
  __syncthreads();
  // slice z spans detector rows z * zo .. z * zo + zo - 1, each with the slice's
  // value (F/XT_ForwardProject.c:8-25)
  for (int q = tid; q < XT_FP_BINS * g.N_t; q += 256) {
    const int jj = q / g.N_t, t = q - jj * g.N_t, j = j0 + jj, z = t / g.zo;
    if (j >= g.N_r) continue;
    const double a = acc[jj * ZL + z];
    if (mode == 0) {
      out64[((size_t)v * g.N_r + j) * g.N_t + t] = a;
    } else {
      const size_t idx = xt_sidx(g, v, j, t);
      e[idx] = (float)((double)y[idx] - a - (double)d[(size_t)j * g.N_t + t]);
    }
  }
}

// ---------------------------------------------------------------------------
// Filtered backprojection (a warm start for ICD; the reference has none).
// Rows p = y - d of each (view, detector row) are cleaned of isolated outliers
// (zingers) by a 3-point median along the bins and filtered with the Ram-Lak
// ramp of sample spacing delta_r: q_j = (1/delta_r) (p_j/4 - sum_{n odd}
// (p_{j-n} + p_{j+n}) / (n pi)^2); q is written [N_p][N_r][N_t].
__global__ void __launch_bounds__(256) k_fbp_filter(XtGeom g, const float *__restrict__ y,
                                                    const float *__restrict__ d,
                                                    float *__restrict__ q) {
  constexpr int TB = 4;  // detector rows per workgroup
  extern __shared__ float fbuf[];
  float *raw = fbuf, *med = fbuf + TB * g.N_r, *coef = med + TB * g.N_r;  // coef[(n-1)/2]
  const int v = blockIdx.y, t0 = blockIdx.x * TB, tid = threadIdx.x, NR = g.N_r;
  for (int i = tid; i < (NR + 1) / 2; i += blockDim.x) {
    const double n = 2.0 * i + 1.0;
    coef[i] = (float)(1.0 / (n * n * XT_PI * XT_PI));
  }
  for (int i = tid; i < TB * NR; i += blockDim.x) {
    const int k = i / NR, r = i - k * NR, t = t0 + k;
    raw[i] = (t < g.N_t) ? y[xt_sidx(g, v, r, t)] - d[(size_t)r * g.N_t + t] : 0.f;
  }
  __syncthreads();
  for (int i = tid; i < TB * NR; i += blockDim.x) {
    const int r = i % NR;
    const float b = raw[i];
    if (r == 0 || r == NR - 1) {
      med[i] = b;
    } else {
      const float a = raw[i - 1], c = raw[i + 1];
      med[i] = fmaxf(fminf(a, b), fminf(fmaxf(a, b), c));
    }
  }
  __syncthreads();
  for (int i = tid; i < TB * NR; i += blockDim.x) {
    const int k = i % TB, j = i / TB, t = t0 + k;
    if (t >= g.N_t) continue;
    const float *m = med + k * NR;
    float acc = 0.25f * m[j];
    for (int n = 1; n < NR; n += 2) {
      float pr = 0.f;
      if (j - n >= 0) pr += m[j - n];
      if (j + n < NR) pr += m[j + n];
      acc -= pr * coef[n >> 1];
    }
    q[((size_t)v * NR + j) * g.N_t + t] = acc / g.delta_r;
  }
}

// Ring-offset guess for the FBP start: per detector pixel the mean over views
// of y without outliers (values beyond 4 standard deviations of the plain mean:
// zingers), minus its 15-bin running mean along the bins (the view-averaged
// projection of an object is smooth in r, per-bin offsets are not).
__global__ void k_offset_guess_mean(XtGeom g, const float *__restrict__ y, float *__restrict__ m) {
  const int rt = blockIdx.x * blockDim.x + threadIdx.x;
  if (rt >= g.N_r * g.N_t) return;
  const int r = rt / g.N_t, t = rt - r * g.N_t;
  double s1 = 0.0, s2 = 0.0;
  for (int p = 0; p < g.N_p; p++) {
    const double v = y[xt_sidx(g, p, r, t)];
    s1 += v;
    s2 += v * v;
  }
  const double mu = s1 / g.N_p, sd = sqrt(fmax(s2 / g.N_p - mu * mu, 0.0));
  double acc = 0.0;
  int n = 0;
  for (int p = 0; p < g.N_p; p++) {
    const double v = y[xt_sidx(g, p, r, t)];
    if (fabs(v - mu) <= 4.0 * sd) {
      acc += v;
      n++;
    }
  }
  m[rt] = (float)(n ? acc / n : mu);
}

__global__ void __launch_bounds__(1024) k_offset_guess_apply(XtGeom g, const float *__restrict__ m,
                                                             int zero_mean, float *__restrict__ d) {
  __shared__ double part[1024];
  const int n = g.N_r * g.N_t, tid = threadIdx.x;
  double acc = 0.0;
  for (int rt = tid; rt < n; rt += blockDim.x) {
    const int r = rt / g.N_t, t = rt - r * g.N_t;
    double trend = 0.0;
    for (int q = r - 7; q <= r + 7; q++) {
      const int qc = q < 0 ? 0 : (q >= g.N_r ? g.N_r - 1 : q);
      trend += m[qc * g.N_t + t];
    }
    const float v = (float)(m[rt] - trend / 15.0);
    d[rt] = v;
    acc += v;
  }
  part[tid] = acc;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (tid < h) part[tid] += part[tid + h];
    __syncthreads();
  }
  if (zero_mean) {
    const float mean = (float)(part[0] / n);
    for (int rt = tid; rt < n; rt += blockDim.x) d[rt] -= mean;
  }
}

// x(line, z) = sum over the frame's views of wv[v] * q_v(u), with u = (r - R0) /
// delta_r - 1/2 the bin coordinate of the voxel centre's projection r (the
// A-matrix geometry, F/XT_AMatrix.c:59-62) and linear interpolation; wv are the
// views' angular weights (sum pi over a frame).
template <int ZL>
__global__ void __launch_bounds__(256) k_fbp_backproject(XtGeom g, const float *__restrict__ q,
                                                         const float *__restrict__ wv,
                                                         const float *__restrict__ cosv,
                                                         const float *__restrict__ sinv,
                                                         const int *__restrict__ views,
                                                         const int *__restrict__ proj_num,
                                                         const uint8_t *__restrict__ mask,
                                                         float *__restrict__ x) {
  constexpr int G = 256 / ZL;
  const int tid = threadIdx.x, z = tid % ZL, line = blockIdx.x * G + tid / ZL, f = blockIdx.y;
  const size_t NYX = (size_t)g.N_y * g.N_x;
  if (line >= (int)NYX || z >= g.N_z) return;
  const int row = line / g.N_x, col = line - row * g.N_x;
  const float xc = (float)((double)g.x0 + ((double)(float)col + 0.5) * (double)g.delta_xy);
  const float yc = (float)((double)g.y0 + ((double)(float)row + 0.5) * (double)g.delta_xy);
  const int nv = proj_num[f];
  const int *vl = views + (size_t)f * g.max_views;
  float acc = 0.f;
  for (int p = 0; p < nv; p++) {
    const int v = vl[p];
    const float r = xc * cosv[v] - yc * sinv[v];
    const float u = (r - g.R0) / g.delta_r - 0.5f;
    const float fl = floorf(u);
    const int i0 = (int)fl;
    const float fr = u - fl;
    if (i0 < -1 || i0 >= g.N_r) continue;
    // a slice of zo detector rows takes the mean of its rows
    float a = 0.f, b = 0.f;
    for (int r = 0; r < g.zo; r++) {
      const int t = z * g.zo + r;
      a += i0 >= 0 ? q[((size_t)v * g.N_r + i0) * g.N_t + t] : 0.f;
      b += i0 + 1 < g.N_r ? q[((size_t)v * g.N_r + i0 + 1) * g.N_t + t] : 0.f;
    }
    if (g.zo > 1) {
      a /= (float)g.zo;
      b /= (float)g.zo;
    }
    acc += wv[v] * (a + fr * (b - a));
  }
  if (!mask[line]) acc = 0.f;
  if (g.positivity && acc < 0.f) acc = 0.f;
  x[((size_t)f * NYX + line) * g.N_z + z] = acc;
}

// ---------------------------------------------------------------------------
// qGGMRF surrogate weight b' = b * rho'(delta)/delta (F/XT_ICD_update.c:89-125,
// 141-168), float.
__device__ __forceinline__ float xt_qggmrf_ratio(float delta, float p, float c, float sq,
                                                 float sqp) {
  if (delta == 0.f) return 2.0f / (sq * c);
  float ad = fabsf(delta);
  float t1 = powf(ad, 2.0f - p) / sqp;
  float t3 = c + t1;
  // rho'(delta)/delta = |delta|^(q-2)/(t3*sq) * (q - (q-p)*t1/t3), q = 2
  return (1.0f / (t3 * sq)) * (2.0f - ((2.0f - p) * t1) / t3);
}

// Per-launch work description of the voxel-update kernel.
struct XtIcdLaunch {
  int mode;        // 0 homogeneous share, 1 NHICD lines, 2 the given list in order (one
                   // workgroup, every z-block; per-voxel checks), 3 the given list in
                   // order, the z-blocks of parity half only (per-voxel checks of the
                   // production pass)
  int half;        // 0/1: the reference's even and odd z-block passes: frame f updates
                   // the slices of z-blocks with parity half ^ (f & 1)
                   // (F/XT_ICD_update.c:799-881); each is followed by a ring-offset update
  int sub_iter;    // j of K
  int K;           // NHICD_Iterations
  uint32_t seed;
  int count_stats; // accumulate total_vox_mag (homogeneous sweeps)
  int zorder;      // 0: slices in red-black order; 1: ascending, one slice at a time
                   // (the reference's slice loop, F/XT_VoxUpdate.c:135; per-voxel checks)
};

struct XtIcdBuffers {
  float *e;
  const float *w;
  float *x;
  const float *Hr;
  const float *cosv, *sinv;
  const int *views;          // [N_time][max_views]
  const int *proj_num;       // [N_time]
  const int *lines;          // masked lines y*N_x + x [n_lines] (mode 2: the list to update)
  int n_lines;
  const int *nhicd_list;     // [N_time][2][N_y*N_x]: NHICD lines of each frame and z-block
                             // parity (raster order)
  const int *nhicd_count;    // [N_time][2]
  const uint8_t *nhicd_flag; // [N_time][2][N_y][N_x]: line selected by some z-block of
                             // that parity
  const uint8_t *nhicd_blk;  // [N_time][nb][N_y][N_x]: selected by that z-block
  float *mag_blk;            // [N_time][nb][N_y][N_x]
  float *magsum;             // [N_time][N_y][N_x]
  double *stats;             // [0] total_vox_mag, [1] voxel updates, [2] footprint entries,
                             // [3] line visits
  const float4 *at_vals;     // A-column table [N_time][N_y*N_x][max_views] (or null)
  const int *at_first;       // first bin of each table entry
  const int *iter;           // ICD iteration number (device memory, so a captured
                             // graph of one iteration replays for every iteration)
  const uint8_t *vmask;      // [N_time][N_y][N_x][N_z] voxels to update (edge-only
                             // updates), or null: every voxel of a listed line
  const int *groups;         // line groups of the homogeneous passes: [n_groups][gl]
                             // member lines (-1: none), adjacent lines of one tile
  int n_groups;
  int gl;                    // lines per group (1: the single-line kernel)
};

// ---------------------------------------------------------------------------
// A-column table: the exact columns of every (frame, line, view), built once
// per reconstruction and read by the voxel update instead of rebuilding each
// column per line visit.  An entry is the first non-zero bin and the next four
// values (20 B; a column of delta_xy <= 1.5 delta_r spans at most 4 bins).
// *overflow is set when some column needs more than four bins; the host then
// keeps building columns on the fly.
__global__ void k_build_atable(XtGeom g, const float *__restrict__ Hr, const float *__restrict__ cosv,
                               const float *__restrict__ sinv, const int *__restrict__ views,
                               const int *__restrict__ proj_num, float4 *__restrict__ at_vals,
                               int *__restrict__ at_first, int *__restrict__ overflow) {
  const size_t NYX = (size_t)g.N_y * g.N_x;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= NYX * g.max_views) return;
  const int c = (int)(i / g.max_views), p = (int)(i % g.max_views);
  const size_t ti = (size_t)f * NYX * g.max_views + i;
  float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
  int fb = 0;
  if (p < proj_num[f]) {
    const int v = views[(size_t)f * g.max_views + p];
    const int row = c / g.N_x, col = c - row * g.N_x;
    int first;
    float vals[XT_MAXA];
    xt_amatrix_column(g, Hr + (size_t)v * XT_HR_STRIDE, cosv[v], sinv[v], row, col, &first, vals);
    int b0 = 0;
    while (b0 < XT_MAXA && !(vals[b0] > 0.f)) b0++;
    if (b0 < XT_MAXA) {
      fb = first + b0;
      float t[4];
      for (int k = 0; k < 4; k++) t[k] = (b0 + k < XT_MAXA) ? vals[b0 + k] : 0.f;
      for (int k = b0 + 4; k < XT_MAXA; k++)
        if (vals[k] > 0.f) atomicOr(overflow, 1);
      out = make_float4(t[0], t[1], t[2], t[3]);
    }
  }
  at_vals[ti] = out;
  at_first[ti] = fb;
}

__device__ __forceinline__ uint32_t xt_hash(uint32_t a) {
  a ^= a >> 16;
  a *= 0x7feb352dU;
  a ^= a >> 15;
  a *= 0x846ca68bU;
  a ^= a >> 16;
  return a;
}

// Key of the homogeneous line order of one iteration, frame and z-block parity.
// The reference draws an independent random order for every (frame, z-block)
// (F/XT_ICD_update.c:599-631); a line visit here updates all slices of one
// parity at once, so the blocks of a parity share an order, but the two
// parities take different ones.  With one order for both parities a slice and
// its z-neighbours (always in the other parity) follow the same in-plane
// update sequence, their errors stay alike, and the z-prior no longer averages
// them out: on a 512x512x8 slab with 1024 views the image after three
// iterations is 21.5% (RMSE / mean attenuation) from the converged one against
// 16.3% with the reference's per-block orders and 18.5% with per-parity orders
// (the reference's sequential schedule in every case).
__device__ __forceinline__ uint32_t xt_order_key(uint32_t seed, int iter, int f, int par) {
  return xt_hash(seed ^ xt_hash((uint32_t)iter * 0x85EBCA6Bu ^
                                xt_hash((uint32_t)f + 17u + 0x10000u * (uint32_t)(par + 1))));
}

// A keyed pseudo-random permutation of [0, n): a 4-round Feistel network on
// the smallest even number of bits covering n, with cycle walking back into
// range (each walk ends: the cycle through i < n returns below n).
__device__ __forceinline__ int xt_permute(int i, int n, uint32_t key) {
  int bits = 2;
  while ((1 << bits) < n) bits += 2;
  const int h = bits >> 1;
  const uint32_t mask = (1u << h) - 1u;
  uint32_t v = (uint32_t)i;
  do {
    uint32_t l = v >> h, r = v & mask;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t t = l ^ (xt_hash(r ^ (key + 0x9E3779B9u * (uint32_t)(k + 1))) & mask);
      l = r;
      r = t;
    }
    v = (l << h) | r;
  } while (v >= (uint32_t)n);
  return (int)v;
}

// ---------------------------------------------------------------------------
// Voxel-line updates (F/XT_VoxUpdate.c:11-254 with the surrogate of
// F/XT_ICD_update.c:130-179).  Grid = (workgroups per frame, frames); each
// workgroup takes lines i = lo + blockIdx.x + k*gridDim.x of the launch's range
// of a random order of the frame's lines (the reference's random_select order
// and its xy_start..xy_end share, F/XT_ICD_update.c:599-631,807-809), or a
// contiguous run of the raster-ordered NHICD list, and updates them one after
// another, each visit updating the line's slices of one z-block parity.  Per
// visit:
//  1. all threads stage the line's A columns (table or built on the fly) and
//     the 3x3 neighbourhood columns plus the temporal neighbours into LDS;
//  2. threads (view group grp, lane zl) sweep the sinogram footprint for
//     theta1/theta2, U views in flight per step, the view walk rotated per line
//     so concurrent workgroups spread over the whole sinogram.  With PACK the
//     active parity's slices are packed into ZL/2 lanes per view group (twice
//     the views per step); buffer loads take the bin offset as a scalar, and a
//     bin outside the column reads 0 through an out-of-range offset;
//  3. wave 0 updates the slices in red-black order (exact Gauss-Seidel: slices
//     of a line share no sinogram entries);
//  4. all threads subtract A*delta from e with no-return buffer atomics, and
//     wait for them before the next line, whose loads then see them.
// Lines updated concurrently by different workgroups read each other's e and
// x as they happen to be (a bounded Jacobi coupling; the concurrency is capped
// on the host).  The body serves one slab (k_icd_lines: bx = blockIdx.x of gx
// workgroups per frame, frame f = blockIdx.y) or one slab of a batch
// (k_icd_lines_multi).
template <int ZL, bool TAB, bool PACK, int MA = XT_MAXA, int NT = 256>
__device__ __forceinline__ void icd_lines_body(const XtGeom &g, const XtIcdBuffers &B,
                                               const XtIcdLaunch &L, const int bx, const int gx,
                                               const int f) {
  // NT threads per workgroup (256, or 64: one wave per line, for short lines)
  static_assert(NT == 64 || NT == 256, "workgroup of one or four waves");
  constexpr bool WIDE = !TAB && MA > XT_MAXA;  // columns of more than 6 bins, in LDS as floats
  constexpr int ZP = PACK ? ZL / 2 : ZL;  // lanes per view group
  constexpr int G = NT / ZP;              // view groups
  constexpr int U = TAB ? 4 : (WIDE ? 1 : 2);  // views in flight per group and step
  constexpr int NA = TAB ? 4 : MA;        // bins a column can hold
  constexpr uint32_t OOB = 0x80000000u;   // buffer offset that reads 0 / drops the atomic
  extern __shared__ float4 smem4[];
  float4 *a_v4 = smem4;                                            // [max_views] bins 0..3
  float2 *a_v2 = (float2 *)(a_v4 + g.max_views);                   // [max_views] bins 4..5
  float *a_w = (float *)smem4;                                     // WIDE: [max_views][MA]
  int *a_base = WIDE ? (int *)(a_w + (size_t)g.max_views * MA)
                     : (int *)(a_v2 + (TAB ? 0 : g.max_views));    // [max_views] e offset
  float *red1 = (float *)(a_base + g.max_views);                   // [NT]
  float *red2 = red1 + NT;                                         // [NT]
  float *xn = red2 + NT;                                           // [11][ZL]
  float *delta = xn + 11 * ZL;                                     // [ZL]
  float *th1z = delta + ZL;                                        // [ZL]
  float *th2z = th1z + ZL;                                         // [ZL]
  float *magb = th2z + ZL;                                         // [64]
  __shared__ int s_cnt_col, s_perf;

  const int tid = threadIdx.x;
  const int nv = B.proj_num[f];
  const int *vlist = B.views + (size_t)f * g.max_views;
  const size_t NYX = (size_t)g.N_y * g.N_x;
  const int iter = *B.iter;
  const int blk_sz = g.N_z / g.nb;
  // z-block parity this launch updates (every block for the list mode)
  const int par = (L.mode == 2) ? -1 : ((L.half ^ f) & 1);
  // this thread's lane: view group grp, slice zs (packed: the k-th slice of
  // the active parity's blocks, blocks being blk_sz slices each)
  const int grp = tid / ZP, lz = tid % ZP;
  int zs = lz;
  if constexpr (PACK) zs = ((lz / blk_sz) * 2 + par) * blk_sz + lz % blk_sz;
  // z-block of a slice: block b holds slices [b*floor(N_z/nb), (b+1)*floor(N_z/nb)),
  // the last one up to N_z-1 (F/XT_ICD_update.c:804-806), so with fewer slices
  // than blocks every slice is in the last block
  auto zblock = [&](int z) {
    int b = blk_sz > 0 ? z / blk_sz : g.nb - 1;
    return b >= g.nb ? g.nb - 1 : b;
  };
  const bool zs_ok = zs < g.N_z && (par < 0 || (zblock(zs) & 1) == par);

  // ---- this launch's range of positions, and the order that maps them to lines
  int lo = 0, hi = 0, n_dom = 0;
  const int *list = B.lines;
  uint32_t key = 0;
  if (L.mode == 0) {
    const int n = B.n_lines;
    const int per = n / L.K;
    lo = L.sub_iter * per;
    hi = (L.sub_iter == L.K - 1) ? n : (L.sub_iter + 1) * per;
    n_dom = n;
    key = xt_order_key(L.seed, iter, f, par);
  } else if (L.mode == 1) {
    lo = 0;
    hi = B.nhicd_count[f * 2 + par];
    list = B.nhicd_list + ((size_t)f * 2 + par) * NYX;
    key = xt_hash(L.seed ^ 0x5BD1E995u ^ xt_hash((uint32_t)iter * 0xC2B2AE35u ^ (uint32_t)f));
  } else {
    lo = 0;
    hi = B.n_lines;
  }
  if (tid == 0) {
    s_cnt_col = 0;
    s_perf = 0;
  }
  float vox_mag = 0.f;     // lane-local sum of |new x| (wave 0)
  double n_upd = 0.0, n_foot = 0.0, n_vis = 0.0;
  const size_t tbl = (size_t)f * NYX * g.max_views;
  const uint32_t sino_bytes = (uint32_t)((size_t)g.N_p * g.N_r * g.N_t * sizeof(float));
  const __amdgpu_buffer_rsrc_t rs_e = __builtin_amdgcn_make_buffer_rsrc(B.e, 0, sino_bytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rs_w =
      __builtin_amdgcn_make_buffer_rsrc((void *)B.w, 0, sino_bytes, 0x00020000);
  const int bin_bytes = (g.pm ? g.H : g.N_t) * (int)sizeof(float);
  // first entry of the launch's parity plane (split layout)
  const int plane = g.pm ? par * g.N_p * g.N_r * g.H : 0;
  const int row_len = g.pm ? g.H : g.N_t;

  // homogeneous lines: positions bx, bx + gx, ... of the random order; NHICD
  // lines: a contiguous run of the raster-ordered list per workgroup (the
  // reference walks its selection lists in raster order, F/XT_NHICD.c:189-199)
  int i0 = lo + bx, i1 = hi, istep = gx;
  if (L.mode == 1) {
    const long long len = hi - lo;
    i0 = lo + (int)(len * bx / gx);
    i1 = lo + (int)(len * (bx + 1) / gx);
    istep = 1;
  }
  for (int i = i0; i < i1; i += istep) {
    const int c = list[n_dom ? xt_permute(i, n_dom, key) : i];
    const int yy = c / g.N_x, xx = c - yy * g.N_x;
    // lane of this thread active: its slice is in the launch's parity, for NHICD
    // lines its own z-block selected the line (F/XT_ICD_update.c:844-881 walk each
    // block's own list), and with edge-only updates its voxel is in the mask (an
    // unmasked voxel is not updated, so its theta is not needed)
    const bool lane_on =
        zs_ok && (L.mode != 1 || B.nhicd_blk[((size_t)f * g.nb + zblock(zs)) * NYX + c] != 0) &&
        (!B.vmask || B.vmask[((size_t)f * NYX + c) * g.N_z + zs] != 0);

    // ---- 1. stage the A columns (scaled by delta_t) and the neighbourhood
    int cnt_local = 0;
    if constexpr (TAB) {
      // four table entries per thread in flight at once
      const size_t row = tbl + (size_t)c * g.max_views;
      for (int p0 = tid; p0 < nv; p0 += 4 * NT) {
        float4 t4[4];
        int first[4], v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int p = p0 + u * NT;
          t4[u] = make_float4(0.f, 0.f, 0.f, 0.f);
          first[u] = 0;
          v[u] = 0;
          if (p < nv) {
            t4[u] = B.at_vals[row + p];
            first[u] = B.at_first[row + p];
            v[u] = vlist[p];
          }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int p = p0 + u * NT;
          if (p < nv) {
            a_v4[p] = make_float4(t4[u].x * g.delta_t, t4[u].y * g.delta_t, t4[u].z * g.delta_t,
                                  t4[u].w * g.delta_t);
            a_base[p] = plane + (v[u] * g.N_r + first[u]) * row_len;
            cnt_local += (t4[u].x > 0.f) + (t4[u].y > 0.f) + (t4[u].z > 0.f) + (t4[u].w > 0.f);
          }
        }
      }
    } else {
      for (int p = tid; p < nv; p += NT) {
        const int v = vlist[p];
        int first = 0;
        float vals[MA];
        xt_amatrix_column(g, B.Hr + (size_t)v * XT_HR_STRIDE, B.cosv[v], B.sinv[v], yy, xx, &first,
                          vals);
        if constexpr (WIDE) {
#pragma unroll
          for (int b = 0; b < MA; b++) a_w[(size_t)p * MA + b] = vals[b] * g.delta_t;
        } else {
          a_v4[p] = make_float4(vals[0] * g.delta_t, vals[1] * g.delta_t, vals[2] * g.delta_t,
                                vals[3] * g.delta_t);
          a_v2[p] = make_float2(vals[4] * g.delta_t, vals[5] * g.delta_t);
        }
        a_base[p] = plane + (v * g.N_r + first) * row_len;
#pragma unroll
        for (int b = 0; b < MA; b++) cnt_local += vals[b] > 0.f;
      }
    }
    for (int q = tid; q < 11 * ZL; q += NT) {
      const int k = q / ZL, z = q - k * ZL;
      float val = 0.f;
      if (z < g.N_z) {
        if (k < 9) {
          const int y2 = yy + k / 3 - 1, x2 = xx + k % 3 - 1;
          if (y2 >= 0 && y2 < g.N_y && x2 >= 0 && x2 < g.N_x)
            val = __hip_atomic_load(&B.x[((size_t)f * NYX + (size_t)y2 * g.N_x + x2) * g.N_z + z],
                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          const int f2 = f + (k == 9 ? -1 : 1);
          if (f2 >= 0 && f2 < g.N_time)
            val = __hip_atomic_load(&B.x[((size_t)f2 * NYX + c) * g.N_z + z], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      xn[q] = val;
    }
    if (cnt_local) atomicAdd(&s_cnt_col, cnt_local);
    __syncthreads();

    // ---- 2. theta pass
    const int rot = (int)(xt_hash((uint32_t)c * 0x27D4EB2Fu ^ key ^ (uint32_t)L.half) %
                          (uint32_t)(nv > 0 ? nv : 1));
    // lane's entry in a row: its slice's first detector row, or in the split layout
    // its packed position; a slice spans zo rows (z_overlap_num; 1 when packed)
    const int zo = PACK ? 1 : g.zo;
    const uint32_t lane_off = (uint32_t)(g.pm ? lz : zs * zo) * sizeof(float);
    const float T2 = g.T * g.T, Td = g.T * g.delta;
    float th1 = 0.f, th2 = 0.f;
    if (lane_on) {
      for (int q0 = grp; q0 < nv; q0 += U * G) {
        uint32_t off[U];
        float av[U][NA];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int q = q0 + u * G;
#pragma unroll
          for (int b = 0; b < NA; b++) av[u][b] = 0.f;
          off[u] = OOB;
          if (q < nv) {
            int p = q + rot;
            if (p >= nv) p -= nv;
            off[u] = (uint32_t)a_base[p] * sizeof(float) + lane_off;
            if constexpr (WIDE) {
#pragma unroll
              for (int b = 0; b < NA; b++) av[u][b] = a_w[(size_t)p * MA + b];
            } else {
              const float4 a4 = a_v4[p];
              av[u][0] = a4.x;
              av[u][1] = a4.y;
              av[u][2] = a4.z;
              av[u][3] = a4.w;
              if constexpr (!TAB) {
                const float2 a2 = a_v2[p];
                av[u][4] = a2.x;
                av[u][5] = a2.y;
              }
            }
          }
        }
        for (int r = 0; r < zo; r++) {  // the slice's detector rows
          float ev[U][NA], wv[U][NA];
#pragma unroll
          for (int u = 0; u < U; u++)
#pragma unroll
            for (int b = 0; b < NA; b++) {
              const uint32_t o = av[u][b] > 0.f ? off[u] : OOB;
              const int so = b * bin_bytes + r * (int)sizeof(float);
              ev[u][b] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs_e, o, so, 0));
              wv[u][b] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs_w, o, so, 0));
            }
#pragma unroll
          for (int u = 0; u < U; u++)
#pragma unroll
            for (int b = 0; b < NA; b++) {
              const float pe = av[u][b], e1 = ev[u][b], w1 = wv[u][b];
              // generalized Huber weight: w if |e*sqrt(w)| < T, else T*delta*sqrt(w)/|e|
              // (F/XT_ICD_update.c:577, F/XT_VoxUpdate.c:38-48,76)
              const float wr = Td * __builtin_amdgcn_sqrtf(w1) * __builtin_amdgcn_rcpf(fabsf(e1));
              const float ws = (e1 * e1 * w1 < T2) ? w1 : wr;
              const float pw = pe * ws;
              th2 += pe * pw;
              th1 -= e1 * pw;
            }
        }
      }
    }
    // ---- theta reduction by slice: over the view groups of a wave with lane
    // shuffles (64 / ZP groups per wave), then over the waves through LDS
    constexpr int WAVES = NT / 64;
#pragma unroll
    for (int m = ZP; m < 64; m <<= 1) {
      th1 += __shfl_xor(th1, m);
      th2 += __shfl_xor(th2, m);
    }
    if ((tid & 63) < ZP) {
      red1[(tid >> 6) * ZP + lz] = th1;
      red2[(tid >> 6) * ZP + lz] = th2;
    }
    __syncthreads();
    if (tid < ZP) {
      float t1 = 0.f, t2 = 0.f;
#pragma unroll
      for (int q = 0; q < WAVES; q++) {
        t1 += red1[q * ZP + tid];
        t2 += red2[q * ZP + tid];
      }
      if (zs < ZL) {
        th1z[zs] = t1;
        th2z[zs] = t2;
      }
    }
    if (tid < ZL) delta[tid] = 0.f;
    if (tid < 64) magb[tid] = 0.f;
    __syncthreads();

    // ---- 3. surrogate update of every active slice, red-black in z (or
    // ascending); the line's own column is xn[4][*] and is updated in place
    float *xcol = xn + 4 * ZL;
    const int n_steps = L.zorder ? g.N_z : 2;
    for (int step = 0; step < n_steps; step++) {
      const int z = tid;
      const bool mine = L.zorder ? (z == step) : ((z & 1) == step);
      bool upd = false;
      if (z < g.N_z && mine) {
        upd = par < 0 || (zblock(z) & 1) == par;
        if (upd && L.mode == 1) upd = B.nhicd_blk[((size_t)f * g.nb + zblock(z)) * NYX + c] != 0;
        if (upd && B.vmask) upd = B.vmask[((size_t)f * NYX + c) * g.N_z + z] != 0;
      }
      if (upd) {
        const float V = xcol[z];
        float nb[27];
        bool fl[27];
        bool any_pos = false;
#pragma unroll
        for (int dz = -1; dz <= 1; dz++)
#pragma unroll
          for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
              const int l = ((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1);
              const int z2 = z + dz, y2 = yy + dy, x2 = xx + dx;
              const bool ok = (l != 13) && z2 >= 0 && z2 < g.N_z && y2 >= 0 && y2 < g.N_y &&
                              x2 >= 0 && x2 < g.N_x;
              const float val = ok ? xn[((dy + 1) * 3 + (dx + 1)) * ZL + z2] : 0.f;
              nb[l] = val;
              fl[l] = ok;
              any_pos |= ok && (val > 0.f);
            }
        const bool tf0 = f > 0, tf1 = f + 1 < g.N_time;
        const float tn0 = xn[9 * ZL + z], tn1 = xn[10 * ZL + z];
        any_pos |= (tf0 && tn0 > 0.f) || (tf1 && tn1 > 0.f);
        // zero skipping, F/XT_VoxUpdate.c:208-232
        const bool skip = (V == 0.f) && (iter > 1) && !any_pos;
        if (!skip) {
          float s1 = 0.f, s2 = 0.f;
#pragma unroll
          for (int l = 0; l < 27; l++) {
            if (fl[l]) {
              const float bq =
                  g.filt[l] * xt_qggmrf_ratio(V - nb[l], g.MRF_P, g.C_S, g.Sigma_S_Q, g.Sigma_S_Q_P);
              s1 += bq * nb[l];
              s2 += bq;
            }
          }
          if (tf0) {
            const float bq =
                g.tfilt * xt_qggmrf_ratio(V - tn0, g.MRF_P, g.C_T, g.Sigma_T_Q, g.Sigma_T_Q_P);
            s1 += bq * tn0;
            s2 += bq;
          }
          if (tf1) {
            const float bq =
                g.tfilt * xt_qggmrf_ratio(V - tn1, g.MRF_P, g.C_T, g.Sigma_T_Q, g.Sigma_T_Q_P);
            s1 += bq * tn1;
            s2 += bq;
          }
          const float T1 = th1z[z], T2z = th2z[z];
          const float u = (s1 + (T2z * V) - T1) / (s2 + T2z);
          float nv2 = V + g.alpha * (u - V);
          if (g.positivity && nv2 <= 0.f) nv2 = 0.f;
          xcol[z] = nv2;
          delta[z] = nv2 - V;
          atomicAdd(&magb[zblock(z)], fabsf(nv2 - V));
          vox_mag += fabsf(nv2);
          atomicAdd(&s_perf, 1);
        }
      }
      __syncthreads();
    }

    // ---- write the line back and its update magnitude
    if (tid < g.N_z)
      __hip_atomic_store(&B.x[((size_t)f * NYX + c) * g.N_z + tid], xcol[tid], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    if (tid < g.nb) {
      // MagUpdateMap of this line: reset and re-accumulated by the blocks that
      // updated it (F/XT_VoxUpdate.c:128,236); other blocks keep theirs
      const size_t mi = ((size_t)f * g.nb + tid) * NYX + c;
      if ((par < 0 || (tid & 1) == par) && (L.mode != 1 || B.nhicd_blk[mi] != 0))
        B.mag_blk[mi] = magb[tid];
      else
        magb[tid] = B.mag_blk[mi];
    }

    // ---- 4. e -= A * delta over the same footprint
    const float dz = (zs < ZL && lane_on) ? delta[zs] : 0.f;
    if (dz != 0.f) {
      for (int q = grp; q < nv; q += G) {
        int p = q + rot;
        if (p >= nv) p -= nv;
        const uint32_t o = (uint32_t)a_base[p] * sizeof(float) + lane_off;
        float av[NA > XT_MAXA ? NA : XT_MAXA] = {};
        if constexpr (WIDE) {
#pragma unroll
          for (int b = 0; b < NA; b++) av[b] = a_w[(size_t)p * MA + b];
        } else {
          const float4 a4 = a_v4[p];
          av[0] = a4.x;
          av[1] = a4.y;
          av[2] = a4.z;
          av[3] = a4.w;
          if constexpr (!TAB) {
            const float2 a2 = a_v2[p];
            av[4] = a2.x;
            av[5] = a2.y;
          }
        }
        for (int r = 0; r < zo; r++)
#pragma unroll
          for (int b = 0; b < NA; b++)
            __builtin_amdgcn_raw_ptr_buffer_atomic_fadd_f32(-(av[b] * dz), rs_e, av[b] > 0.f ? o : OOB,
                                                        b * bin_bytes + r * (int)sizeof(float), 0);
      }
    }
    // the next line's loads must see these atomics: wait for all of this
    // thread's memory operations, then for every thread of the workgroup, and
    // drop any stale copies of e from this CU's cache
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (tid == 0) {
      float sm = 0.f;
      for (int b = 0; b < g.nb; b++) sm += magb[b];
      B.magsum[(size_t)f * NYX + c] = sm;
      n_upd += (double)s_perf;
      n_foot += (double)s_perf * (double)s_cnt_col * (double)g.zo;
      n_vis += 1.0;
      s_cnt_col = 0;
      s_perf = 0;
    }
    __syncthreads();
  }
  // ---- per-workgroup statistics
  if (tid < ZL && L.count_stats && vox_mag != 0.f) atomicAdd(&B.stats[0], (double)vox_mag);
  if (tid == 0 && n_vis != 0.0) {
    atomicAdd(&B.stats[1], n_upd);
    atomicAdd(&B.stats[2], n_foot);
    atomicAdd(&B.stats[3], n_vis);
  }
}

// ---------------------------------------------------------------------------
// Voxel-line updates in groups of GL lines (homogeneous passes: the lines of a
// 1 x 2 or 2 x 2 tile of the image, groups taken in a random order; NHICD and
// listed lines: GL consecutive entries of the list), with the A columns read
// from the table.  A group's lines are updated one after another exactly as
// single lines are (same neighbourhood, theta sums and surrogate), except that
// the error sinogram is written once per group: line l reads e and subtracts,
// entry by entry, the changes A_m * delta_m of the group's earlier lines m < l
// -- the value e would hold had they been applied -- and after the last line
// one atomic per entry of the union of the group's columns applies the sum of
// all their changes.  Adjacent lines' columns share most bins, so a 2 x 2
// group issues about half the atomics of four single lines.  Columns of one
// view that do not fit the union window of SPAN bins get atomics of their own,
// so any grouping is exact.  Needs slices of one detector row (zo == 1).
template <int ZL, bool PACK, int NT, int GL>
__device__ __forceinline__ void icd_groups_body(const XtGeom &g, const XtIcdBuffers &B,
                                                const XtIcdLaunch &L, const int bx, const int gx,
                                                const int f) {
  static_assert(NT == 64 || NT == 256, "workgroup of one or four waves");
  static_assert(GL == 2 || GL == 4, "groups of two or four lines");
  constexpr int ZP = PACK ? ZL / 2 : ZL;  // lanes per view group
  constexpr int G = NT / ZP;              // view groups
  constexpr int U = 4;                    // views in flight per group and step
  constexpr int SPAN = 8;                 // bins of a view's union window
  constexpr uint32_t OOB = 0x80000000u;   // buffer offset that reads 0 / drops the atomic
  extern __shared__ float4 smem4[];
  int *vlist_s = (int *)smem4;                                     // [max_views] view index
  float *red1 = (float *)(vlist_s + g.max_views);                  // [NT]
  float *red2 = red1 + NT;                                         // [NT]
  float *xn = red2 + NT;                                           // [11][ZL]
  float *dxg = xn + 11 * ZL;                                       // [GL][ZL] changes
  float *th1z = dxg + GL * ZL;                                     // [ZL]
  float *th2z = th1z + ZL;                                         // [ZL]
  float *magb = th2z + ZL;                                         // [64]
  __shared__ int s_cnt_col, s_perf;

  const int tid = threadIdx.x;
  const int nv = B.proj_num[f];
  const size_t NYX = (size_t)g.N_y * g.N_x;
  const int iter = *B.iter;
  const int blk_sz = g.N_z / g.nb;
  const int par = (L.mode == 2) ? -1 : ((L.half ^ f) & 1);
  const int grp = tid / ZP, lz = tid % ZP;
  int zs = lz;
  if constexpr (PACK) zs = ((lz / blk_sz) * 2 + par) * blk_sz + lz % blk_sz;
  auto zblock = [&](int z) {
    int b = blk_sz > 0 ? z / blk_sz : g.nb - 1;
    return b >= g.nb ? g.nb - 1 : b;
  };
  const bool zs_ok = zs < g.N_z && (par < 0 || (zblock(zs) & 1) == par);
  for (int p = tid; p < nv; p += NT) vlist_s[p] = B.views[(size_t)f * g.max_views + p];

  // ---- this launch's items: groups of the random order (mode 0); runs of GL
  // consecutive entries of the workgroup's contiguous share of the NHICD list
  // (mode 1) or of the line list (modes 2, 3: items bx, bx + gx, ...)
  int n_items = 0, i0 = 0, istep = 1, r0 = 0, r1 = 0, n_dom = 0;
  const int *list = B.lines;
  uint32_t key = 0;
  if (L.mode == 0) {
    const int n = B.n_groups;
    const int per = n / L.K;
    const int lo = L.sub_iter * per;
    n_items = (L.sub_iter == L.K - 1) ? n : (L.sub_iter + 1) * per;
    i0 = lo + bx;
    istep = gx;
    n_dom = n;
    key = xt_order_key(L.seed, iter, f, par);
  } else if (L.mode == 1) {
    const int hi = B.nhicd_count[f * 2 + par];
    list = B.nhicd_list + ((size_t)f * 2 + par) * NYX;
    key = xt_hash(L.seed ^ 0x5BD1E995u ^ xt_hash((uint32_t)iter * 0xC2B2AE35u ^ (uint32_t)f));
    r0 = (int)((long long)hi * bx / gx);
    r1 = (int)((long long)hi * (bx + 1) / gx);
    n_items = (r1 - r0 + GL - 1) / GL;
  } else {
    r1 = B.n_lines;
    n_items = (r1 + GL - 1) / GL;
    i0 = bx;
    istep = gx;
  }
  if (tid == 0) {
    s_cnt_col = 0;
    s_perf = 0;
  }
  float vox_mag = 0.f;
  double n_upd = 0.0, n_foot = 0.0, n_vis = 0.0;
  const size_t tbl = (size_t)f * NYX * g.max_views;
  const uint32_t sino_bytes = (uint32_t)((size_t)g.N_p * g.N_r * g.N_t * sizeof(float));
  const __amdgpu_buffer_rsrc_t rs_e = __builtin_amdgcn_make_buffer_rsrc(B.e, 0, sino_bytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rs_w =
      __builtin_amdgcn_make_buffer_rsrc((void *)B.w, 0, sino_bytes, 0x00020000);
  const int bin_bytes = (g.pm ? g.H : g.N_t) * (int)sizeof(float);
  const int plane = g.pm ? par * g.N_p * g.N_r * g.H : 0;
  const int row_len = g.pm ? g.H : g.N_t;
  const uint32_t lane_off = (uint32_t)(g.pm ? lz : zs) * sizeof(float);
  const float T2 = g.T * g.T, Td = g.T * g.delta, dt = g.delta_t;
  __syncthreads();

  for (int i = i0; i < n_items; i += istep) {
    int cl[GL];
    uint32_t gkey;
    if (L.mode == 0) {
      const int gi = xt_permute(i, n_dom, key);
#pragma unroll
      for (int m = 0; m < GL; m++) cl[m] = B.groups[(size_t)gi * GL + m];
      gkey = (uint32_t)gi;
    } else {
      const int e0 = r0 + i * GL;
#pragma unroll
      for (int m = 0; m < GL; m++) cl[m] = e0 + m < r1 ? list[e0 + m] : -1;
      gkey = (uint32_t)cl[0];
    }
    const int rot = (int)(xt_hash(gkey * 0x27D4EB2Fu ^ key ^ (uint32_t)L.half) %
                          (uint32_t)(nv > 0 ? nv : 1));
    for (int q = tid; q < GL * ZL; q += NT) dxg[q] = 0.f;

    for (int l = 0; l < GL; l++) {
      const int c = cl[l];
      if (c < 0) continue;  // the same for every thread of the workgroup
      const int yy = c / g.N_x, xx = c - yy * g.N_x;
      const bool lane_on =
          zs_ok && (L.mode != 1 || B.nhicd_blk[((size_t)f * g.nb + zblock(zs)) * NYX + c] != 0);
      const size_t row_l = tbl + (size_t)c * g.max_views;
      // ---- neighbourhood (the group's earlier lines are already written back)
      // and the column's entry count
      int cnt_local = 0;
      for (int p = tid; p < nv; p += NT) {
        const float4 t4 = B.at_vals[row_l + p];
        cnt_local += (t4.x > 0.f) + (t4.y > 0.f) + (t4.z > 0.f) + (t4.w > 0.f);
      }
      for (int q = tid; q < 11 * ZL; q += NT) {
        const int k = q / ZL, z = q - k * ZL;
        float val = 0.f;
        if (z < g.N_z) {
          if (k < 9) {
            const int y2 = yy + k / 3 - 1, x2 = xx + k % 3 - 1;
            if (y2 >= 0 && y2 < g.N_y && x2 >= 0 && x2 < g.N_x)
              val = __hip_atomic_load(&B.x[((size_t)f * NYX + (size_t)y2 * g.N_x + x2) * g.N_z + z],
                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          } else {
            const int f2 = f + (k == 9 ? -1 : 1);
            if (f2 >= 0 && f2 < g.N_time)
              val = __hip_atomic_load(&B.x[((size_t)f2 * NYX + c) * g.N_z + z], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
          }
        }
        xn[q] = val;
      }
      if (cnt_local) atomicAdd(&s_cnt_col, cnt_local);
      __syncthreads();

      // ---- theta pass over e less the group's earlier lines' changes
      float dprev[GL];
#pragma unroll
      for (int m = 0; m < GL; m++) dprev[m] = (m < l && zs < ZL) ? dxg[m * ZL + zs] * dt : 0.f;
      bool corr = false;
#pragma unroll
      for (int m = 0; m < GL; m++) corr |= dprev[m] != 0.f;
      float th1 = 0.f, th2 = 0.f;
      if (lane_on) {
        for (int q0 = grp; q0 < nv; q0 += U * G) {
          uint32_t off[U];
          int first[U], pv[U];
          float av[U][4], ev[U][4], wv[U][4];
#pragma unroll
          for (int u = 0; u < U; u++) {
            const int q = q0 + u * G;
            int p = q + rot;
            if (p >= nv) p -= nv;
            const bool ok = q < nv;
            pv[u] = ok ? p : -1;
            const float4 t4 = ok ? B.at_vals[row_l + p] : make_float4(0.f, 0.f, 0.f, 0.f);
            first[u] = ok ? B.at_first[row_l + p] : 0;
            av[u][0] = t4.x * dt;
            av[u][1] = t4.y * dt;
            av[u][2] = t4.z * dt;
            av[u][3] = t4.w * dt;
            off[u] = ok ? (uint32_t)(plane + (vlist_s[p] * g.N_r + first[u]) * row_len) *
                                  (uint32_t)sizeof(float) + lane_off
                        : OOB;
          }
#pragma unroll
          for (int u = 0; u < U; u++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
              const uint32_t o = av[u][b] > 0.f ? off[u] : OOB;
              ev[u][b] =
                  __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs_e, o, b * bin_bytes, 0));
              wv[u][b] =
                  __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs_w, o, b * bin_bytes, 0));
            }
          if (corr) {
            // e -= A_m * delta_m of the earlier lines where their columns meet this one's
#pragma unroll
            for (int m = 0; m < GL - 1; m++) {
              if (m < l && cl[m] >= 0) {
                const size_t rm = tbl + (size_t)cl[m] * g.max_views;
#pragma unroll
                for (int u = 0; u < U; u++) {
                  if (pv[u] >= 0) {
                    const float4 t4 = B.at_vals[rm + pv[u]];
                    const int k0 = first[u] - B.at_first[rm + pv[u]];
                    const float am[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                      const int k = k0 + b;
                      float a = 0.f;
#pragma unroll
                      for (int kk = 0; kk < 4; kk++) a = (k == kk) ? am[kk] : a;
                      ev[u][b] -= a * dprev[m];
                    }
                  }
                }
              }
            }
          }
#pragma unroll
          for (int u = 0; u < U; u++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
              const float pe = av[u][b], e1 = ev[u][b], w1 = wv[u][b];
              // generalized Huber weight as in the single-line kernel
              const float wr = Td * __builtin_amdgcn_sqrtf(w1) * __builtin_amdgcn_rcpf(fabsf(e1));
              const float ws = (e1 * e1 * w1 < T2) ? w1 : wr;
              const float pw = pe * ws;
              th2 += pe * pw;
              th1 -= e1 * pw;
            }
        }
      }
      red1[tid] = th1;
      red2[tid] = th2;
      __syncthreads();
      if (tid < ZP) {
        float t1 = 0.f, t2 = 0.f;
        for (int q = 0; q < G; q++) {
          t1 += red1[q * ZP + tid];
          t2 += red2[q * ZP + tid];
        }
        if (zs < ZL) {
          th1z[zs] = t1;
          th2z[zs] = t2;
        }
      }
      if (tid < 64) magb[tid] = 0.f;
      __syncthreads();

      // ---- surrogate update of the active slices, red-black in z (or ascending)
      float *xcol = xn + 4 * ZL;
      const int n_steps = L.zorder ? g.N_z : 2;
      for (int step = 0; step < n_steps; step++) {
        const int z = tid;
        bool upd = false;
        if (z < g.N_z && (L.zorder ? (z == step) : ((z & 1) == step))) {
          upd = par < 0 || (zblock(z) & 1) == par;
          if (upd && L.mode == 1) upd = B.nhicd_blk[((size_t)f * g.nb + zblock(z)) * NYX + c] != 0;
          if (upd && B.vmask) upd = B.vmask[((size_t)f * NYX + c) * g.N_z + z] != 0;
        }
        if (upd) {
          const float V = xcol[z];
          float nb[27];
          bool fl[27];
          bool any_pos = false;
#pragma unroll
          for (int dz = -1; dz <= 1; dz++)
#pragma unroll
            for (int dy = -1; dy <= 1; dy++)
#pragma unroll
              for (int dx = -1; dx <= 1; dx++) {
                const int ll = ((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1);
                const int z2 = z + dz, y2 = yy + dy, x2 = xx + dx;
                const bool ok = (ll != 13) && z2 >= 0 && z2 < g.N_z && y2 >= 0 && y2 < g.N_y &&
                                x2 >= 0 && x2 < g.N_x;
                const float val = ok ? xn[((dy + 1) * 3 + (dx + 1)) * ZL + z2] : 0.f;
                nb[ll] = val;
                fl[ll] = ok;
                any_pos |= ok && (val > 0.f);
              }
          const bool tf0 = f > 0, tf1 = f + 1 < g.N_time;
          const float tn0 = xn[9 * ZL + z], tn1 = xn[10 * ZL + z];
          any_pos |= (tf0 && tn0 > 0.f) || (tf1 && tn1 > 0.f);
          const bool skip = (V == 0.f) && (iter > 1) && !any_pos;
          if (!skip) {
            float s1 = 0.f, s2 = 0.f;
#pragma unroll
            for (int ll = 0; ll < 27; ll++) {
              if (fl[ll]) {
                const float bq = g.filt[ll] *
                                 xt_qggmrf_ratio(V - nb[ll], g.MRF_P, g.C_S, g.Sigma_S_Q, g.Sigma_S_Q_P);
                s1 += bq * nb[ll];
                s2 += bq;
              }
            }
            if (tf0) {
              const float bq =
                  g.tfilt * xt_qggmrf_ratio(V - tn0, g.MRF_P, g.C_T, g.Sigma_T_Q, g.Sigma_T_Q_P);
              s1 += bq * tn0;
              s2 += bq;
            }
            if (tf1) {
              const float bq =
                  g.tfilt * xt_qggmrf_ratio(V - tn1, g.MRF_P, g.C_T, g.Sigma_T_Q, g.Sigma_T_Q_P);
              s1 += bq * tn1;
              s2 += bq;
            }
            const float T1 = th1z[z], T2z = th2z[z];
            const float u = (s1 + (T2z * V) - T1) / (s2 + T2z);
            float nv2 = V + g.alpha * (u - V);
            if (g.positivity && nv2 <= 0.f) nv2 = 0.f;
            xcol[z] = nv2;
            dxg[l * ZL + z] = nv2 - V;
            atomicAdd(&magb[zblock(z)], fabsf(nv2 - V));
            vox_mag += fabsf(nv2);
            atomicAdd(&s_perf, 1);
          }
        }
        __syncthreads();
      }
      // ---- write the line back and its update magnitude
      if (tid < g.N_z)
        __hip_atomic_store(&B.x[((size_t)f * NYX + c) * g.N_z + tid], xcol[tid], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      if (tid < g.nb) {
        const size_t mi = ((size_t)f * g.nb + tid) * NYX + c;
        if ((par < 0 || (tid & 1) == par) && (L.mode != 1 || B.nhicd_blk[mi] != 0))
          B.mag_blk[mi] = magb[tid];
        else
          magb[tid] = B.mag_blk[mi];
      }
      __syncthreads();
      if (tid == 0) {
        float sm = 0.f;
        for (int b = 0; b < g.nb; b++) sm += magb[b];
        B.magsum[(size_t)f * NYX + c] = sm;
        n_upd += (double)s_perf;
        n_foot += (double)s_perf * (double)s_cnt_col;
        n_vis += 1.0;
        s_cnt_col = 0;
        s_perf = 0;
      }
      // the next line's neighbourhood loads must see this line's values
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }

    // ---- e -= sum over the group's lines of A_m * delta_m: per view one atomic
    // per bin of the window [u0, u0 + SPAN) from the leftmost column, columns
    // reaching past it on their own
    float dzm[GL];
    bool any = false;
#pragma unroll
    for (int m = 0; m < GL; m++) {
      dzm[m] = (zs < ZL && zs_ok) ? dxg[m * ZL + zs] * dt : 0.f;
      any |= dzm[m] != 0.f;
    }
    if (any) {
      for (int q = grp; q < nv; q += G) {
        int p = q + rot;
        if (p >= nv) p -= nv;
        float am[GL][4];
        int fm[GL];
        int u0 = 0x7fffffff;
#pragma unroll
        for (int m = 0; m < GL; m++) {
          float4 t4 = make_float4(0.f, 0.f, 0.f, 0.f);
          fm[m] = 0;
          if (cl[m] >= 0) {
            const size_t rm = tbl + (size_t)cl[m] * g.max_views + p;
            t4 = B.at_vals[rm];
            fm[m] = B.at_first[rm];
            u0 = min(u0, fm[m]);
          }
          am[m][0] = t4.x;
          am[m][1] = t4.y;
          am[m][2] = t4.z;
          am[m][3] = t4.w;
        }
        const uint32_t vrow = (uint32_t)(plane + vlist_s[p] * g.N_r * row_len);
        const uint32_t o = (vrow + (uint32_t)(u0 * row_len)) * (uint32_t)sizeof(float) + lane_off;
        bool out[GL];
#pragma unroll
        for (int m = 0; m < GL; m++) out[m] = fm[m] + 4 > u0 + SPAN;
#pragma unroll
        for (int b = 0; b < SPAN; b++) {
          float d = 0.f;
#pragma unroll
          for (int m = 0; m < GL; m++) {
            const int k = u0 + b - fm[m];
            float a = 0.f;
#pragma unroll
            for (int kk = 0; kk < 4; kk++) a = (k == kk) ? am[m][kk] : a;
            d += out[m] ? 0.f : a * dzm[m];
          }
          __builtin_amdgcn_raw_ptr_buffer_atomic_fadd_f32(-d, rs_e, d != 0.f ? o : OOB,
                                                        b * bin_bytes, 0);
        }
#pragma unroll
        for (int m = 0; m < GL; m++) {
          if (out[m]) {
            const uint32_t om = (vrow + (uint32_t)(fm[m] * row_len)) * (uint32_t)sizeof(float) + lane_off;
#pragma unroll
            for (int b = 0; b < 4; b++) {
              const float d = am[m][b] * dzm[m];
              __builtin_amdgcn_raw_ptr_buffer_atomic_fadd_f32(-d, rs_e, d != 0.f ? om : OOB,
                                                            b * bin_bytes, 0);
            }
          }
        }
      }
    }
    // the next group's loads must see these atomics
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  if (tid < ZL && L.count_stats && vox_mag != 0.f) atomicAdd(&B.stats[0], (double)vox_mag);
  if (tid == 0 && n_vis != 0.0) {
    atomicAdd(&B.stats[1], n_upd);
    atomicAdd(&B.stats[2], n_foot);
    atomicAdd(&B.stats[3], n_vis);
  }
}

template <int ZL, bool PACK, int NT, int GL>
__global__ void __launch_bounds__(NT) k_icd_groups(XtGeom g, XtIcdBuffers B, XtIcdLaunch L) {
  icd_groups_body<ZL, PACK, NT, GL>(g, B, L, (int)blockIdx.x, (int)gridDim.x, (int)blockIdx.y);
}

template <int ZL, bool TAB, bool PACK, int MA = XT_MAXA, int NT = 256>
__global__ void __launch_bounds__(NT) k_icd_lines(XtGeom g, XtIcdBuffers B, XtIcdLaunch L) {
  icd_lines_body<ZL, TAB, PACK, MA, NT>(g, B, L, (int)blockIdx.x, (int)gridDim.x,
                                        (int)blockIdx.y);
}

// One slab of a batch launch: its state, its launch parameters, its workgroups
// per frame gx and its first row y0 of the grid (rows y0 .. y0 + N_time - 1).
struct XtIcdSlab {
  XtGeom g;
  XtIcdBuffers B;
  XtIcdLaunch L;
  int gx;
  int y0;
};

// Every slab of a batch in one launch: grid = (max gx, sum of frames); the
// workgroups beyond a slab's own gx return at once.  Slabs are listed by
// increasing y0.
template <int ZL, bool TAB, bool PACK, int NT = 256>
__global__ void __launch_bounds__(NT) k_icd_lines_multi(const XtIcdSlab *__restrict__ slabs,
                                                        int n) {
  int i = 0;
  while (i + 1 < n && (int)blockIdx.y >= slabs[i + 1].y0) i++;
  const int gx = slabs[i].gx;
  if ((int)blockIdx.x >= gx) return;
  const XtGeom g = slabs[i].g;
  const XtIcdBuffers B = slabs[i].B;
  const XtIcdLaunch L = slabs[i].L;
  icd_lines_body<ZL, TAB, PACK, XT_MAXA, NT>(g, B, L, (int)blockIdx.x, gx,
                                             (int)blockIdx.y - slabs[i].y0);
}

// Line selection per (frame, z-block parity): the update magnitude summed over
// the parity's blocks, magpar[f][par][c]; after the selection on it, every block
// of the parity takes the line's flag: blk[f][b][c] = flag[f][b & 1][c].
__global__ void k_mag_parity(int nb, size_t NYX, int n_time, const float *__restrict__ mag,
                             float *__restrict__ magpar) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NYX * n_time) return;
  const size_t f = i / NYX, c = i % NYX;
  float a[2] = {0.f, 0.f};
  for (int b = 0; b < nb; b++) a[b & 1] += mag[(f * nb + b) * NYX + c];
  magpar[(f * 2 + 0) * NYX + c] = a[0];
  magpar[(f * 2 + 1) * NYX + c] = a[1];
}

__global__ void k_flag_blocks(int nb, size_t NYX, int n_time, const uint8_t *__restrict__ flag,
                              uint8_t *__restrict__ blk) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NYX * n_time) return;
  const size_t f = i / NYX, c = i % NYX;
  for (int b = 0; b < nb; b++) blk[(f * nb + b) * NYX + c] = flag[(f * 2 + (b & 1)) * NYX + c];
}

// NHICD line list of each (frame, parity) map: the selected lines inside the
// mask, in raster order, and their number.  The map is cut into chunks of
// XT_COMPACT_CHUNK lines, one 256-thread workgroup each (grid = (chunks, maps)):
// k_nhicd_chunk_count counts every chunk's lines into cnt[map][chunk], then
// k_nhicd_chunk_write places a chunk's lines after those of the earlier chunks
// and its last chunk stores the map's total.
#define XT_COMPACT_CHUNK 4096

__device__ __forceinline__ int xt_block256_sum(int v, int *wsum) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  return wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ void __launch_bounds__(256) k_nhicd_chunk_count(int NYX, int nch,
                                                           const uint8_t *__restrict__ flag,
                                                           const uint8_t *__restrict__ mask,
                                                           int *__restrict__ cnt) {
  __shared__ int wsum[4];
  const int ch = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
  const uint8_t *fl = flag + (size_t)f * NYX;
  const int c0 = ch * XT_COMPACT_CHUNK;
  int n = 0;
  for (int k = 0; k < XT_COMPACT_CHUNK / 256; k++) {
    const int c = c0 + k * 256 + tid;
    n += (c < NYX && fl[c] && mask[c]) ? 1 : 0;
  }
  n = xt_block256_sum(n, wsum);
  if (tid == 0) cnt[(size_t)f * nch + ch] = n;
}

__global__ void __launch_bounds__(256) k_nhicd_chunk_write(int NYX, int nch,
                                                           const uint8_t *__restrict__ flag,
                                                           const uint8_t *__restrict__ mask,
                                                           const int *__restrict__ cnt,
                                                           int *__restrict__ list,
                                                           int *__restrict__ count) {
  __shared__ int wsum[4];
  const int ch = blockIdx.x, f = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int *cf = cnt + (size_t)f * nch;
  int b = 0;
  for (int j = tid; j < ch; j += 256) b += cf[j];
  int base = xt_block256_sum(b, wsum);
  const uint8_t *fl = flag + (size_t)f * NYX;
  int *out = list + (size_t)f * NYX;
  const int c0 = ch * XT_COMPACT_CHUNK;
  for (int k = 0; k < XT_COMPACT_CHUNK / 256; k++) {
    const int c = c0 + k * 256 + tid;
    const bool on = c < NYX && fl[c] && mask[c];
    const unsigned long long bal = __ballot(on);
    __syncthreads();  // every wave has read wsum of the previous step
    if (lane == 0) wsum[wid] = __popcll(bal);
    __syncthreads();
    int off = base;
    for (int q = 0; q < wid; q++) off += wsum[q];
    if (on) out[off + __popcll(bal & ((1ull << lane) - 1ull))] = c;
    base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
  }
  if (ch == nch - 1 && tid == 0) count[f] = base;
}

// ---------------------------------------------------------------------------
// Ring-offset update, F/XT_ICD_update.c:666-759.  The view sum of every
// detector pixel is split over view chunks (enough workgroups to fill the
// chip), then combined, then the zero-mean constraint gamma is formed, then
// the correction is applied to d and to e.  Sums are FP64.
// One sinogram entry's terms of the offset step's numerator and denominator,
// F/XT_ICD_update.c:703-717 (ProjSelect is the threshold test).
__device__ __forceinline__ void offset_accumulate(const XtGeom &g, float ev, float wv, double &num,
                                                  double &den) {
  if (ev * ev * wv < g.T * g.T) {
    num += (double)ev * (double)wv;
    den += (double)wv;
  } else {
    const double temp = (double)g.T * (double)g.delta * sqrt((double)wv);
    const double sign = (ev > 0.f) - (ev < 0.f);
    num += temp * sign;
    den += temp / fabs((double)ev);
  }
}

__global__ void k_offset_partial(XtGeom g, const float *__restrict__ e, const float *__restrict__ w,
                                 int chunk_views, double *__restrict__ part /*[nch][2][NRT]*/) {
  const int NRT = g.N_r * g.N_t;
  const int rt = blockIdx.x * blockDim.x + threadIdx.x;
  const int ch = blockIdx.y;
  if (rt >= NRT) return;
  const int p0 = ch * chunk_views, p1 = min(g.N_p, p0 + chunk_views);
  double num = 0.0, den = 0.0;
  const int r = rt / g.N_t, t = rt - (rt / g.N_t) * g.N_t;
  // views in groups of 8 (16 measured slower): the group's loads are issued
  // before its sums, which keep the view order
  int p = p0;
  for (; p + 8 <= p1; p += 8) {
    float ev[8], wv[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const size_t idx = xt_sidx(g, p + k, r, t);
      ev[k] = e[idx];
      wv[k] = w[idx];
    }
#pragma unroll
    for (int k = 0; k < 8; k++) offset_accumulate(g, ev[k], wv[k], num, den);
  }
  for (; p < p1; p++) {
    const size_t idx = xt_sidx(g, p, r, t);
    offset_accumulate(g, e[idx], w[idx], num, den);
  }
  part[((size_t)ch * 2 + 0) * NRT + rt] = num;
  part[((size_t)ch * 2 + 1) * NRT + rt] = den;
}

__global__ void k_offset_combine(XtGeom g, int nch, const double *__restrict__ part,
                                 const float *__restrict__ d, double *__restrict__ numden /*[2][NRT]*/,
                                 double *__restrict__ wg_sums /*[nblk][2]*/) {
  __shared__ double sa[256], sb[256];
  const int NRT = g.N_r * g.N_t;
  const int rt = blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0, b = 0.0;
  if (rt < NRT) {
    double num = 0.0, den = 0.0;
    for (int ch = 0; ch < nch; ch++) {
      num += part[((size_t)ch * 2 + 0) * NRT + rt];
      den += part[((size_t)ch * 2 + 1) * NRT + rt];
    }
    numden[rt] = num;
    numden[NRT + rt] = den;
    a = (double)d[rt] + num / den;
    b = 1.0 / den;
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    wg_sums[blockIdx.x * 2 + 0] = sa[0];
    wg_sums[blockIdx.x * 2 + 1] = sb[0];
  }
}

__global__ void k_offset_gamma(int nblk, const double *__restrict__ wg_sums, int zero_mean,
                               double *__restrict__ gamma) {
  __shared__ double sa[256], sb[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
    a += wg_sums[i * 2];
    b += wg_sums[i * 2 + 1];
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *gamma = zero_mean ? sa[0] / sb[0] : 0.0;
}

__global__ void k_offset_apply(XtGeom g, int chunk_views, const double *__restrict__ numden,
                               const double *__restrict__ gamma, float *__restrict__ d,
                               float *__restrict__ e) {
  const int NRT = g.N_r * g.N_t;
  const int rt = blockIdx.x * blockDim.x + threadIdx.x;
  const int ch = blockIdx.y;
  if (rt >= NRT) return;
  const float step = (float)((numden[rt] - *gamma) / numden[NRT + rt]);
  if (ch == 0) d[rt] = d[rt] + step;
  const int p0 = ch * chunk_views, p1 = min(g.N_p, p0 + chunk_views);
  const int r = rt / g.N_t, t = rt - (rt / g.N_t) * g.N_t;
  int p = p0;
  for (; p + 8 <= p1; p += 8) {
    float ev[8];
#pragma unroll
    for (int k = 0; k < 8; k++) ev[k] = e[xt_sidx(g, p + k, r, t)];
#pragma unroll
    for (int k = 0; k < 8; k++) e[xt_sidx(g, p + k, r, t)] = ev[k] - step;
  }
  for (; p < p1; p++) e[xt_sidx(g, p, r, t)] -= step;
}

// ---------------------------------------------------------------------------
// Variance re-estimate, F/XT_ICD_update.c:633-664.
__global__ void k_variance_partial(XtGeom g, const float *__restrict__ e, const float *__restrict__ w,
                                   float var_old, size_t n, double *__restrict__ wg) {
  __shared__ double sa[256];
  double a = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float ev = e[i], wv = w[i];
    const float wr = wv * var_old;  // Weight restored to 1/variance units
    if (ev * ev * wv < g.T * g.T)
      a += (double)(ev * ev * wr);
    else
      a += (double)fabsf(ev) * (double)g.delta * (double)g.T * sqrt((double)wr * (double)var_old);
  }
  sa[threadIdx.x] = a;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sa[threadIdx.x] += sa[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) wg[blockIdx.x] = sa[0];
}

__global__ void k_variance_apply(float *__restrict__ w, size_t n, float var_old, float var_new) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    w[i] = (w[i] * var_old) / var_new;
}

// ---------------------------------------------------------------------------
// Generic FP64 block reduction of a float array (|x| optional) into wg[blockIdx].
__global__ void k_sum_f32(const float *__restrict__ a, size_t n, int absval, double *__restrict__ wg) {
  __shared__ double sa[256];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    s += absval ? fabs((double)a[i]) : (double)a[i];
  sa[threadIdx.x] = s;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sa[threadIdx.x] += sa[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) wg[blockIdx.x] = sa[0];
}

__global__ void k_sum_f64(const double *__restrict__ a, int n, double *__restrict__ out) {
  __shared__ double sa[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += a[i];
  sa[threadIdx.x] = s;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sa[threadIdx.x] += sa[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sa[0];
}

// ---------------------------------------------------------------------------
// MAP cost in FP64, F/XT_ICD_update.c:181-314: data term over the sinogram and
// the prior over the 13 causal neighbour pairs plus the next frame.
__global__ void k_cost_data(XtGeom g, const float *__restrict__ e, const float *__restrict__ w,
                            size_t n, double *__restrict__ wg) {
  __shared__ double sa[256];
  double s = 0.0;
  const double T = g.T, dl = g.delta;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double u = (double)e[i] * sqrt((double)w[i]);
    const bool sel = e[i] * e[i] * w[i] < g.T * g.T;
    s += sel ? u * u : 2.0 * dl * T * fabs(u) + T * T * (1.0 - 2.0 * dl);
  }
  sa[threadIdx.x] = s;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sa[threadIdx.x] += sa[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) wg[blockIdx.x] = sa[0] * 0.5;
}

__device__ __forceinline__ double xt_qggmrf_value(double delta, double p, double c, double sq,
                                                  double sqp) {
  const double ad = fabs(delta);
  return (ad * ad / sq) / (c + pow(ad, 2.0 - p) / sqp);
}

__global__ void k_cost_prior(XtGeom g, const float *__restrict__ x, double *__restrict__ wg) {
  __shared__ double sa[256];
  const size_t NYX = (size_t)g.N_y * g.N_x;
  const size_t n = (size_t)g.N_time * NYX * g.N_z;
  double s = 0.0;
  const double P = g.MRF_P;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % g.N_z);
    const size_t col = (i / g.N_z) % NYX;
    const int f = (int)(i / ((size_t)g.N_z * NYX));
    const int yy = (int)(col / g.N_x), xx = (int)(col % g.N_x);
    const double cv = x[i];
    // causal neighbours: (dz,dy,dx) with dz=+1 (all 9), or dz=0 and (dy=+1 or dy=0,dx=+1)
    for (int dz = 0; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          if (dz == 0 && (dy < 0 || (dy == 0 && dx <= 0))) continue;
          const int z2 = z + dz, y2 = yy + dy, x2 = xx + dx;
          if (z2 >= g.N_z || y2 < 0 || y2 >= g.N_y || x2 < 0 || x2 >= g.N_x) continue;
          const double nv = x[((size_t)f * NYX + (size_t)y2 * g.N_x + x2) * g.N_z + z2];
          const int l = ((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1);
          s += (double)g.filt[l] *
               xt_qggmrf_value(cv - nv, P, g.C_S, g.Sigma_S_Q, g.Sigma_S_Q_P);
        }
    if (f + 1 < g.N_time) {
      const double nv = x[i + NYX * g.N_z];
      s += (double)g.tfilt * xt_qggmrf_value(cv - nv, P, g.C_T, g.Sigma_T_Q, g.Sigma_T_Q_P);
    }
  }
  sa[threadIdx.x] = s;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sa[threadIdx.x] += sa[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) wg[blockIdx.x] = sa[0];
}

// ---------------------------------------------------------------------------
// NHICD selection, F/XT_NHICD.c:128-216: 5x5 Hamming filter of each magnitude
// map (float, reference order), then the order statistic of rank
// N - n_select by radix select on the (non-negative) float bit patterns, then
// the lines strictly above it.  One map per (blockIdx.y).
__global__ void k_vsc_filter(int NY, int NX, const float *__restrict__ mag, float *__restrict__ filt) {
#pragma clang fp contract(off)
  const float Hm[5][5] = {{0.0064f, 0.0432f, 0.0800f, 0.0432f, 0.0064f},
                          {0.0432f, 0.2916f, 0.5400f, 0.2916f, 0.0432f},
                          {0.0800f, 0.5400f, 1.0000f, 0.5400f, 0.0800f},
                          {0.0432f, 0.2916f, 0.5400f, 0.2916f, 0.0432f},
                          {0.0064f, 0.0432f, 0.0800f, 0.0432f, 0.0064f}};
  const size_t NYX = (size_t)NY * NX;
  const int m = blockIdx.y;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int)NYX) return;
  const int i = idx / NX, j = idx % NX;
  const float *M = mag + (size_t)m * NYX;
  float acc = 0;
  for (int p = -2; p <= 2; p++)
    for (int q = -2; q <= 2; q++)
      if (i + p >= 0 && i + p < NY && j + q >= 0 && j + q < NX)
        acc += Hm[p + 2][q + 2] * M[(size_t)(i + p) * NX + (j + q)];
  filt[(size_t)m * NYX + idx] = acc;
}

// NHICD threshold per map m (F/XT_NHICD.c:178-205): the order statistic of rank
// NYX - n_select (1-indexed from the smallest) of the filtered map, found by a
// radix selection on the float bits (non-negative floats order as unsigned
// integers), 8 bits a pass; lines strictly above it are selected.  The passes run
// over many workgroups: state st[m] = {prefix, mask, rank left}, histogram
// hist[m][256] (cleared by k_vsc_init and by each pick, so a captured graph can
// replay them).
__global__ void k_vsc_init(int NYX, const int *__restrict__ n_select, unsigned *__restrict__ st,
                           unsigned *__restrict__ hist, int *__restrict__ n_out) {
  const int m = blockIdx.x;
  hist[(size_t)m * 256 + threadIdx.x] = 0u;
  if (threadIdx.x == 0) {
    int target = NYX - n_select[m] - 1;  // 0-indexed
    if (target < 0) target = 0;
    st[m * 4 + 0] = 0u;
    st[m * 4 + 1] = 0u;
    st[m * 4 + 2] = (unsigned)target;
    n_out[m] = 0;
  }
}

__global__ void k_vsc_hist(int NYX, const float *__restrict__ filt, const unsigned *__restrict__ st,
                           unsigned *__restrict__ hist, int shift) {
  __shared__ unsigned h[256];
  const int m = blockIdx.y;
  h[threadIdx.x] = 0u;
  __syncthreads();
  const unsigned pre = st[m * 4 + 0], msk = st[m * 4 + 1];
  const unsigned *K = (const unsigned *)(filt + (size_t)m * NYX);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < NYX; i += gridDim.x * blockDim.x) {
    const unsigned k = K[i];
    if ((k & msk) == pre) atomicAdd(&h[(k >> shift) & 255u], 1u);
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[(size_t)m * 256 + threadIdx.x], h[threadIdx.x]);
}

__global__ void k_vsc_pick(unsigned *__restrict__ st, unsigned *__restrict__ hist, int shift) {
  const int m = blockIdx.x;
  unsigned *H = hist + (size_t)m * 256;
  if (threadIdx.x == 0) {
    const unsigned t = st[m * 4 + 2];
    unsigned cum = 0u;
    int b = 0;
    for (; b < 256; b++) {
      if (cum + H[b] > t) break;
      cum += H[b];
    }
    if (b > 255) b = 255;
    st[m * 4 + 2] = t - cum;
    st[m * 4 + 0] |= (unsigned)b << shift;
    st[m * 4 + 1] |= 255u << shift;
  }
  __syncthreads();
  H[threadIdx.x] = 0u;
}

__global__ void k_vsc_flag(int NYX, const float *__restrict__ filt, const unsigned *__restrict__ st,
                           uint8_t *__restrict__ flag, int *__restrict__ n_out) {
  __shared__ int cnt;
  const int m = blockIdx.y;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const float thresh = __uint_as_float(st[m * 4 + 0]);
  const float *F = filt + (size_t)m * NYX;
  int local = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < NYX; i += gridDim.x * blockDim.x) {
    const uint8_t sel = F[i] > thresh ? 1 : 0;
    flag[(size_t)m * NYX + i] = sel;
    local += sel;
  }
  if (local) atomicAdd(&cnt, local);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(&n_out[m], cnt);
}

// A line is an NHICD line of a z-block parity when a z-block of that parity
// selected it: any[f][par][c].
__global__ void k_vsc_any(int nb, size_t NYX, int n_time, const uint8_t *__restrict__ blk,
                          uint8_t *__restrict__ any) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NYX * n_time) return;
  const size_t f = i / NYX, c = i % NYX;
  uint8_t a[2] = {0, 0};
  for (int b = 0; b < nb; b++) a[b & 1] |= blk[(f * nb + b) * NYX + c];
  any[(f * 2 + 0) * NYX + c] = a[0];
  any[(f * 2 + 1) * NYX + c] = a[1];
}

// Multi-resolution upsampling (initObject with initICD 2/3,
// F/XT_ICD_update.c:316-435): one axis of a [A][n][B] array, n coarse samples
// at stride B, to [A][2n][B].  Along the axis f[0] = c[0], f[1] = (3c0 + c1)/4,
// f[2n-2] = (c[n-2] + 3c[n-1])/4, f[2n-1] = c[n-1], interior f[2k] =
// (c[k-1] + 3c[k])/4, f[2k+1] = (3c[k] + c[k+1])/4; each sum in double and
// rounded to float once, except the interior odd samples of the y pass
// (real_odd), which the reference sums in float (its integer 3, :339,:411).
__device__ __forceinline__ float up_even(float p, float c) {
  return (float)(((double)p + 3.0 * (double)c) / 4.0);
}
__device__ __forceinline__ float up_odd(float c, float n) {
  return (float)((3.0 * (double)c + (double)n) / 4.0);
}
__device__ __forceinline__ float up_odd_real(float c, float n) {
#pragma clang fp contract(off)
  // two roundings, as the reference's float expression (an FMA would round once)
  const float t = 3.f * c;
  return (t + n) * 0.25f;
}

__global__ void k_upsample_axis(const float *__restrict__ in, float *__restrict__ out, int A,
                                int n, int B, int real_odd) {
  const size_t total = (size_t)A * 2 * n * B;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(i % B);
    const int f = (int)((i / B) % (2 * n));
    const size_t a = i / ((size_t)B * 2 * n);
    const float *c = in + a * n * B + b;
    float v;
    if (f == 0) {
      v = c[0];
    } else if (f == 2 * n - 1) {
      v = c[(size_t)(n - 1) * B];
    } else if (f == 1) {
      v = up_odd(c[0], c[B]);
    } else if (f == 2 * n - 2) {
      v = up_even(c[(size_t)(n - 2) * B], c[(size_t)(n - 1) * B]);
    } else {
      const int k = f / 2;
      if ((f & 1) == 0)
        v = up_even(c[(size_t)(k - 1) * B], c[(size_t)k * B]);
      else if (real_odd)
        v = up_odd_real(c[(size_t)k * B], c[(size_t)(k + 1) * B]);
      else
        v = up_odd(c[(size_t)k * B], c[(size_t)(k + 1) * B]);
    }
    out[i] = v;
  }
}

// ---------------------------------------------------------------------------
// Counts to sinogram, gen_projection_4m_HDF's arithmetic (F/XT_HDFIO.c:160-199),
// on the region of interest read from the file.  Flat fields: per detector
// pixel the mean of the white (dark) frames in float, stored [R_in][T_in] as
// the reference; bright[j][k] and weight[i][j][k] sum |white - dark| and
// |counts - dark| over the ratio_r x ratio_t pixels of bin (j, k) in the
// reference's order (r offset outer, t offset inner), in float.
__global__ void k_flat_mean(const uint16_t *__restrict__ frames, int n, int T_in, int R_in,
                            float *__restrict__ out) {
  const size_t plane = (size_t)T_in * R_in;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < plane;
       q += (size_t)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int i = 0; i < n; i++) s += (float)frames[(size_t)i * plane + q];
    const int t = (int)(q / R_in), r = (int)(q % R_in);
    out[(size_t)r * T_in + t] = s / (float)n;
  }
}

__global__ void k_bright(const float *__restrict__ w2, const float *__restrict__ d2, int T_in,
                         int N_r, int N_t, int rr, int rt, float *__restrict__ bright) {
  const size_t n = (size_t)N_r * N_t;
  for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < n;
       o += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(o / N_t), k = (int)(o % N_t);
    float b = 0.f;
    for (int m = 0; m < rr; m++)
      for (int l = 0; l < rt; l++) {
        const size_t q = (size_t)(j * rr + m) * T_in + (k * rt + l);
        b += fabsf(w2[q] - d2[q]);
      }
    bright[o] = b;
  }
}

// projection = BH * temp^2 + temp in double with temp = log(bright / weight)
// rounded to float (the reference's Real_t temp and double macro literal)
__global__ void k_counts_sino(const uint16_t *__restrict__ counts, const float *__restrict__ d2,
                              const float *__restrict__ bright, int N_p, int T_in, int R_in,
                              int N_r, int N_t, int rr, int rt, double bh,
                              float *__restrict__ proj, float *__restrict__ weight) {
#pragma clang fp contract(off)
  const size_t n = (size_t)N_p * N_t * N_r;
  const size_t plane = (size_t)T_in * R_in;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (size_t)gridDim.x * blockDim.x) {
    // bins fastest: neighbouring threads read neighbouring detector pixels
    const int j = (int)(idx % N_r);
    const int k = (int)((idx / N_r) % N_t);
    const int i = (int)(idx / ((size_t)N_r * N_t));
    float w = 0.f;
    for (int m = 0; m < rr; m++)
      for (int l = 0; l < rt; l++) {
        const int r = j * rr + m, t = k * rt + l;
        w += fabsf((float)counts[(size_t)i * plane + (size_t)t * R_in + r] -
                   d2[(size_t)r * T_in + t]);
      }
    const size_t o = ((size_t)i * N_r + j) * N_t + k;
    const float temp = (float)log((double)(bright[(size_t)j * N_t + k] / w));
    weight[o] = w;
    proj[o] = (float)(bh * (double)temp * (double)temp + (double)temp);
  }
}

// ---------------------------------------------------------------------------
// Phase contrast (the reference's -DPHASE_CONTRAST_TOMOGRAPHY build).  A voxel's
// footprint is the detector Laplacian of its absorption footprint trimmed to
// the detector (F/XT_AMatrix.c:113-187), so by linearity A_pc x = Lap * (A x)
// on the detector: mode 0 out64 = Lap * ax, mode 1 e = y - Lap * ax - d.
__global__ void k_pc_laplace(XtGeom g, const double *__restrict__ ax, int mode,
                             double *__restrict__ out64, const float *__restrict__ y,
                             const float *__restrict__ d, float *__restrict__ e) {
  const size_t n = (size_t)g.N_p * g.N_r * g.N_t;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int t = (int)(i % g.N_t), r = (int)((i / g.N_t) % g.N_r);
    const size_t v = i / ((size_t)g.N_r * g.N_t);
    double s = 0.0;
#pragma unroll
    for (int m = -1; m <= 1; m++)
#pragma unroll
      for (int k = -1; k <= 1; k++) {
        const int rr = r + m, tt = t + k;
        if (rr >= 0 && rr < g.N_r && tt >= 0 && tt < g.N_t)
          s += (double)g.lap[(m + 1) * 3 + (k + 1)] * ax[(v * g.N_r + rr) * g.N_t + tt];
      }
    if (mode == 0) {
      out64[i] = s;
    } else {
      const size_t idx = xt_sidx(g, (int)v, r, t);
      e[idx] = (float)((double)y[idx] - s - (double)d[(size_t)r * g.N_t + t]);
    }
  }
}

// Laplacian footprint entry of a column (values vals[0..XT_MAXA) times delta_t,
// vals[k] at bin first + k) at bin first + rr, for a row whose Laplacian weights
// are l0, l1, l2 (Lap[m + 1][1 - dt] for m = -1, 0, 1, dt the row's offset from
// the voxel's): the sum in the reference's order and float steps
// (compute_LapMatrix_4m_AMatrix; zero entries of the line add nothing).  rr is
// known at compile time where it is called, so vals stays in registers.
__device__ __forceinline__ float xt_pc_entry(const float (&vals)[XT_MAXA], int rr, float l0,
                                             float l1, float l2) {
#pragma clang fp contract(off)
  const float lm = (rr - 1 >= 0 && rr - 1 < XT_MAXA) ? vals[rr - 1] : 0.f;
  const float lc = (rr >= 0 && rr < XT_MAXA) ? vals[rr] : 0.f;
  const float lp = (rr + 1 >= 0 && rr + 1 < XT_MAXA) ? vals[rr + 1] : 0.f;
  float acc = 0.f;
  acc += l0 * lm;
  acc += l1 * lc;
  acc += l2 * lp;
  return acc;
}

// Phase-contrast voxel update (F/XT_VoxUpdate.c:590-845).  A slice's footprint
// covers detector rows t-1..t+1, so slices three apart have disjoint footprints
// and no prior term between them: a line's slices are updated in three phases
// (z mod 3 = 0, 1, 2; or, for the per-voxel checks, one slice per phase in
// ascending order, the reference's slice loop), the slices of a phase at once.
// Lanes run over detector rows: lane t of a view group serves the phase's slice
// among t-1..t+1, so a wave reads and updates whole rows of e.  Theta is summed
// over the view groups in LDS and over the slice's three rows by the slice's own
// thread, which takes the qGGMRF step; then e -= A * delta and a wait before the
// next phase reads e.  Lines, the z-block parity and NHICD lists as k_icd_lines.
template <int ZL>
__global__ void __launch_bounds__(256) k_icd_lines_pc(XtGeom g, XtIcdBuffers B, XtIcdLaunch L) {
  constexpr int NT = 256, RL = ZL, G = NT / RL;  // row lanes of a view group, view groups
  extern __shared__ float4 smem4[];
  float4 *a_v4 = smem4;                           // [max_views] bins 0..3 (x delta_t)
  float2 *a_v2 = (float2 *)(a_v4 + g.max_views);  // [max_views] bins 4..5
  int *a_first = (int *)(a_v2 + g.max_views);     // [max_views] first candidate bin
  float *red1 = (float *)(a_first + g.max_views);  // [G][RL]
  float *red2 = red1 + NT;                         // [G][RL]
  float *xn = red2 + NT;                           // [11][ZL]
  float *delta = xn + 11 * ZL;                     // [ZL]
  float *magb = delta + ZL;                        // [64]
  __shared__ int s_cnt;
  const int tid = threadIdx.x, grp = tid / RL, lt = tid % RL;
  const int f = blockIdx.y;
  const int nv = B.proj_num[f];
  const int *vlist = B.views + (size_t)f * g.max_views;
  const size_t NYX = (size_t)g.N_y * g.N_x;
  const int iter = *B.iter;
  const int blk_sz = g.N_z / g.nb;
  const int par = (L.mode == 2) ? -1 : ((L.half ^ f) & 1);
  auto zblock = [&](int z) {
    int b = blk_sz > 0 ? z / blk_sz : g.nb - 1;
    return b >= g.nb ? g.nb - 1 : b;
  };
  int lo = 0, hi = 0, n_dom = 0;
  const int *list = B.lines;
  uint32_t key = 0;
  if (L.mode == 0) {
    const int n = B.n_lines, per = n / L.K;
    lo = L.sub_iter * per;
    hi = (L.sub_iter == L.K - 1) ? n : (L.sub_iter + 1) * per;
    n_dom = n;
    key = xt_order_key(L.seed, iter, f, par);
  } else if (L.mode == 1) {
    hi = B.nhicd_count[f * 2 + par];
    list = B.nhicd_list + ((size_t)f * 2 + par) * NYX;
  } else {
    hi = B.n_lines;
  }
  int i0 = lo + (int)blockIdx.x, i1 = hi, istep = (int)gridDim.x;
  if (L.mode == 1) {
    const long long len = hi - lo;
    i0 = lo + (int)(len * blockIdx.x / gridDim.x);
    i1 = lo + (int)(len * (blockIdx.x + 1) / gridDim.x);
    istep = 1;
  }
  const float T2 = g.T * g.T, Td = g.T * g.delta;
  // the Laplacian weights of the lane's row for each row offset dt = -1, 0, 1
  // come from column 1 - dt of Lap_Kernel
  const int nph = L.zorder ? (g.N_z > 3 ? g.N_z : 3) : 3;
  float vox_mag = 0.f;
  double n_upd = 0.0, n_foot = 0.0, n_vis = 0.0;
  for (int i = i0; i < i1; i += istep) {
    const int c = list[n_dom ? xt_permute(i, n_dom, key) : i];
    const int yy = c / g.N_x, xx = c - yy * g.N_x;
    // a slice is updated when it is in the launch's parity, its block selected
    // the line (NHICD) and the edge mask (if any) holds it
    auto slice_on = [&](int z) {
      bool upd = par < 0 || (zblock(z) & 1) == par;
      if (upd && L.mode == 1) upd = B.nhicd_blk[((size_t)f * g.nb + zblock(z)) * NYX + c] != 0;
      if (upd && B.vmask) upd = B.vmask[((size_t)f * NYX + c) * g.N_z + z] != 0;
      return upd;
    };
    for (int p = tid; p < nv; p += NT) {
      const int v = vlist[p];
      int first = 0;
      float vals[XT_MAXA];
      xt_amatrix_column(g, B.Hr + (size_t)v * XT_HR_STRIDE, B.cosv[v], B.sinv[v], yy, xx, &first,
                        vals);
      a_v4[p] = make_float4(vals[0] * g.delta_t, vals[1] * g.delta_t, vals[2] * g.delta_t,
                            vals[3] * g.delta_t);
      a_v2[p] = make_float2(vals[4] * g.delta_t, vals[5] * g.delta_t);
      a_first[p] = first;
    }
    for (int q = tid; q < 11 * ZL; q += NT) {
      const int k = q / ZL, z = q - k * ZL;
      float val = 0.f;
      if (z < g.N_z) {
        if (k < 9) {
          const int y2 = yy + k / 3 - 1, x2 = xx + k % 3 - 1;
          if (y2 >= 0 && y2 < g.N_y && x2 >= 0 && x2 < g.N_x)
            val = __hip_atomic_load(&B.x[((size_t)f * NYX + (size_t)y2 * g.N_x + x2) * g.N_z + z],
                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          const int f2 = f + (k == 9 ? -1 : 1);
          if (f2 >= 0 && f2 < g.N_time)
            val = __hip_atomic_load(&B.x[((size_t)f2 * NYX + c) * g.N_z + z], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      xn[q] = val;
    }
    if (tid < 64) magb[tid] = 0.f;
    if (tid < ZL) delta[tid] = 0.f;
    if (tid == 0) s_cnt = 0;
    __syncthreads();
    float *xcol = xn + 4 * ZL;
    for (int ph = 0; ph < nph; ph++) {
      // the lane's slice in this phase: the one of rows lt-1..lt+1 whose index
      // is ph modulo the phase count (at most one), and the row's offset dt
      int zt = -1, dt = 0;
#pragma unroll
      for (int q = -1; q <= 1; q++) {
        const int z = lt + q;
        if (z >= 0 && z < g.N_z && z % nph == ph) {
          zt = z;
          dt = -q;
        }
      }
      const bool act = zt >= 0 && lt < g.N_t && slice_on(zt);
      const int col = 1 - dt;
      const float l0 = col == 0 ? g.lap[0] : (col == 1 ? g.lap[1] : g.lap[2]);
      const float l1 = col == 0 ? g.lap[3] : (col == 1 ? g.lap[4] : g.lap[5]);
      const float l2 = col == 0 ? g.lap[6] : (col == 1 ? g.lap[7] : g.lap[8]);
      float th1 = 0.f, th2 = 0.f;
      int cnt = 0;
      if (act) {
        for (int p = grp; p < nv; p += G) {
          const int v = vlist[p];
          const float4 a4 = a_v4[p];
          const float2 a2 = a_v2[p];
          const float vals[XT_MAXA] = {a4.x, a4.y, a4.z, a4.w, a2.x, a2.y};
          const int first = a_first[p];
#pragma unroll
          for (int rr = -1; rr <= XT_MAXA; rr++) {
            const int r = first + rr;
            if (r < 0 || r >= g.N_r) continue;
            const float pe = xt_pc_entry(vals, rr, l0, l1, l2);
            if (pe == 0.f) continue;
            const size_t idx = xt_sidx(g, v, r, lt);
            const float e1 = B.e[idx], w1 = B.w[idx];
            const float wr = Td * __builtin_amdgcn_sqrtf(w1) * __builtin_amdgcn_rcpf(fabsf(e1));
            const float ws = (e1 * e1 * w1 < T2) ? w1 : wr;
            const float pw = pe * ws;
            th2 += pe * pw;
            th1 -= e1 * pw;
            cnt++;
          }
        }
      }
      red1[tid] = th1;
      red2[tid] = th2;
      if (cnt) atomicAdd(&s_cnt, cnt);
      __syncthreads();
      // the phase's slices, one thread each: theta over its three rows and the
      // view groups, zero skipping (F/XT_VoxUpdate.c:805-826), the qGGMRF step
      if (tid < g.N_z && tid % nph == ph && slice_on(tid)) {
        const int z = tid;
        float T1 = 0.f, T2z = 0.f;
        for (int q = -1; q <= 1; q++) {
          const int t = z + q;
          if (t < 0 || t >= g.N_t) continue;
          for (int gg = 0; gg < G; gg++) {
            T1 += red1[gg * RL + t];
            T2z += red2[gg * RL + t];
          }
        }
        const float V = xcol[z];
        bool any_pos = false;
        float s1 = 0.f, s2 = 0.f;
        for (int dz = -1; dz <= 1; dz++)
          for (int k = 0; k < 9; k++) {
            const int l = (dz + 1) * 9 + k;
            const int z2 = z + dz, y2 = yy + k / 3 - 1, x2 = xx + k % 3 - 1;
            if (l == 13 || z2 < 0 || z2 >= g.N_z || y2 < 0 || y2 >= g.N_y || x2 < 0 || x2 >= g.N_x)
              continue;
            const float nb = xn[k * ZL + z2];
            any_pos |= nb > 0.f;
            const float bq =
                g.filt[l] * xt_qggmrf_ratio(V - nb, g.MRF_P, g.C_S, g.Sigma_S_Q, g.Sigma_S_Q_P);
            s1 += bq * nb;
            s2 += bq;
          }
        if (f > 0) {
          const float tn = xn[9 * ZL + z];
          any_pos |= tn > 0.f;
          const float bq = g.tfilt * xt_qggmrf_ratio(V - tn, g.MRF_P, g.C_T, g.Sigma_T_Q, g.Sigma_T_Q_P);
          s1 += bq * tn;
          s2 += bq;
        }
        if (f + 1 < g.N_time) {
          const float tn = xn[10 * ZL + z];
          any_pos |= tn > 0.f;
          const float bq = g.tfilt * xt_qggmrf_ratio(V - tn, g.MRF_P, g.C_T, g.Sigma_T_Q, g.Sigma_T_Q_P);
          s1 += bq * tn;
          s2 += bq;
        }
        if (!((V == 0.f) && (iter > 1) && !any_pos)) {
          const float u = (s1 + (T2z * V) - T1) / (s2 + T2z);
          float nv2 = V + g.alpha * (u - V);
          if (g.positivity && nv2 <= 0.f) nv2 = 0.f;
          xcol[z] = nv2;
          delta[z] = nv2 - V;
          atomicAdd(&magb[zblock(z)], fabsf(nv2 - V));
          vox_mag += fabsf(nv2);
          n_upd += 1.0;
        }
      }
      __syncthreads();
      // e -= A * delta over the lane's row
      const float dl = act ? delta[zt] : 0.f;
      if (dl != 0.f) {
        for (int p = grp; p < nv; p += G) {
          const int v = vlist[p];
          const float4 a4 = a_v4[p];
          const float2 a2 = a_v2[p];
          const float vals[XT_MAXA] = {a4.x, a4.y, a4.z, a4.w, a2.x, a2.y};
          const int first = a_first[p];
#pragma unroll
          for (int rr = -1; rr <= XT_MAXA; rr++) {
            const int r = first + rr;
            if (r < 0 || r >= g.N_r) continue;
            const float pe = xt_pc_entry(vals, rr, l0, l1, l2);
            if (pe != 0.f) atomicAdd(&B.e[xt_sidx(g, v, r, lt)], -(pe * dl));
          }
        }
      }
      // the next phase's theta reads these e entries
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    // ---- write the line back and its update magnitude (as k_icd_lines)
    if (tid < g.N_z)
      __hip_atomic_store(&B.x[((size_t)f * NYX + c) * g.N_z + tid], xcol[tid], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    if (tid < g.nb) {
      const size_t mi = ((size_t)f * g.nb + tid) * NYX + c;
      if ((par < 0 || (tid & 1) == par) && (L.mode != 1 || B.nhicd_blk[mi] != 0))
        B.mag_blk[mi] = magb[tid];
      else
        magb[tid] = B.mag_blk[mi];
    }
    __syncthreads();
    if (tid == 0) {
      float sm = 0.f;
      for (int b = 0; b < g.nb; b++) sm += magb[b];
      B.magsum[(size_t)f * NYX + c] = sm;
      n_vis += 1.0;
      n_foot += (double)s_cnt;
    }
    __syncthreads();
  }
  if (L.count_stats && vox_mag != 0.f) atomicAdd(&B.stats[0], (double)vox_mag);
  if (n_upd != 0.0) atomicAdd(&B.stats[1], n_upd);
  if (tid == 0 && n_vis != 0.0) {
    atomicAdd(&B.stats[2], n_foot);
    atomicAdd(&B.stats[3], n_vis);
  }
}

// ---------------------------------------------------------------------------
// Edge-only updates (F/XT_ImageProc.c:40-149): the voxel update mask of the
// current object, per frame and slice: a 5x5 Gaussian blur, the 3x3 gradient
// pair, |gradient| > 0.0002 inside the object circle, then a 5x5 dilation.
// Each sum runs in the reference's order in float with contraction off, so the
// mask is bit-exact.  Threads run over the device layout [N_time][N_y][N_x][N_z]
// (slices fastest: neighbouring threads read neighbouring words).
struct XtEdgeParams {
  float hblur[25];  // normalised blur weights [k + 2][l + 2], row offset k, column offset l
  float half_x, half_y;  // -Length_X / 2, -Length_Y / 2
  float delta_xy;
  float r2;  // radius_obj * radius_obj
};

__global__ void k_edge_blur(XtGeom g, XtEdgeParams P, const float *__restrict__ x,
                            float *__restrict__ blur) {
#pragma clang fp contract(off)
  const size_t NYX = (size_t)g.N_y * g.N_x, n = (size_t)g.N_time * NYX * g.N_z;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % g.N_z);
    const size_t fc = i / g.N_z;
    const int c = (int)(fc % NYX), row = c / g.N_x, col = c - row * g.N_x;
    const size_t base = (fc - c) * g.N_z;  // frame start
    float acc = 0.f;
    for (int k = -2; k <= 2; k++) {
      const int r2 = row + k;
      if (r2 < 0 || r2 >= g.N_y) continue;
      for (int l = -2; l <= 2; l++) {
        const int c2 = col + l;
        if (c2 < 0 || c2 >= g.N_x) continue;
        acc += P.hblur[(k + 2) * 5 + (l + 2)] * x[base + ((size_t)r2 * g.N_x + c2) * g.N_z + z];
      }
    }
    blur[i] = acc;
  }
}

__global__ void k_edge_gradient(XtGeom g, XtEdgeParams P, const float *__restrict__ blur,
                                uint8_t *__restrict__ edge) {
#pragma clang fp contract(off)
  const float hx[9] = {3.f, 0.f, -3.f, 10.f, 0.f, -10.f, 3.f, 0.f, -3.f};
  const float hy[9] = {3.f, 10.f, 3.f, 0.f, 0.f, 0.f, -3.f, -10.f, -3.f};
  const size_t NYX = (size_t)g.N_y * g.N_x, n = (size_t)g.N_time * NYX * g.N_z;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % g.N_z);
    const size_t fc = i / g.N_z;
    const int c = (int)(fc % NYX), row = c / g.N_x, col = c - row * g.N_x;
    const size_t base = (fc - c) * g.N_z;
    float gx = 0.f, gy = 0.f;
    for (int k = -1; k <= 1; k++) {
      const int r2 = row + k;
      if (r2 < 0 || r2 >= g.N_y) continue;
      for (int l = -1; l <= 1; l++) {
        const int c2 = col + l;
        if (c2 < 0 || c2 >= g.N_x) continue;
        const float v = blur[base + ((size_t)r2 * g.N_x + c2) * g.N_z + z];
        gx += hx[(k + 1) * 3 + (l + 1)] * v;
        gy += hy[(k + 1) * 3 + (l + 1)] * v;
      }
    }
    // the circle of F/XT_ImageProc.c:122-124: centre from -Length / 2, double sums
    const float xc = (float)((double)P.half_x + ((double)(float)col + 0.5) * (double)P.delta_xy);
    const float yc = (float)((double)P.half_y + ((double)(float)row + 0.5) * (double)P.delta_xy);
    uint8_t on = 0;
    if (xc * xc + yc * yc < P.r2) {
      const float mag = (float)sqrt((double)(gx * gx + gy * gy));
      on = mag > 0.0002f;
    }
    edge[i] = on;
  }
}

__global__ void k_edge_dilate(XtGeom g, const uint8_t *__restrict__ edge,
                              uint8_t *__restrict__ mask) {
  const size_t NYX = (size_t)g.N_y * g.N_x, n = (size_t)g.N_time * NYX * g.N_z;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % g.N_z);
    const size_t fc = i / g.N_z;
    const int c = (int)(fc % NYX), row = c / g.N_x, col = c - row * g.N_x;
    const size_t base = (fc - c) * g.N_z;
    uint8_t v = 0;
    for (int k = -2; k <= 2; k++) {
      const int r2 = row + k;
      if (r2 < 0 || r2 >= g.N_y) continue;
      for (int l = -2; l <= 2; l++) {
        const int c2 = col + l;
        if (c2 >= 0 && c2 < g.N_x) v |= edge[base + ((size_t)r2 * g.N_x + c2) * g.N_z + z];
      }
    }
    mask[i] = v;
  }
}

// x := max(x, 0) (the positivity constraint applied to a starting volume)
__global__ void k_clamp_nonneg(float *__restrict__ x, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] = fmaxf(x[i], 0.f);
}

// ---------------------------------------------------------------------------
// Layout changes between the host's [N_time][N_z(+2)][N_y][N_x] object and the
// device's [N_time][N_y][N_x][N_z].
__global__ void k_object_to_device(XtGeom g, const float *__restrict__ host_layout /*[T][Z][Y][X]*/,
                                   float *__restrict__ x) {
  const size_t NYX = (size_t)g.N_y * g.N_x, n = (size_t)g.N_time * NYX * g.N_z;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t col = i % NYX;
    const int z = (int)((i / NYX) % g.N_z);
    const size_t f = i / (NYX * g.N_z);
    x[(f * NYX + col) * g.N_z + z] = host_layout[i];
  }
}

__global__ void k_object_to_host(XtGeom g, const float *__restrict__ x,
                                 float *__restrict__ host_layout) {
  const size_t NYX = (size_t)g.N_y * g.N_x, n = (size_t)g.N_time * NYX * g.N_z;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t col = i % NYX;
    const int z = (int)((i / NYX) % g.N_z);
    const size_t f = i / (NYX * g.N_z);
    host_layout[i] = x[(f * NYX + col) * g.N_z + z];
  }
}

// ProjSelect from the final e and w (F/XT_ICD_update.c:577).
// ProjSelect in the host's [N_p][N_r][N_t] order.
__global__ void k_select(XtGeom g, const float *__restrict__ e, const float *__restrict__ w, size_t n,
                         uint8_t *__restrict__ sel) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int t = (int)(i % g.N_t), r = (int)((i / g.N_t) % g.N_r), p = (int)(i / ((size_t)g.N_t * g.N_r));
    const size_t k = xt_sidx(g, p, r, t);
    sel[i] = (fabs((double)e[k] * sqrt((double)w[k])) < (double)g.T) ? 1 : 0;
  }
}

// Between the host's [N_p][N_r][N_t] order and the HBM layout.
__global__ void k_sino_layout(XtGeom g, const float *__restrict__ src, float *__restrict__ dst,
                              size_t n, int to_hbm) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int t = (int)(i % g.N_t), r = (int)((i / g.N_t) % g.N_r), p = (int)(i / ((size_t)g.N_t * g.N_r));
    const size_t k = xt_sidx(g, p, r, t);
    if (to_hbm)
      dst[k] = src[i];
    else
      dst[i] = src[k];
  }
}

__global__ void k_zero_f64(double *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Start of an iteration inside a captured graph: advance the iteration number
// and clear the per-iteration statistics.
__global__ void k_begin_iteration(int *iter, double *stats, int n) {
  const int i = threadIdx.x;
  if (i == 0) *iter += 1;
  if (i < n) stats[i] = 0.0;
}
