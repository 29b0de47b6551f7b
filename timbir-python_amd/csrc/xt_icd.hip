// xt_icd.hip -- host side of libxt_hip.so: ICD_BackProject and the stepwise
// and seam entry points of include/xt_icd.h.
//
// One process, one GPU, one stream.  Per ICD iteration the host enqueues, for
// each of K = NHICD_Iterations sub-iterations: two launches of the voxel-update
// kernel over the two halves of this sub-iteration's 1/K share of a random
// order of every frame's lines, a ring-offset update after each, the NHICD
// selection, and two launches over the selected lines (from the second
// iteration on) -- the schedule of Source_Code_4D_Fast/XT_ICD_update.c:767-926
// with z-blocks replaced by whole voxel lines and OpenMP threads by up to 1024
// workgroups per frame, each updating one line at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xt_icd.h"
#include "xt_device.h"
#include "xt_unet_internal.h"

// Real_t is float (the reference's production build; libxt_hip.so) or, with
// XT_REAL_IS_DOUBLE, double (its REAL_IS_DOUBLE build; libxt_hip_f64.so).  Only the
// host side sees it: the device keeps float arrays, and the double build keeps the
// float build's system matrix (detector table and A columns from float geometry).
#ifdef XT_REAL_IS_DOUBLE
#define XT_REAL_NAME "double"
#else
#define XT_REAL_NAME "float"
#endif

// ---------------------------------------------------------------------------
// errors, options
static thread_local char g_err[1024] = "";

static int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int xt_report_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define XT_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t _e = (call);                                                            \
    if (_e != hipSuccess)                                                              \
      return set_err(XT_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                     __FILE__, __LINE__);                                              \
  } while (0)

#define XT_TRY(expr)          \
  do {                        \
    int _rc = (expr);         \
    if (_rc != XT_OK) return _rc; \
  } while (0)

static double g_opt[20] = {0, 0, 0, 0, 761521, 0, 0, 1, 0, 1, 1, 1, 1, 0, 1, 0, 0, 0, 0, 1};  // XT_OPT_*
// bumped whenever an option changes value: a captured iteration graph (which
// froze the options it was recorded under) is recaptured when it differs
static uint64_t g_opt_gen = 1;
std::string g_xt_run_dir;  // directory of ICD_BackProject's own files ("" = cwd)

extern "C" const char *xt_last_error(void) { return g_err; }
extern "C" const char *xt_version(void) { return "libxt_hip gfx950 Real_t=" XT_REAL_NAME; }

extern "C" int xt_set_option(int option, double value) {
  if (option < 1 || option > 19) return set_err(XT_E_ARG, "unknown option %d", option);
  if (option == XT_OPT_LINE_THREADS && value != 0 && value != 64 && value != 256)
    return set_err(XT_E_ARG, "XT_OPT_LINE_THREADS must be 0, 64 or 256");
  // groups of 2 or 4 lines (k_icd_groups) are switched off: their first run on the
  // GPU took a memory fault whose cause is not found (the kernel is memory-safe by
  // construction and passes the per-voxel checks in a CPU build of this file)
  if (option == XT_OPT_LINE_GROUP && value != 1)
    return set_err(XT_E_ARG, "XT_OPT_LINE_GROUP: only 1 (single lines) is enabled in this build");
  if (g_opt[option] != value) g_opt_gen++;
  g_opt[option] = value;
  return XT_OK;
}
extern "C" double xt_get_option(int option) {
  return (option >= 1 && option <= 19) ? g_opt[option] : NAN;
}

static std::string run_path(const char *name) {
  if (g_xt_run_dir.empty()) return std::string(name);
  return g_xt_run_dir + "/" + name;
}

static void dbg(TomoInputs *T, const char *fmt, ...) {
  if (!T || !T->debug_file_ptr || g_opt[XT_OPT_VERBOSE] == 0) return;
  va_list ap;
  va_start(ap, fmt);
  vfprintf(T->debug_file_ptr, fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------------------
// pointer trees with the reference layout (F/allocate.c:91-188)
extern "C" void *xt_multialloc(size_t s, int d, const int64_t *dims) {
  if (d < 1) return nullptr;
  if (d == 1) return calloc(dims[0] ? (size_t)dims[0] : 1, s);
  size_t count = 1;
  void **prev = nullptr;
  void *top = nullptr;
  for (int lvl = 0; lvl < d - 1; lvl++) {
    size_t n = count * (size_t)dims[lvl];
    void **arr = (void **)calloc(n ? n : 1, sizeof(void *));
    if (!arr) return nullptr;
    if (lvl == 0)
      top = arr;
    else
      for (size_t j = 0; j < count; j++) prev[j] = (void *)(arr + j * (size_t)dims[lvl]);
    prev = arr;
    count = n;
  }
  size_t row = (size_t)dims[d - 1] * s;
  char *data = (char *)calloc(count * (size_t)dims[d - 1] ? count * (size_t)dims[d - 1] : 1, s);
  if (!data) return nullptr;
  for (size_t j = 0; j < count; j++) prev[j] = data + j * row;
  return top;
}

extern "C" void xt_multifree(void *r, int d) {
  void **p = (void **)r;
  for (int i = 0; i < d && p != nullptr; i++) {
    void *next = *p;
    free(p);
    p = (void **)next;
  }
}

// ---------------------------------------------------------------------------
// geometry from the engine structs
static int zl_for(int nz) { return nz <= 8 ? 8 : nz <= 16 ? 16 : nz <= 32 ? 32 : 64; }

static int make_geom(const Sinogram *S, const ScannedObject *O, const TomoInputs *T, XtGeom *g) {
  memset(g, 0, sizeof *g);
  if (S->N_p <= 0 || S->N_r <= 0 || S->N_t <= 0 || O->N_x <= 0 || O->N_y <= 0 || O->N_z <= 0 ||
      O->N_time <= 0)
    return set_err(XT_E_ARG, "empty problem (N_p=%d N_r=%d N_t=%d N_x=%d N_y=%d N_z=%d N_time=%d)",
                   S->N_p, S->N_r, S->N_t, O->N_x, O->N_y, O->N_z, O->N_time);
  // the voxel-update kernels address e and w through 32-bit buffer offsets, and an
  // offset of 2^31 is their "absent bin" value (reads 0, drops the atomic): every
  // byte offset into the sinogram must stay below it
  if ((size_t)S->N_p * (size_t)S->N_r * (size_t)S->N_t * sizeof(float) >= ((size_t)1 << 31))
    return set_err(XT_E_ARG, "sinogram of %d x %d x %d entries: at most 2^29 - 1 entries per GPU",
                   S->N_p, S->N_r, S->N_t);
  // a slice spans z_overlap_num = mult_z detector rows (F/XT_Init.c:617-621)
  if (S->z_overlap_num < 1 || S->z_overlap_num > 8 || S->N_t != O->N_z * S->z_overlap_num)
    return set_err(XT_E_ARG, "z_overlap_num=%d with N_t=%d, N_z=%d: need N_t = N_z * "
                   "z_overlap_num, 1 <= z_overlap_num <= 8", S->z_overlap_num, S->N_t, O->N_z);
  if (O->N_z > 64) return set_err(XT_E_ARG, "N_z=%d > 64 slices per GPU", O->N_z);
  if (2.0 * O->delta_xy / S->delta_r + 2.0 > XT_MAXA_WIDE)
    return set_err(XT_E_ARG, "delta_xy/delta_r = %g too large for one A column (at most 6)",
                   (double)(O->delta_xy / S->delta_r));
  g->wide = 2.0 * O->delta_xy / S->delta_r + 2.0 > XT_MAXA;
  g->pc = g_opt[XT_OPT_PHASE_CONTRAST] != 0;
  if (g->pc && g->wide)
    return set_err(XT_E_ARG, "phase contrast needs voxels at most 2 detector bins wide");
  if (g->pc && S->z_overlap_num != 1)
    return set_err(XT_E_ARG, "phase contrast needs slices of one detector row (delta_z = 1)");
  for (int i = 0; i < 9; i++) g->lap[i] = S->Lap_Kernel[i / 3][i % 3];
  if (T && T->node_num > 1)
    return set_err(XT_E_ARG, "node_num=%d: this engine runs one slab on one GPU", T->node_num);
  g->N_p = S->N_p;
  g->N_r = S->N_r;
  g->N_t = S->N_t;
  g->N_x = O->N_x;
  g->N_y = O->N_y;
  g->N_z = O->N_z;
  g->N_time = O->N_time;
  g->zo = S->z_overlap_num;
  g->nb = (T && T->num_z_blocks > 0) ? T->num_z_blocks : 1;
  if (g->nb > 64) g->nb = 64;
  g->pm = 0;  // [N_p][N_r][N_t] unless the ICD state splits it by parity
  g->bs = O->N_z / g->nb > 0 ? O->N_z / g->nb : 1;
  g->H = S->N_t;
  g->delta_r = S->delta_r;
  g->delta_t = S->delta_t;
  g->R0 = S->R0;
  g->RMax = S->RMax;
  g->OffsetR = S->OffsetR;
  g->x0 = O->x0;
  g->y0 = O->y0;
  g->delta_xy = O->delta_xy;
  g->MRF_P = O->MRF_P;
  g->C_S = O->C_S;
  g->C_T = O->C_T;
  if (T) {
    g->Sigma_S_Q = T->Sigma_S_Q;
    g->Sigma_S_Q_P = T->Sigma_S_Q_P;
    g->Sigma_T_Q = T->Sigma_T_Q;
    g->Sigma_T_Q_P = T->Sigma_T_Q_P;
    const Real_t *sf = &T->Spatial_Filter[0][0][0];
    for (int k = 0; k < 27; k++) g->filt[k] = (float)sf[k];
    g->tfilt = T->Time_Filter[0];
    g->alpha = T->alpha;
    g->T = T->ErrorSinoThresh;
    g->delta = T->ErrorSinoDelta;
  }
  g->positivity = g_opt[XT_OPT_POSITIVITY] != 0;
  int mv = 0;
  for (int i = 0; i < O->N_time; i++) mv = O->ProjNum[i] > mv ? O->ProjNum[i] : mv;
  g->max_views = mv > 0 ? mv : 1;
  return XT_OK;
}

// dynamic LDS of k_icd_lines: per view the A column (float4, plus float2 for
// bins 4..5 without the table) and its e offset; theta partials; neighbourhood.
static size_t icd_lds_bytes(const XtGeom &g, int ZL, bool table, int nt = 256) {
  const size_t col = g.wide ? XT_MAXA_WIDE * sizeof(float)
                            : sizeof(float4) + (table ? 0 : sizeof(float2));
  return (size_t)g.max_views * (col + sizeof(int)) + 2 * (size_t)nt * sizeof(float) +
         14 * ZL * sizeof(float) + 64 * sizeof(float);
}

template <typename T>
static int dev_alloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess)
    return set_err(XT_E_HIP, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  return XT_OK;
}

// Host arrays hold Real_t, device arrays float: n values between them, a plain
// copy when Real_t is float, converted through a host buffer when it is double.
static hipError_t real_to_dev(float *dst, const Real_t *src, size_t n) {
  if (sizeof(Real_t) == sizeof(float))
    return hipMemcpy(dst, src, n * sizeof(float), hipMemcpyHostToDevice);
  std::vector<float> t(src, src + n);
  return hipMemcpy(dst, t.data(), n * sizeof(float), hipMemcpyHostToDevice);
}

static hipError_t dev_to_real(Real_t *dst, const float *src, size_t n) {
  if (sizeof(Real_t) == sizeof(float))
    return hipMemcpy(dst, src, n * sizeof(float), hipMemcpyDeviceToHost);
  std::vector<float> t(n);
  hipError_t e = hipMemcpy(t.data(), src, n * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) std::copy(t.begin(), t.end(), dst);
  return e;
}

// The views' cosines, sines (either may be null) and angles on the device.  The
// float build takes the host's (calculateSinCos in Real_t = float, F/XT_Init.c:
// 259-272); the double build uses the float build's values -- the angle rounded to
// float, its cosine in double rounded to float -- so both keep one system matrix.
static int upload_view_tables(float *cosv, float *sinv, float *viewang, const Sinogram *S) {
  const int n = S->N_p;
  std::vector<float> c(n), sn(n), a(n);
  for (int i = 0; i < n; i++) {
    a[i] = (float)S->ViewPtr[i];
    c[i] = sizeof(Real_t) == sizeof(float) ? (float)S->cosine[i] : (float)cos((double)a[i]);
    sn[i] = sizeof(Real_t) == sizeof(float) ? (float)S->sine[i] : (float)sin((double)a[i]);
  }
  hipError_t e = hipSuccess;
  if (cosv) e = hipMemcpy(cosv, c.data(), n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess && sinv) e = hipMemcpy(sinv, sn.data(), n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess && viewang)
    e = hipMemcpy(viewang, a.data(), n * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) return set_err(XT_E_HIP, "view tables: %s", hipGetErrorString(e));
  return XT_OK;
}

static int grid_1d(size_t n, int bs = 256, int cap = 8192) {
  size_t b = (n + bs - 1) / bs;
  if (b < 1) b = 1;
  if (b > (size_t)cap) b = cap;
  return (int)b;
}

// ---------------------------------------------------------------------------
// stepwise state
struct xt_icd_state {
  int dev = 0;
  hipStream_t st = nullptr;
  XtGeom g;
  int ZL = 32;
  int nt = 256;  // threads per voxel-update workgroup (256 or 64)
  int K = 10, NumIter = 0, update_offset = 0, zero_mean = 0, update_var = 0, no_nhicd = 0;
  float var_est = 1.f, stop_threshold = 0.f;
  int iter = 0, stopped = 0;
  int n_masked = 0, conc = 16;  // masked lines; workgroups (concurrent lines) per frame
  size_t nsino = 0, nobj = 0, NYX = 0;
  TomoInputs *T = nullptr;  // for debug lines only
  // device buffers
  float *e = nullptr, *w = nullptr, *y = nullptr, *d = nullptr, *x = nullptr, *Hr = nullptr;
  float *cosv = nullptr, *sinv = nullptr, *viewang = nullptr, *fbp_w = nullptr;
  int *views = nullptr, *proj_num = nullptr, *view_frame = nullptr;
  int *vpos = nullptr;  // each view's position in its frame's list
  int *lines = nullptr, *nhicd_list = nullptr, *nhicd_count = nullptr;
  // the homogeneous passes' line groups: 1 x 2 tiles (groups2) and 2 x 2 tiles
  // (groups4) of the line list, members first and -1 after them
  int *groups2 = nullptr, *groups4 = nullptr;
  int n_groups2 = 0, n_groups4 = 0;
  int *nhicd_chunk = nullptr;  // NHICD list compaction: lines per chunk of each map
  uint8_t *line_mask = nullptr;
  uint8_t *nhicd_flag = nullptr, *nhicd_blk = nullptr;
  float *mag_blk = nullptr, *magsum = nullptr, *filt = nullptr, *magpar = nullptr;
  int *nsel = nullptr, *nsel_out = nullptr;  // NHICD lines per (frame, z-block) map
  int *nsel_par = nullptr;  // NHICD lines per (frame, z-block parity) map
  unsigned *vsc_st = nullptr, *vsc_hist = nullptr;  // NHICD radix selection scratch
  double *stats = nullptr, *red_wg = nullptr, *scal = nullptr;
  double *off_part = nullptr, *off_numden = nullptr, *off_wg = nullptr;
  float4 *at_vals = nullptr;  // A-column table (null: columns built on the fly)
  int *at_first = nullptr;
  int *iter_dev = nullptr;      // iteration number read by the voxel-update kernel
  double *pc_tmp = nullptr;     // phase contrast: the absorption projection (FP64)
  int edge = 0;                 // edge-only updates (only_Edge_Updates)
  XtEdgeParams edge_par{};
  uint8_t *vmask = nullptr;     // edge-only: [N_time][N_y][N_x][N_z] voxels to update
  hipGraph_t graph = nullptr;   // one captured iteration (iterations >= 2)
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_opt_gen = 0;   // g_opt_gen when the graph was captured
  uint64_t lines_gen = 1;       // bumped when the line list or its NHICD shares change
  int64_t graph_launches = 0;   // voxel-update launches in the captured iteration
  int off_nch = 1, off_chunk = 1, off_nblk = 1;
  // statistics
  double total_updates = 0, total_foot = 0, total_visits = 0, icd_ms = 0, last_cost = NAN,
         stop_pct = NAN;
  int64_t launches = 0;
  int timing = 0;
  std::vector<hipEvent_t> ev;
  int ev_used = 0;
};

static void free_state(xt_icd_state *s) {
  if (!s) return;
  void *bufs[] = {s->e,        s->w,          s->y,        s->d,          s->x,          s->Hr,
                  s->fbp_w,
                  s->cosv,     s->sinv,       s->viewang,  s->views,      s->proj_num,   s->view_frame,
                  s->vpos,
                  s->lines,    s->nhicd_list, s->nhicd_count, s->nhicd_chunk, s->line_mask, s->nhicd_flag,
                  s->nhicd_blk,
                  s->mag_blk,  s->magsum,     s->filt,     s->nsel,       s->nsel_out,   s->stats,
                  s->nsel_par,
                  s->magpar,
                  s->red_wg,   s->scal,       s->off_part, s->off_numden, s->off_wg,
                  s->at_vals,  s->at_first,   s->iter_dev, s->pc_tmp, s->vmask, s->vsc_st,
                  s->vsc_hist, s->groups2, s->groups4};
  if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
}

static int launch_detector_response(xt_icd_state *s, float beam_width) {
  hipLaunchKernelGGL(k_detector_response, dim3(s->g.N_p), dim3(256), 0, s->st, s->g, s->viewang,
                     beam_width, s->Hr);
  XT_HIP(hipGetLastError());
  return XT_OK;
}

template <int ZL>
static void fp_launch(hipStream_t st, const XtGeom &g, const float *x, const float *Hr,
                      const float *c, const float *s, const int *vf, int mode, double *out64,
                      const float *y, const float *d, float *e, const XtFpTable &tab) {
  dim3 grid((g.N_r + XT_FP_BINS - 1) / XT_FP_BINS, g.N_p);
  if (tab.vals && !g.wide)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_forward_project<ZL, 4, true>), grid, dim3(256), 0, st, g,
                       x, Hr, c, s, vf, mode, out64, y, d, e, tab);
  else if (g.wide)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_forward_project<ZL, XT_MAXA_WIDE>), grid, dim3(256), 0, st,
                       g, x, Hr, c, s, vf, mode, out64, y, d, e);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_forward_project<ZL, XT_MAXA>), grid, dim3(256), 0, st, g,
                       x, Hr, c, s, vf, mode, out64, y, d, e);
}

// Ax (mode 0, FP64 into out64) or e = y - Ax - d (mode 1).  Phase contrast:
// the absorption projection into pc_tmp (FP64, a sinogram's size), then its
// detector Laplacian.
static int launch_forward(hipStream_t st, int ZL, const XtGeom &g, const float *x, const float *Hr,
                          const float *c, const float *s, const int *vf, int mode, double *out64,
                          const float *y, const float *d, float *e, double *pc_tmp = nullptr,
                          const XtFpTable &tab = {nullptr, nullptr, nullptr}) {
  double *ax = out64;
  int m = mode;
  if (g.pc) {
    if (!pc_tmp) return set_err(XT_E_ARG, "phase contrast without its projection buffer");
    ax = pc_tmp;
    m = 0;
  }
  switch (ZL) {
    case 8: fp_launch<8>(st, g, x, Hr, c, s, vf, m, ax, y, d, e, tab); break;
    case 16: fp_launch<16>(st, g, x, Hr, c, s, vf, m, ax, y, d, e, tab); break;
    case 32: fp_launch<32>(st, g, x, Hr, c, s, vf, m, ax, y, d, e, tab); break;
    default: fp_launch<64>(st, g, x, Hr, c, s, vf, m, ax, y, d, e, tab); break;
  }
  if (g.pc) {
    const size_t n = (size_t)g.N_p * g.N_r * g.N_t;
    hipLaunchKernelGGL(k_pc_laplace, dim3((unsigned)std::min<size_t>((n + 255) / 256, 16384)),
                       dim3(256), 0, st, g, pc_tmp, mode, out64, y, d, e);
  }
  XT_HIP(hipGetLastError());
  return XT_OK;
}

// Packed lanes (the active z-block parity's slices in ZL/2 lanes) need blocks of
// equal size and an even number of them.
static bool icd_packable(const XtGeom &g, int ZL, const XtIcdLaunch &L) {
  (void)ZL;
  return g.pm != 0 && L.mode != 2;  // mode 2 updates every block's slices
}

template <int ZL, int NT>
static void icd_launch_nt(hipStream_t st, const XtGeom &g, const XtIcdBuffers &B,
                          const XtIcdLaunch &L, int gx, size_t lds) {
  const dim3 grid(gx, g.N_time), block(NT);
  const bool pack = icd_packable(g, ZL, L);
  if (B.gl == 2 || B.gl == 4) {  // groups of lines, from the table
    if (B.gl == 2) {
      if (pack)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_groups<ZL, true, NT, 2>), grid, block, lds, st, g,
                           B, L);
      else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_groups<ZL, false, NT, 2>), grid, block, lds, st, g,
                           B, L);
    } else {
      if (pack)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_groups<ZL, true, NT, 4>), grid, block, lds, st, g,
                           B, L);
      else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_groups<ZL, false, NT, 4>), grid, block, lds, st, g,
                           B, L);
    }
  } else if (g.wide) {  // never with the table (it holds four bins per column)
    if (pack)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_lines<ZL, false, true, XT_MAXA_WIDE, NT>), grid,
                         block, lds, st, g, B, L);
    else
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_lines<ZL, false, false, XT_MAXA_WIDE, NT>), grid,
                         block, lds, st, g, B, L);
  } else if (B.at_vals) {
    if (pack)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_lines<ZL, true, true, XT_MAXA, NT>), grid, block,
                         lds, st, g, B, L);
    else
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_lines<ZL, true, false, XT_MAXA, NT>), grid, block,
                         lds, st, g, B, L);
  } else {
    if (pack)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_lines<ZL, false, true, XT_MAXA, NT>), grid, block,
                         lds, st, g, B, L);
    else
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_lines<ZL, false, false, XT_MAXA, NT>), grid, block,
                         lds, st, g, B, L);
  }
}

template <int ZL>
static void icd_launch(hipStream_t st, const XtGeom &g, const XtIcdBuffers &B, const XtIcdLaunch &L,
                       int gx, size_t lds, int nt) {
  if (nt == 64)
    icd_launch_nt<ZL, 64>(st, g, B, L, gx, lds);
  else
    icd_launch_nt<ZL, 256>(st, g, B, L, gx, lds);
}

static XtIcdBuffers buffers_of(xt_icd_state *s) {
  XtIcdBuffers B;
  B.e = s->e;
  B.w = s->w;
  B.x = s->x;
  B.Hr = s->Hr;
  B.cosv = s->cosv;
  B.sinv = s->sinv;
  B.views = s->views;
  B.proj_num = s->proj_num;
  B.lines = s->lines;
  B.n_lines = s->n_masked;
  B.nhicd_list = s->nhicd_list;
  B.nhicd_count = s->nhicd_count;
  B.nhicd_flag = s->nhicd_flag;
  B.nhicd_blk = s->nhicd_blk;
  B.mag_blk = s->mag_blk;
  B.magsum = s->magsum;
  B.stats = s->stats;
  B.at_vals = s->at_vals;
  B.at_first = s->at_first;
  B.iter = s->iter_dev;
  B.vmask = s->vmask;
  B.groups = nullptr;
  B.n_groups = 0;
  B.gl = 1;
  return B;
}

// One voxel-update launch over a range of lines; gx workgroups per frame.
// Lines per workgroup item of the voxel-update launches (XT_OPT_LINE_GROUP): 2
// or 4 with the A-column table and slices of one detector row, else 1.
static int line_group(const xt_icd_state *s) {
  const int gl = (int)g_opt[XT_OPT_LINE_GROUP];
  if ((gl != 2 && gl != 4) || !s->at_vals || s->g.wide || s->g.pc || s->g.zo != 1) return 1;
  return gl;
}

static size_t group_lds_bytes(const XtGeom &g, int ZL, int gl, int nt) {
  return (size_t)g.max_views * sizeof(int) + 2 * (size_t)nt * sizeof(float) +
         (size_t)(13 + gl) * ZL * sizeof(float) + 64 * sizeof(float);
}

static int launch_icd(xt_icd_state *s, const XtIcdLaunch &L, int gx) {
  XtIcdBuffers B = buffers_of(s);
  size_t lds = icd_lds_bytes(s->g, s->ZL, B.at_vals != nullptr, s->nt);
  const int gl = line_group(s);
  if (gl > 1) {
    B.gl = gl;
    B.groups = gl == 2 ? s->groups2 : s->groups4;
    B.n_groups = gl == 2 ? s->n_groups2 : s->n_groups4;
    lds = group_lds_bytes(s->g, s->ZL, gl, s->nt);
    if (L.mode == 0) {  // no more workgroups than the sub-iteration has groups
      const int per = B.n_groups / L.K;
      const int share = (L.sub_iter == L.K - 1) ? B.n_groups - L.sub_iter * per : per;
      if (gx > share) gx = share;
    }
  }
  if (gx < 1) gx = 1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->timing) {
    if (s->ev_used + 2 > (int)s->ev.size()) {
      for (int i = 0; i < 64; i++) {
        hipEvent_t e;
        XT_HIP(hipEventCreate(&e));
        s->ev.push_back(e);
      }
    }
    e0 = s->ev[s->ev_used];
    e1 = s->ev[s->ev_used + 1];
    s->ev_used += 2;
    XT_HIP(hipEventRecord(e0, s->st));
  }
  if (s->g.pc) {
    const dim3 grid(gx, s->g.N_time), block(256);
    switch (s->ZL) {
      case 8: hipLaunchKernelGGL(k_icd_lines_pc<8>, grid, block, lds, s->st, s->g, B, L); break;
      case 16: hipLaunchKernelGGL(k_icd_lines_pc<16>, grid, block, lds, s->st, s->g, B, L); break;
      case 32: hipLaunchKernelGGL(k_icd_lines_pc<32>, grid, block, lds, s->st, s->g, B, L); break;
      default: hipLaunchKernelGGL(k_icd_lines_pc<64>, grid, block, lds, s->st, s->g, B, L); break;
    }
    XT_HIP(hipGetLastError());
    if (s->timing) XT_HIP(hipEventRecord(e1, s->st));
    s->launches++;
    return XT_OK;
  }
  switch (s->ZL) {
    case 8: icd_launch<8>(s->st, s->g, B, L, gx, lds, s->nt); break;
    case 16: icd_launch<16>(s->st, s->g, B, L, gx, lds, s->nt); break;
    case 32: icd_launch<32>(s->st, s->g, B, L, gx, lds, s->nt); break;
    default: icd_launch<64>(s->st, s->g, B, L, gx, lds, s->nt); break;
  }
  XT_HIP(hipGetLastError());
  if (s->timing) XT_HIP(hipEventRecord(e1, s->st));
  s->launches++;
  return XT_OK;
}

static int offset_update(xt_icd_state *s) {
  const int NRT = s->g.N_r * s->g.N_t;
  dim3 g2((NRT + 255) / 256, s->off_nch);
  hipLaunchKernelGGL(k_offset_partial, g2, dim3(256), 0, s->st, s->g, s->e, s->w, s->off_chunk,
                     s->off_part);
  hipLaunchKernelGGL(k_offset_combine, dim3(s->off_nblk), dim3(256), 0, s->st, s->g, s->off_nch,
                     s->off_part, s->d, s->off_numden, s->off_wg);
  hipLaunchKernelGGL(k_offset_gamma, dim3(1), dim3(256), 0, s->st, s->off_nblk, s->off_wg,
                     s->zero_mean, s->scal);
  hipLaunchKernelGGL(k_offset_apply, g2, dim3(256), 0, s->st, s->g, s->off_chunk, s->off_numden,
                     s->scal, s->d, s->e);
  XT_HIP(hipGetLastError());
  return XT_OK;
}

static int variance_update(xt_icd_state *s) {
  int nb = grid_1d(s->nsino, 256, 4096);
  hipLaunchKernelGGL(k_variance_partial, dim3(nb), dim3(256), 0, s->st, s->g, s->e, s->w, s->var_est,
                     s->nsino, s->red_wg);
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, s->st, s->red_wg, nb, s->scal + 1);
  XT_HIP(hipGetLastError());
  double sum = 0;
  XT_HIP(hipMemcpyAsync(&sum, s->scal + 1, sizeof(double), hipMemcpyDeviceToHost, s->st));
  XT_HIP(hipStreamSynchronize(s->st));
  float var_old = s->var_est;
  float var_new = (float)(sum / ((double)s->g.N_p * s->g.N_r * s->g.N_t));
  hipLaunchKernelGGL(k_variance_apply, dim3(nb), dim3(256), 0, s->st, s->w, s->nsino, var_old,
                     var_new);
  XT_HIP(hipGetLastError());
  s->var_est = var_new;
  return XT_OK;
}

// The NHICD threshold and flags of M filtered maps (k_vsc_init, four radix passes,
// k_vsc_flag); st holds 4 and hist 256 words per map.
static void launch_vsc_select(hipStream_t st, int NYX, int M, const float *filt, const int *nsel,
                              uint8_t *flag, int *nout, unsigned *vst, unsigned *hist) {
  const dim3 grid((unsigned)std::max(1, std::min(64, (NYX + 4095) / 4096)), M);
  hipLaunchKernelGGL(k_vsc_init, dim3(M), dim3(256), 0, st, NYX, nsel, vst, hist, nout);
  for (int shift = 24; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(k_vsc_hist, grid, dim3(256), 0, st, NYX, filt, vst, hist, shift);
    hipLaunchKernelGGL(k_vsc_pick, dim3(M), dim3(256), 0, st, vst, hist, shift);
  }
  hipLaunchKernelGGL(k_vsc_flag, grid, dim3(256), 0, st, NYX, filt, vst, flag, nout);
}

// NHICD selection per (frame, z-block) map as F/XT_NHICD.c:178-205, then the
// per-line union that decides which lines the NHICD launches visit.
static int vsc_select(xt_icd_state *s) {
  const unsigned nl = (unsigned)((s->NYX * s->g.N_time + 255) / 256);
  if (g_opt[XT_OPT_NHICD_LINES] != 0) {
    // one selection per (frame, parity) on the parity's summed map; every
    // block of the parity takes it
    const int M = s->g.N_time * 2;
    hipLaunchKernelGGL(k_mag_parity, dim3(nl), dim3(256), 0, s->st, s->g.nb, s->NYX, s->g.N_time,
                       s->mag_blk, s->magpar);
    dim3 gf((unsigned)((s->NYX + 255) / 256), M);
    hipLaunchKernelGGL(k_vsc_filter, gf, dim3(256), 0, s->st, s->g.N_y, s->g.N_x, s->magpar,
                       s->filt);
    launch_vsc_select(s->st, (int)s->NYX, M, s->filt, s->nsel_par, s->nhicd_flag, s->nsel_out,
                      s->vsc_st, s->vsc_hist);
    hipLaunchKernelGGL(k_flag_blocks, dim3(nl), dim3(256), 0, s->st, s->g.nb, s->NYX,
                       s->g.N_time, s->nhicd_flag, s->nhicd_blk);
  } else {
    const int M = s->g.N_time * s->g.nb;
    dim3 gf((unsigned)((s->NYX + 255) / 256), M);
    hipLaunchKernelGGL(k_vsc_filter, gf, dim3(256), 0, s->st, s->g.N_y, s->g.N_x, s->mag_blk,
                       s->filt);
    launch_vsc_select(s->st, (int)s->NYX, M, s->filt, s->nsel, s->nhicd_blk, s->nsel_out,
                      s->vsc_st, s->vsc_hist);
    hipLaunchKernelGGL(k_vsc_any, dim3(nl), dim3(256), 0, s->st, s->g.nb, s->NYX, s->g.N_time,
                       s->nhicd_blk, s->nhicd_flag);
  }
  const int nch = (int)((s->NYX + XT_COMPACT_CHUNK - 1) / XT_COMPACT_CHUNK);
  const dim3 gc((unsigned)nch, (unsigned)(s->g.N_time * 2));
  hipLaunchKernelGGL(k_nhicd_chunk_count, gc, dim3(256), 0, s->st, (int)s->NYX, nch,
                     s->nhicd_flag, s->line_mask, s->nhicd_chunk);
  hipLaunchKernelGGL(k_nhicd_chunk_write, gc, dim3(256), 0, s->st, (int)s->NYX, nch,
                     s->nhicd_flag, s->line_mask, s->nhicd_chunk, s->nhicd_list, s->nhicd_count);
  XT_HIP(hipGetLastError());
  return XT_OK;
}

// The masked lines on the device, and how many lines of a frame are updated
// at once: at most 1024 and at most 1/64 of the lines (concurrently updated
// lines read each other's stale error sinogram; a CPU simulation of that
// coupling kept the reference's trajectory up to ~6% of the lines at once
// and diverged at 25%), or XT_OPT_LINES when set.
// The groups of the homogeneous passes (XT_OPT_LINE_GROUP 2 and 4): the listed
// lines of each 1 x 2 and of each 2 x 2 tile of the image, tiles in raster order.
static int build_groups(xt_icd_state *s, const std::vector<int> &lines) {
  const int NX = s->g.N_x, NY = s->g.N_y;
  std::vector<uint8_t> in(s->NYX, 0);
  for (int c : lines) in[c] = 1;
  for (int gl : {2, 4}) {
    const int th = gl == 4 ? 2 : 1;  // tile rows
    std::vector<int> grp;
    for (int ty = 0; ty < NY; ty += th)
      for (int tx = 0; tx < NX; tx += 2) {
        int mem[4], n = 0;
        for (int dy = 0; dy < th; dy++)
          for (int dx = 0; dx < 2; dx++) {
            const int y = ty + dy, x = tx + dx;
            if (y < NY && x < NX && in[(size_t)y * NX + x]) mem[n++] = y * NX + x;
          }
        if (n == 0) continue;
        for (int m = 0; m < gl; m++) grp.push_back(m < n ? mem[m] : -1);
      }
    int *&dst = gl == 2 ? s->groups2 : s->groups4;
    if (dst) (void)hipFree(dst);
    dst = nullptr;
    XT_TRY(dev_alloc(&dst, grp.empty() ? 1 : grp.size()));
    if (!grp.empty())
      XT_HIP(hipMemcpy(dst, grp.data(), grp.size() * sizeof(int), hipMemcpyHostToDevice));
    (gl == 2 ? s->n_groups2 : s->n_groups4) = (int)(grp.size() / gl);
  }
  return XT_OK;
}

static int setup_lines(xt_icd_state *s, const std::vector<int> &masked) {
  int conc = (int)g_opt[XT_OPT_LINES];
  if (conc <= 0) {
    conc = s->n_masked / 64;
    if (conc > 1024) conc = 1024;
    if (conc < 16) conc = 16;
  }
  if (conc > 65535) return set_err(XT_E_ARG, "XT_OPT_LINES %d > 65535", conc);
  s->conc = conc;
  std::vector<uint8_t> in(s->NYX, 0);
  for (int c : masked) in[c] = 1;
  XT_TRY(dev_alloc(&s->lines, masked.empty() ? 1 : masked.size()));
  XT_TRY(dev_alloc(&s->line_mask, s->NYX));
  XT_TRY(dev_alloc(&s->nhicd_list, (size_t)s->g.N_time * 2 * s->NYX));
  XT_TRY(dev_alloc(&s->nhicd_count, s->g.N_time * 2));
  XT_TRY(dev_alloc(&s->nhicd_chunk, (size_t)s->g.N_time * 2 *
                                        ((s->NYX + XT_COMPACT_CHUNK - 1) / XT_COMPACT_CHUNK)));
  if (!masked.empty())
    XT_HIP(hipMemcpy(s->lines, masked.data(), masked.size() * sizeof(int), hipMemcpyHostToDevice));
  XT_HIP(hipMemcpy(s->line_mask, in.data(), s->NYX, hipMemcpyHostToDevice));
  XT_HIP(hipMemset(s->nhicd_count, 0, s->g.N_time * 2 * sizeof(int)));
  return build_groups(s, masked);
}

// the detector-table kernel reduces angles by repeated subtraction of pi/2
static int check_angles(const Sinogram *S) {
  for (int i = 0; i < S->N_p; i++)
    if (!std::isfinite(S->ViewPtr[i]) || fabsf(S->ViewPtr[i]) > 1.0e5f)
      return set_err(XT_E_ARG, "view %d angle %g rad outside +-1e5", i, (double)S->ViewPtr[i]);
  return XT_OK;
}

static int upload_views(xt_icd_state *s, const Sinogram *S, const ScannedObject *O) {
  const XtGeom &g = s->g;
  XT_TRY(check_angles(S));
  std::vector<int> flat((size_t)g.N_time * g.max_views, 0), num(g.N_time), vf(g.N_p, -1),
      vp(g.N_p, 0);
  for (int f = 0; f < g.N_time; f++) {
    num[f] = O->ProjNum[f];
    for (int p = 0; p < O->ProjNum[f]; p++) {
      int v = O->ProjIdxPtr[f][p];
      if (v < 0 || v >= g.N_p) return set_err(XT_E_ARG, "view index %d out of range", v);
      flat[(size_t)f * g.max_views + p] = v;
      vf[v] = f;
      vp[v] = p;
    }
  }
  XT_TRY(dev_alloc(&s->vpos, g.N_p));
  XT_HIP(hipMemcpy(s->vpos, vp.data(), vp.size() * sizeof(int), hipMemcpyHostToDevice));
  XT_TRY(dev_alloc(&s->views, flat.size()));
  XT_TRY(dev_alloc(&s->proj_num, g.N_time));
  XT_TRY(dev_alloc(&s->view_frame, g.N_p));
  XT_TRY(dev_alloc(&s->cosv, g.N_p));
  XT_TRY(dev_alloc(&s->sinv, g.N_p));
  XT_TRY(dev_alloc(&s->viewang, g.N_p));
  XT_HIP(hipMemcpy(s->views, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice));
  XT_HIP(hipMemcpy(s->proj_num, num.data(), num.size() * sizeof(int), hipMemcpyHostToDevice));
  XT_HIP(hipMemcpy(s->view_frame, vf.data(), vf.size() * sizeof(int), hipMemcpyHostToDevice));
  XT_TRY(upload_view_tables(s->cosv, s->sinv, s->viewang, S));
  // backprojection weights: each view's share of [0, pi) among its frame's view
  // angles taken modulo pi (half the gap to each neighbour, circularly)
  std::vector<float> wv(g.N_p, 0.f);
  for (int f = 0; f < g.N_time; f++) {
    const int n = O->ProjNum[f];
    std::vector<std::pair<double, int>> a(n);
    for (int p = 0; p < n; p++) {
      const int v = O->ProjIdxPtr[f][p];
      double t = fmod((double)S->ViewPtr[v], XT_PI);
      if (t < 0) t += XT_PI;
      a[p] = {t, v};
    }
    std::sort(a.begin(), a.end());
    for (int p = 0; p < n; p++) {
      const double lo = p > 0 ? a[p - 1].first : a[n - 1].first - XT_PI;
      const double hi = p + 1 < n ? a[p + 1].first : a[0].first + XT_PI;
      wv[a[p].second] = (float)(n == 1 ? XT_PI : 0.5 * (hi - lo));
    }
  }
  XT_TRY(dev_alloc(&s->fbp_w, g.N_p));
  XT_HIP(hipMemcpy(s->fbp_w, wv.data(), g.N_p * sizeof(float), hipMemcpyHostToDevice));
  return XT_OK;
}

// A host [N_p][N_r][N_t] array into dst in the HBM layout (through e when the
// layout is split by parity; e is set afterwards by the resync).
static int upload_sino(xt_icd_state *s, float *dst, const Real_t *host) {
  if (!s->g.pm) {
    XT_HIP(real_to_dev(dst, host, s->nsino));
    return XT_OK;
  }
  XT_HIP(real_to_dev(s->e, host, s->nsino));
  hipLaunchKernelGGL(k_sino_layout, dim3(grid_1d(s->nsino)), dim3(256), 0, s->st, s->g, s->e, dst,
                     s->nsino, 1);
  XT_HIP(hipGetLastError());
  XT_HIP(hipStreamSynchronize(s->st));
  return XT_OK;
}

static int upload_object(xt_icd_state *s, const ScannedObject *O, float *x_dev) {
  const XtGeom &g = s->g;
  size_t per = (size_t)g.N_z * g.N_y * g.N_x;
  float *tmp = nullptr;
  XT_TRY(dev_alloc(&tmp, per * g.N_time));
  for (int f = 0; f < g.N_time; f++) {
    hipError_t e = real_to_dev(tmp + per * f, &O->Object[f][1][0][0], per);
    if (e != hipSuccess) {
      (void)hipFree(tmp);
      return set_err(XT_E_HIP, "object upload: %s", hipGetErrorString(e));
    }
  }
  hipLaunchKernelGGL(k_object_to_device, dim3(grid_1d(per * g.N_time)), dim3(256), 0, s->st, g, tmp,
                     x_dev);
  hipError_t e1 = hipStreamSynchronize(s->st);
  (void)hipFree(tmp);
  if (e1 != hipSuccess) return set_err(XT_E_HIP, "object layout: %s", hipGetErrorString(e1));
  return XT_OK;
}

static int download_object(xt_icd_state *s, ScannedObject *O) {
  const XtGeom &g = s->g;
  size_t per = (size_t)g.N_z * g.N_y * g.N_x;
  float *tmp = nullptr;
  XT_TRY(dev_alloc(&tmp, per * g.N_time));
  hipLaunchKernelGGL(k_object_to_host, dim3(grid_1d(per * g.N_time)), dim3(256), 0, s->st, g, s->x,
                     tmp);
  hipError_t e = hipStreamSynchronize(s->st);
  for (int f = 0; f < g.N_time && e == hipSuccess; f++)
    e = dev_to_real(&O->Object[f][1][0][0], tmp + per * f, per);
  (void)hipFree(tmp);
  if (e != hipSuccess) return set_err(XT_E_HIP, "object download: %s", hipGetErrorString(e));
  return XT_OK;
}

static int init_device(int *dev) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return set_err(XT_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  *dev = (int)g_opt[XT_OPT_DEVICE];
  if (*dev < 0 || *dev >= n) return set_err(XT_E_ARG, "device %d of %d", *dev, n);
  XT_HIP(hipSetDevice(*dev));
  return XT_OK;
}

static XtFpTable fp_table(const xt_icd_state *s) {
  return XtFpTable{s->at_vals, s->at_first, s->vpos};
}

static int icd_resync(xt_icd_state *s) {
  return launch_forward(s->st, s->ZL, s->g, s->x, s->Hr, s->cosv, s->sinv, s->view_frame, 1,
                        nullptr, s->y, s->d, s->e, s->pc_tmp, fp_table(s));
}

template <int ZL>
static void fbp_bp_launch(xt_icd_state *s, const float *q) {
  constexpr int G = 256 / ZL;
  dim3 grid((unsigned)((s->NYX + G - 1) / G), s->g.N_time);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fbp_backproject<ZL>), grid, dim3(256), 0, s->st, s->g, q,
                     s->fbp_w, s->cosv, s->sinv, s->views, s->proj_num, s->line_mask, s->x);
}

// Launch-shape autotuner: times one homogeneous sub-iteration (its two parity
// launches) with W = cap, cap/2 and cap/4 workgroups per frame on the real
// data, restoring the reconstruction state after each trial, and keeps a
// smaller W only when it is more than 3% faster than the best larger one (the
// NHICD launches, which cannot be tried before the first iteration, gain from
// more lines in flight).
static int icd_autotune(xt_icd_state *s) {
  // (lines in flight, threads per workgroup); the first is the default
  std::vector<std::pair<int, int>> cand;
  const int nt_opt = (int)g_opt[XT_OPT_LINE_THREADS];
  for (int nt : {256, 64}) {
    if ((nt_opt && nt != nt_opt) || (s->g.pc && nt != 256)) continue;
    int k = 0;
    for (int w = s->conc; w >= 16 && k < 3; w /= 2, k++) cand.push_back({w, nt});
  }
  if (cand.size() < 2) return XT_OK;
  const XtGeom &g = s->g;
  const size_t nmag = (size_t)g.N_time * g.nb * s->NYX;
  float *e0 = nullptr, *x0 = nullptr, *m0 = nullptr;
  int rc = dev_alloc(&e0, s->nsino);
  if (!rc) rc = dev_alloc(&x0, s->nobj);
  if (!rc) rc = dev_alloc(&m0, nmag);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (!rc && (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess))
    rc = set_err(XT_E_HIP, "autotune events");
  auto copy = [&](float *dst, const float *src, size_t n) {
    return hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, s->st) == hipSuccess;
  };
  if (!rc && !(copy(e0, s->e, s->nsino) && copy(x0, s->x, s->nobj) && copy(m0, s->mag_blk, nmag)))
    rc = set_err(XT_E_HIP, "autotune snapshot");
  const int share = s->n_masked / s->K;
  float best_ms = 1e30f;
  std::pair<int, int> best = cand[0];
  const int saved_timing = s->timing;
  s->timing = 0;
  for (size_t i = 0; !rc && i < cand.size(); i++) {
    s->conc = cand[i].first;
    s->nt = cand[i].second;
    XtIcdLaunch L;
    L.mode = 0;
    L.sub_iter = 0;
    L.K = s->K;
    L.seed = (uint32_t)(int64_t)g_opt[XT_OPT_SEED];
    L.count_stats = 0;
    L.zorder = 0;
    if (hipEventRecord(t0, s->st) != hipSuccess) rc = set_err(XT_E_HIP, "autotune record");
    for (int h = 0; !rc && h < 2; h++) {
      L.half = h;
      rc = launch_icd(s, L, share < s->conc ? share : s->conc);
    }
    float ms = 0.f;
    if (!rc && (hipEventRecord(t1, s->st) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                hipEventElapsedTime(&ms, t0, t1) != hipSuccess))
      rc = set_err(XT_E_HIP, "autotune timing");
    if (!rc && !(copy(s->e, e0, s->nsino) && copy(s->x, x0, s->nobj) && copy(s->mag_blk, m0, nmag)))
      rc = set_err(XT_E_HIP, "autotune restore");
    dbg(s->T, "autotune: %d lines in flight per frame, %d threads each: %.3f ms\n",
        cand[i].first, cand[i].second, ms);
    if (!rc && ms < best_ms * 0.97f) {
      best_ms = ms;
      best = cand[i];
    }
  }
  s->timing = saved_timing;
  s->launches = 0;
  s->conc = best.first;
  s->nt = best.second;
  if (!rc && hipStreamSynchronize(s->st) != hipSuccess) rc = set_err(XT_E_HIP, "autotune sync");
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  (void)hipFree(e0);
  (void)hipFree(x0);
  (void)hipFree(m0);
  return rc;
}

// x = filtered backprojection of y - d, then e = y - Ax - d.  When the ring
// offsets are estimated, d first gets a guess from the view-averaged sinogram,
// so the rings go into d rather than into the starting object.
static int select_edges(xt_icd_state *s);

static int icd_fbp(xt_icd_state *s, bool resync = true) {
  const XtGeom &g = s->g;
  if (g.pc) return set_err(XT_E_ARG, "the filtered backprojection start is for absorption data");
  float *q = nullptr;
  XT_TRY(dev_alloc(&q, s->nsino));
  if (s->update_offset > 0) {
    const int n = g.N_r * g.N_t;
    hipLaunchKernelGGL(k_offset_guess_mean, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->st,
                       g, s->y, q);
    hipLaunchKernelGGL(k_offset_guess_apply, dim3(1), dim3(1024), 0, s->st, g, q, s->zero_mean,
                       s->d);
  }
  const size_t lds = (size_t)(2 * 4 * g.N_r + (g.N_r + 1) / 2) * sizeof(float);
  if (lds > 64 * 1024) {
    (void)hipFree(q);
    return set_err(XT_E_ARG, "N_r=%d too large for the FBP filter", g.N_r);
  }
  hipLaunchKernelGGL(k_fbp_filter, dim3((unsigned)((g.N_t + 3) / 4), g.N_p), dim3(256), lds, s->st,
                     g, s->y, s->d, q);
  switch (s->ZL) {
    case 8: fbp_bp_launch<8>(s, q); break;
    case 16: fbp_bp_launch<16>(s, q); break;
    case 32: fbp_bp_launch<32>(s, q); break;
    default: fbp_bp_launch<64>(s, q); break;
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  (void)hipFree(q);
  if (e != hipSuccess) return set_err(XT_E_HIP, "FBP: %s", hipGetErrorString(e));
  if (s->edge && resync) XT_TRY(select_edges(s));
  return resync ? icd_resync(s) : XT_OK;
}

// The A-column table is used when it fits in half of the free device memory
// (XT_OPT_ATABLE = 1, the default) and every column spans at most four bins.
static int build_atable(xt_icd_state *s) {
  if (g_opt[XT_OPT_ATABLE] == 0 || s->g.wide || s->g.pc) return XT_OK;
  const XtGeom &g = s->g;
  const size_t n = (size_t)g.N_time * s->NYX * g.max_views;
  size_t free_b = 0, total_b = 0;
  XT_HIP(hipMemGetInfo(&free_b, &total_b));
  if (n * (sizeof(float4) + sizeof(int)) > free_b / 2) return XT_OK;
  int *ovf = nullptr;
  XT_TRY(dev_alloc(&s->at_vals, n));
  XT_TRY(dev_alloc(&s->at_first, n));
  XT_TRY(dev_alloc(&ovf, 1));
  XT_HIP(hipMemsetAsync(ovf, 0, sizeof(int), s->st));
  dim3 grid((unsigned)((s->NYX * g.max_views + 255) / 256), g.N_time);
  hipLaunchKernelGGL(k_build_atable, grid, dim3(256), 0, s->st, g, s->Hr, s->cosv, s->sinv, s->views,
                     s->proj_num, s->at_vals, s->at_first, ovf);
  int h_ovf = 0;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&h_ovf, ovf, sizeof(int), hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  (void)hipFree(ovf);
  if (e != hipSuccess) return set_err(XT_E_HIP, "A-column table: %s", hipGetErrorString(e));
  if (h_ovf) {
    (void)hipFree(s->at_vals);
    (void)hipFree(s->at_first);
    s->at_vals = nullptr;
    s->at_first = nullptr;
  }
  return XT_OK;
}

// Edge-detection constants of F/XT_ImageProc.c:85-101,122-124: the 5x5 Gaussian
// (variance 3) in float, normalised by its float sum, and the object circle.
static XtEdgeParams edge_params(const ScannedObject *O, const TomoInputs *T) {
  XtEdgeParams P{};
  const float var = 3.0f;
  float sum = 0.f;
  for (int i = -2; i <= 2; i++)
    for (int j = -2; j <= 2; j++) {
      P.hblur[(i + 2) * 5 + (j + 2)] = (float)exp((double)(-(i * i + j * j) / (2 * var)));
      sum += P.hblur[(i + 2) * 5 + (j + 2)];
    }
  for (int k = 0; k < 25; k++) P.hblur[k] /= sum;
  P.half_x = -O->Length_X / 2;
  P.half_y = -O->Length_Y / 2;
  P.delta_xy = O->delta_xy;
  P.r2 = T->radius_obj * T->radius_obj;
  return P;
}

// The edge-only voxel mask of the object x into mask (device layout, which is
// also the reference's Mask[N_time][N_y][N_x][N_z]).
static int launch_edge_mask(xt_icd_state *s, const float *x, uint8_t *mask) {
  float *blur = nullptr;
  uint8_t *edge = nullptr;
  XT_TRY(dev_alloc(&blur, s->nobj));
  int rc = dev_alloc(&edge, s->nobj);
  if (rc == XT_OK) {
    const dim3 grid(grid_1d(s->nobj)), block(256);
    hipLaunchKernelGGL(k_edge_blur, grid, block, 0, s->st, s->g, s->edge_par, x, blur);
    hipLaunchKernelGGL(k_edge_gradient, grid, block, 0, s->st, s->g, s->edge_par, blur, edge);
    hipLaunchKernelGGL(k_edge_dilate, grid, block, 0, s->st, s->g, edge, mask);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) rc = set_err(XT_E_HIP, "edge mask: %s", hipGetErrorString(e));
  }
  (void)hipFree(blur);
  if (edge) (void)hipFree(edge);
  return rc;
}

static void drop_graph(xt_icd_state *s) {
  if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->graph_exec = nullptr;
  s->graph = nullptr;
}

// Edge-only updates, F/XT_ICD_update.c:996-997 with F/XT_ImageProc.c:151-170 and
// the line selection of F/XT_ICD_update.c:599-631: the voxel mask of the current
// object; the line list becomes every line with a masked voxel (in any frame),
// and each (frame, z-block)'s NHICD share is its own count of such lines over
// NHICD_Iterations.  The line kernel skips unmasked voxels.
static int select_edges(xt_icd_state *s) {
  const XtGeom &g = s->g;
  if (!s->vmask) {
    XT_TRY(dev_alloc(&s->vmask, s->nobj));
    // the list may grow beyond the radius mask (the dilation reaches outside it)
    if (s->lines) (void)hipFree(s->lines);
    s->lines = nullptr;
    XT_TRY(dev_alloc(&s->lines, s->NYX));
  }
  XT_TRY(launch_edge_mask(s, s->x, s->vmask));
  std::vector<uint8_t> m(s->nobj);
  XT_HIP(hipMemcpy(m.data(), s->vmask, s->nobj, hipMemcpyDeviceToHost));
  // per (frame, z-block) map: that block's lines with a masked voxel; per (frame,
  // parity) map (XT_OPT_NHICD_LINES = 1, where a selected line is updated in every
  // block of the parity): the lines with a masked voxel in some block of the parity
  const int nb = g.nb;
  auto zblock = [&](int z) {  // the line kernel's block of slice z
    const int bs = g.N_z / g.nb;
    const int b = bs > 0 ? z / bs : g.nb - 1;
    return b >= g.nb ? g.nb - 1 : b;
  };
  std::vector<int> lines, ns((size_t)g.N_time * (nb > 2 ? nb : 2), 0), np((size_t)g.N_time * 2, 0);
  std::vector<uint8_t> has(nb);
  for (size_t c = 0; c < s->NYX; c++) {
    bool any = false;
    for (int f = 0; f < g.N_time; f++) {
      const uint8_t *v = &m[((size_t)f * s->NYX + c) * g.N_z];
      std::fill(has.begin(), has.end(), 0);
      for (int z = 0; z < g.N_z; z++)
        if (v[z]) has[zblock(z)] = 1;
      bool par_has[2] = {false, false};
      for (int b = 0; b < nb; b++) {
        ns[(size_t)f * nb + b] += has[b];
        par_has[b & 1] |= has[b] != 0;
        any |= has[b] != 0;
      }
      np[(size_t)f * 2] += par_has[0];
      np[(size_t)f * 2 + 1] += par_has[1];
    }
    if (any) lines.push_back((int)c);
  }
  for (int &n : ns) n /= s->K;
  for (int &n : np) n /= s->K;
  s->n_masked = (int)lines.size();
  if (!lines.empty())
    XT_HIP(hipMemcpy(s->lines, lines.data(), lines.size() * sizeof(int), hipMemcpyHostToDevice));
  XT_TRY(build_groups(s, lines));
  XT_HIP(hipMemcpy(s->nsel, ns.data(), ns.size() * sizeof(int), hipMemcpyHostToDevice));
  XT_HIP(hipMemcpy(s->nsel_par, np.data(), np.size() * sizeof(int), hipMemcpyHostToDevice));
  drop_graph(s);
  s->lines_gen++;
  return XT_OK;
}

static int create_impl(Sinogram *S, ScannedObject *O, TomoInputs *T, xt_icd_state **out,
                       const xt_icd_state *src = nullptr) {
  *out = nullptr;
  xt_icd_state *s = new xt_icd_state();
  auto fail = [&](int rc) {
    free_state(s);
    return rc;
  };
  // the geometry is checked before any device call, so a bad one is refused the
  // same way with or without a GPU
  int rc = make_geom(S, O, T, &s->g);
  if (rc) return fail(rc);
  rc = init_device(&s->dev);
  if (rc) return fail(rc);
  if (hipStreamCreate(&s->st) != hipSuccess) return fail(set_err(XT_E_HIP, "hipStreamCreate"));
  XtGeom &g = s->g;
  s->ZL = zl_for(g.N_z);
  s->T = T;
  s->K = O->NHICD_Iterations > 0 ? O->NHICD_Iterations : 1;
  s->NumIter = T->NumIter;
  s->edge = T->only_Edge_Updates != 0;
  s->edge_par = edge_params(O, T);
  s->update_offset = T->updateProjOffset;
  s->zero_mean = T->EnforceZeroMeanOffset;
  s->update_var = T->updateVar;
  s->no_nhicd = T->no_NHICD;
  s->var_est = T->var_est;
  s->stop_threshold = T->StopThreshold;
  if (icd_lds_bytes(g, s->ZL, false) > 64 * 1024)
    return fail(set_err(XT_E_ARG, "%d views in one frame exceed the LDS footprint cache",
                        g.max_views));
  s->nsino = (size_t)g.N_p * g.N_r * g.N_t;
  s->NYX = (size_t)g.N_y * g.N_x;
  s->nobj = (size_t)g.N_time * s->NYX * g.N_z;
  // masked lines, F/XT_ICD_update.c:958-974
  std::vector<int> masked;
  for (int j = 0; j < g.N_y; j++)
    for (int k = 0; k < g.N_x; k++) {
      Real_t xc = (Real_t)((double)O->x0 + ((double)(Real_t)k + 0.5) * (double)O->delta_xy);
      Real_t yc = (Real_t)((double)O->y0 + ((double)(Real_t)j + 0.5) * (double)O->delta_xy);
      if (xc * xc + yc * yc < T->radius_obj * T->radius_obj) masked.push_back(j * g.N_x + k);
    }
  s->n_masked = (int)masked.size();
  // split the sinogram by z-block parity when the passes can pack their lanes
  if (g_opt[XT_OPT_PACK] != 0 && !g.pc && g.zo == 1 && g.nb % 2 == 0 && g.N_z % g.nb == 0 &&
      g.N_t % 2 == 0) {
    g.pm = 1;
    g.bs = g.N_z / g.nb;
    g.H = g.N_t / 2;
  }
  if ((rc = setup_lines(s, masked))) return fail(rc);
  if ((rc = upload_views(s, S, O))) return fail(rc);
  if (g.pc && (rc = dev_alloc(&s->pc_tmp, s->nsino))) return fail(rc);
  if ((rc = dev_alloc(&s->e, s->nsino)) || (rc = dev_alloc(&s->w, s->nsino)) ||
      (rc = dev_alloc(&s->y, s->nsino)) || (rc = dev_alloc(&s->d, (size_t)g.N_r * g.N_t)) ||
      (rc = dev_alloc(&s->x, s->nobj)) || (rc = dev_alloc(&s->Hr, (size_t)g.N_p * XT_HR_STRIDE)) ||
      (rc = dev_alloc(&s->nhicd_flag, (size_t)g.N_time * 2 * s->NYX)) ||
      (rc = dev_alloc(&s->nhicd_blk, (size_t)g.N_time * g.nb * s->NYX)) ||
      (rc = dev_alloc(&s->mag_blk, (size_t)g.N_time * g.nb * s->NYX)) ||
      (rc = dev_alloc(&s->magsum, (size_t)g.N_time * s->NYX)) ||
      (rc = dev_alloc(&s->filt, (size_t)g.N_time * (g.nb > 2 ? g.nb : 2) * s->NYX)) ||
      (rc = dev_alloc(&s->magpar, (size_t)g.N_time * 2 * s->NYX)) ||
      (rc = dev_alloc(&s->nsel, g.N_time * (g.nb > 2 ? g.nb : 2))) ||
      (rc = dev_alloc(&s->nsel_par, g.N_time * 2)) ||
      (rc = dev_alloc(&s->nsel_out, g.N_time * (g.nb > 2 ? g.nb : 2))) ||
      (rc = dev_alloc(&s->vsc_st, (size_t)4 * g.N_time * (g.nb > 2 ? g.nb : 2))) ||
      (rc = dev_alloc(&s->vsc_hist, (size_t)256 * g.N_time * (g.nb > 2 ? g.nb : 2))) ||
      (rc = dev_alloc(&s->stats, 8)) || (rc = dev_alloc(&s->red_wg, 8192)) ||
      (rc = dev_alloc(&s->scal, 8)) || (rc = dev_alloc(&s->iter_dev, 1)))
    return fail(rc);
  if (hipMemset(s->iter_dev, 0, sizeof(int)) != hipSuccess)
    return fail(set_err(XT_E_HIP, "iteration counter clear failed"));
  // ring-offset scratch: view chunks so that N_r*N_t*chunks >= ~256k threads
  {
    const int NRT = g.N_r * g.N_t;
    int nch = (262144 + NRT - 1) / NRT;
    if (nch > g.N_p) nch = g.N_p;
    if (nch < 1) nch = 1;
    s->off_chunk = (g.N_p + nch - 1) / nch;
    s->off_nch = (g.N_p + s->off_chunk - 1) / s->off_chunk;
    s->off_nblk = (NRT + 255) / 256;
    if ((rc = dev_alloc(&s->off_part, (size_t)s->off_nch * 2 * NRT)) ||
        (rc = dev_alloc(&s->off_numden, (size_t)2 * NRT)) ||
        (rc = dev_alloc(&s->off_wg, (size_t)2 * s->off_nblk)))
      return fail(rc);
  }
  if (src) {
    // y and w of a state over the same sinogram, already in the HBM layout
    const XtGeom &h = src->g;
    if (src->dev != s->dev || h.N_p != g.N_p || h.N_r != g.N_r || h.N_t != g.N_t ||
        h.N_time != g.N_time || h.pm != g.pm || (g.pm && (h.bs != g.bs || h.H != g.H)) ||
        src->nsino != s->nsino)
      return fail(set_err(XT_E_ARG, "sinogram source: another device, size or layout"));
    if (hipStreamSynchronize(src->st) != hipSuccess ||
        hipMemcpyAsync(s->y, src->y, s->nsino * sizeof(float), hipMemcpyDeviceToDevice, s->st) !=
            hipSuccess ||
        hipMemcpyAsync(s->w, src->w, s->nsino * sizeof(float), hipMemcpyDeviceToDevice, s->st) !=
            hipSuccess ||
        hipStreamSynchronize(s->st) != hipSuccess)
      return fail(set_err(XT_E_HIP, "sinogram copy failed"));
  } else {
    // the host's y and w, put into the HBM layout through e (set later by the resync)
    if ((rc = upload_sino(s, s->y, &S->Projection[0][0][0])) ||
        (rc = upload_sino(s, s->w, &T->Weight[0][0][0])))
      return fail(rc);
  }
  if (real_to_dev(s->d, &S->ProjOffset[0][0], (size_t)g.N_r * g.N_t) != hipSuccess)
    return fail(set_err(XT_E_HIP, "sinogram upload failed"));
  if (hipMemset(s->mag_blk, 0, (size_t)g.N_time * g.nb * s->NYX * sizeof(float)) != hipSuccess ||
      hipMemset(s->magsum, 0, (size_t)g.N_time * s->NYX * sizeof(float)) != hipSuccess ||
      hipMemset(s->nhicd_flag, 0, (size_t)g.N_time * 2 * s->NYX) != hipSuccess ||
      hipMemset(s->nhicd_blk, 0, (size_t)g.N_time * g.nb * s->NYX) != hipSuccess)
    return fail(set_err(XT_E_HIP, "scratch clear failed"));
  {
    // NHICD lines per (frame, z-block): floor(UpdateSelectNum / NHICD_Iterations),
    // F/XT_NHICD.c:182 (every block holds all masked lines)
    std::vector<int> ns((size_t)g.N_time * (g.nb > 2 ? g.nb : 2), s->n_masked / s->K);
    if (hipMemcpy(s->nsel, ns.data(), ns.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->nsel_par, ns.data(), (size_t)g.N_time * 2 * sizeof(int),
                  hipMemcpyHostToDevice) != hipSuccess)
      return fail(set_err(XT_E_HIP, "nsel upload"));
  }
  if ((rc = upload_object(s, O, s->x))) return fail(rc);
  if (s->edge && (rc = select_edges(s))) return fail(rc);
  if ((rc = launch_detector_response(s, O->BeamWidth))) return fail(rc);
  if ((rc = build_atable(s))) return fail(rc);
  {
    // the line kernel stages every view's A column of a frame in LDS
    int lds_max = 0;
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, s->dev);
    const size_t need = icd_lds_bytes(g, s->ZL, s->at_vals != nullptr);
    if (lds_max > 0 && need > (size_t)lds_max)
      return fail(set_err(XT_E_ARG, "%d views per frame need %zu B of LDS per workgroup (device "
                          "limit %d)", g.max_views, need, lds_max));
  }
  if ((rc = icd_resync(s))) return fail(rc);
  if (hipStreamSynchronize(s->st) != hipSuccess)
    return fail(set_err(XT_E_HIP, "initialisation kernels failed"));
  s->nt = ((int)g_opt[XT_OPT_LINE_THREADS] == 64 && !g.pc) ? 64 : 256;
  if (g_opt[XT_OPT_AUTOTUNE] != 0 && g_opt[XT_OPT_LINES] == 0 && (rc = icd_autotune(s)))
    return fail(rc);
  *out = s;
  return XT_OK;
}

extern "C" int xt_icd_create(Sinogram *S, ScannedObject *O, TomoInputs *T, xt_icd_state **state) {
  if (!S || !O || !T || !state) return set_err(XT_E_ARG, "null argument");
  return create_impl(S, O, T, state);
}

extern "C" int xt_icd_create_from(Sinogram *S, ScannedObject *O, TomoInputs *T,
                                  const xt_icd_state *sino_from, xt_icd_state **state) {
  if (!S || !O || !T || !sino_from || !state) return set_err(XT_E_ARG, "null argument");
  return create_impl(S, O, T, state, sino_from);
}

extern "C" int xt_icd_set_timing(xt_icd_state *s, int32_t on) {
  if (!s) return set_err(XT_E_ARG, "null state");
  s->timing = on;
  return XT_OK;
}

static int collect_timing(xt_icd_state *s) {
  if (!s->timing) return XT_OK;
  for (int i = 0; i + 1 < s->ev_used; i += 2) {
    float ms = 0;
    XT_HIP(hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]));
    s->icd_ms += ms;
  }
  s->ev_used = 0;
  return XT_OK;
}

extern "C" int xt_icd_cost(xt_icd_state *s, double *cost) {
  if (!s || !cost) return set_err(XT_E_ARG, "null argument");
  int nb = grid_1d(s->nsino, 256, 4096);
  hipLaunchKernelGGL(k_cost_data, dim3(nb), dim3(256), 0, s->st, s->g, s->e, s->w, s->nsino,
                     s->red_wg);
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, s->st, s->red_wg, nb, s->scal + 2);
  int nbp = grid_1d(s->nobj, 256, 4096);
  hipLaunchKernelGGL(k_cost_prior, dim3(nbp), dim3(256), 0, s->st, s->g, s->x, s->red_wg);
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, s->st, s->red_wg, nbp, s->scal + 3);
  XT_HIP(hipGetLastError());
  double v[2];
  XT_HIP(hipMemcpyAsync(v, s->scal + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  XT_HIP(hipStreamSynchronize(s->st));
  const double n = (double)s->g.N_p * s->g.N_r * s->g.N_t;
  *cost = v[0] + n * log((double)s->var_est) / 2 + v[1];
  return XT_OK;
}

// The kernels of one iteration, enqueued on s->st (eagerly, or into a graph
// under capture): the schedule of F/XT_ICD_update.c:794-882.
static int enqueue_iteration(xt_icd_state *s, bool nhicd) {
  hipLaunchKernelGGL(k_begin_iteration, dim3(1), dim3(64), 0, s->st, s->iter_dev, s->stats, 8);
  XtIcdLaunch L;
  L.zorder = 0;
  L.K = s->K;
  L.seed = (uint32_t)(int64_t)g_opt[XT_OPT_SEED];
  // per sub-iteration: the even and the odd z-block pass over the homogeneous
  // share, then over the NHICD lines, each followed by a ring-offset update
  // (F/XT_ICD_update.c:799-881)
  const int per = s->n_masked / s->K;
  for (int j = 0; j < s->K; j++) {
    L.sub_iter = j;
    L.mode = 0;
    L.count_stats = 1;
    const int share = (j == s->K - 1) ? s->n_masked - j * per : per;
    for (int h = 0; h < 2; h++) {
      L.half = h;
      XT_TRY(launch_icd(s, L, share < s->conc ? share : s->conc));
      if (s->update_offset > 1) XT_TRY(offset_update(s));
    }
    if (nhicd) {
      XT_TRY(vsc_select(s));
      L.mode = 1;
      L.count_stats = 0;
      for (int h = 0; h < 2; h++) {
        L.half = h;
        XT_TRY(launch_icd(s, L, s->conc));
        if (s->update_offset > 1) XT_TRY(offset_update(s));
      }
    }
  }
  return XT_OK;
}

// Stop statistic of the iteration, F/XT_ICD_update.c:887-925: sum of |dx| over the
// lines into scal[4] (graph-capturable).
static int enqueue_stop_stats(xt_icd_state *s) {
  int nb = grid_1d((size_t)s->g.N_time * s->NYX, 256, 4096);
  hipLaunchKernelGGL(k_sum_f32, dim3(nb), dim3(256), 0, s->st, s->magsum,
                     (size_t)s->g.N_time * s->NYX, 0, s->red_wg);
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, s->st, s->red_wg, nb, s->scal + 4);
  XT_HIP(hipGetLastError());
  return XT_OK;
}

static bool resync_due(int Iter) {
  const int every = (int)g_opt[XT_OPT_RESYNC_EVERY];
  return every > 0 && Iter % every == 0;
}

// Host side of the end of an iteration: read the statistics back, decide the
// stop test, and run what is due and was not already enqueued.
static int iteration_host(xt_icd_state *s, int Iter, bool resynced) {
  double st[4], avg = 0;
  XT_HIP(hipMemcpyAsync(st, s->stats, 4 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  XT_HIP(hipMemcpyAsync(&avg, s->scal + 4, sizeof(double), hipMemcpyDeviceToHost, s->st));
  XT_HIP(hipStreamSynchronize(s->st));
  XT_TRY(collect_timing(s));
  s->total_updates += st[1];
  s->total_foot += st[2];
  s->total_visits += st[3];
  s->stop_pct = 100.0 * avg / st[0];
  s->stopped = (s->stop_pct < (double)s->stop_threshold) ? 1 : 0;
  if (!resynced && resync_due(Iter)) XT_TRY(icd_resync(s));
  if (g_opt[XT_OPT_COST] != 0) XT_TRY(xt_icd_cost(s, &s->last_cost));
  dbg(s->T, "ICD iteration %d: updates %.0f in %.0f line visits, average update %.6f%%, cost %.9g\n",
      Iter, st[1], st[3], s->stop_pct, s->last_cost);
  return XT_OK;
}

static int iterate_one(xt_icd_state *s) {
  const int Iter = ++s->iter;
  const bool nhicd = Iter > 1 && !s->no_nhicd;
  // iterations from the second on are all the same launch sequence: capture it
  // once and replay it (no memset nodes; the iteration number lives on the device)
  // (per-launch timing needs eager launches: HIP does not report elapsed times
  // of events recorded inside a graph replay)
  const bool use_graph = g_opt[XT_OPT_GRAPH] != 0 && Iter > 1 && !s->timing;
  if (use_graph) {
    if (s->graph_exec && s->graph_opt_gen != g_opt_gen) drop_graph(s);
    if (!s->graph_exec) {
      s->graph_opt_gen = g_opt_gen;
      XT_HIP(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
      const int64_t l0 = s->launches;
      int rc = enqueue_iteration(s, nhicd);
      s->graph_launches = s->launches - l0;
      s->launches = l0;
      hipGraph_t gr = nullptr;
      hipError_t e = hipStreamEndCapture(s->st, &gr);
      if (rc != XT_OK) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
      }
      if (e != hipSuccess) return set_err(XT_E_HIP, "graph capture: %s", hipGetErrorString(e));
      s->graph = gr;
      XT_HIP(hipGraphInstantiate(&s->graph_exec, s->graph, nullptr, nullptr, 0));
    }
    XT_HIP(hipGraphLaunch(s->graph_exec, s->st));
    s->launches += s->graph_launches;
  } else {
    XT_TRY(enqueue_iteration(s, nhicd));
  }
  if (s->update_var == 1) XT_TRY(variance_update(s));
  XT_TRY(enqueue_stop_stats(s));
  return iteration_host(s, Iter, false);
}

extern "C" int xt_icd_iterate(xt_icd_state *s, int32_t n_iter, int32_t *stop) {
  if (!s) return set_err(XT_E_ARG, "null state");
  for (int i = 0; i < n_iter; i++) XT_TRY(iterate_one(s));
  if (stop) *stop = s->stopped;
  return XT_OK;
}

extern "C" int xt_icd_get_stats(xt_icd_state *s, xt_icd_stats *out) {
  if (!s || !out) return set_err(XT_E_ARG, "null argument");
  out->iterations = s->iter;
  out->voxel_updates = s->total_updates;
  out->footprint_entries = s->total_foot;
  out->icd_kernel_ms = s->icd_ms;
  out->icd_launches = s->launches;
  out->last_cost = s->last_cost;
  out->stop_percentage = s->stop_pct;
  out->stopped = s->stopped;
  out->line_visits = s->total_visits;
  out->lines_in_flight = s->conc;
  out->line_threads = s->nt;
  return XT_OK;
}

extern "C" int xt_icd_resync(xt_icd_state *s) {
  if (!s) return set_err(XT_E_ARG, "null state");
  XT_TRY(icd_resync(s));
  XT_HIP(hipStreamSynchronize(s->st));
  return XT_OK;
}

extern "C" int xt_icd_fbp(xt_icd_state *s) {
  if (!s) return set_err(XT_E_ARG, "null state");
  XT_TRY(icd_fbp(s));
  XT_HIP(hipStreamSynchronize(s->st));
  return XT_OK;
}

extern "C" int xt_icd_upload(xt_icd_state *s, Sinogram *S, ScannedObject *O, TomoInputs *T) {
  (void)T;
  if (!s || !S || !O) return set_err(XT_E_ARG, "null argument");
  if (O->N_time != s->g.N_time || O->N_z != s->g.N_z || O->N_y != s->g.N_y ||
      O->N_x != s->g.N_x || S->N_r != s->g.N_r || S->N_t != s->g.N_t)
    return set_err(XT_E_ARG, "structs do not match the state's sizes");
  XT_TRY(upload_object(s, O, s->x));
  if (s->edge) XT_TRY(select_edges(s));
  XT_HIP(real_to_dev(s->d, &S->ProjOffset[0][0], (size_t)s->g.N_r * s->g.N_t));
  XT_TRY(icd_resync(s));
  XT_HIP(hipStreamSynchronize(s->st));
  return XT_OK;
}

extern "C" int xt_icd_unet_start(xt_icd_state *s, xt_unet *net) {
  if (!s || !net) return set_err(XT_E_ARG, "null argument");
  int d0, d1, d2;
  xt_unet_shape(net, &d0, &d1, &d2);
  const XtGeom &g = s->g;
  if (g.N_time != 1 || d0 != g.N_y || d1 != g.N_x || d2 != g.N_z)
    return set_err(XT_E_ARG, "network volume %dx%dx%d does not match the object %dx%dx%d x %d frames "
                   "(one frame of N_y x N_x x N_z required)", d0, d1, d2, g.N_y, g.N_x, g.N_z,
                   g.N_time);
  XT_TRY(icd_fbp(s, false));
  XT_TRY(xt_unet_forward_on(net, s->x, s->x, s->st));
  if (g.positivity)
    hipLaunchKernelGGL(k_clamp_nonneg, dim3(1024), dim3(256), 0, s->st, s->x, s->nobj);
  XT_HIP(hipGetLastError());
  if (s->edge) XT_TRY(select_edges(s));
  XT_TRY(icd_resync(s));
  XT_HIP(hipStreamSynchronize(s->st));
  return XT_OK;
}

extern "C" int xt_icd_synchronize(xt_icd_state *s) {
  if (!s) return set_err(XT_E_ARG, "null state");
  XT_HIP(hipStreamSynchronize(s->st));
  return XT_OK;
}

extern "C" int xt_icd_download(xt_icd_state *s, Sinogram *S, ScannedObject *O, TomoInputs *T) {
  if (!s || !S || !O || !T) return set_err(XT_E_ARG, "null argument");
  XT_HIP(hipStreamSynchronize(s->st));
  XT_TRY(download_object(s, O));
  XT_HIP(dev_to_real(&S->ProjOffset[0][0], s->d, (size_t)s->g.N_r * s->g.N_t));
  if (s->update_var == 1) {
    if (s->g.pm) {
      float *tmp = nullptr;
      XT_TRY(dev_alloc(&tmp, s->nsino));
      hipLaunchKernelGGL(k_sino_layout, dim3(grid_1d(s->nsino)), dim3(256), 0, s->st, s->g, s->w,
                         tmp, s->nsino, 0);
      hipError_t e = hipStreamSynchronize(s->st);
      if (e == hipSuccess)
        e = dev_to_real(&T->Weight[0][0][0], tmp, s->nsino);
      (void)hipFree(tmp);
      if (e != hipSuccess) return set_err(XT_E_HIP, "Weight download: %s", hipGetErrorString(e));
    } else {
      XT_HIP(dev_to_real(&T->Weight[0][0][0], s->w, s->nsino));
    }
    T->var_est = s->var_est;
  }
  uint8_t *sel = nullptr;
  XT_TRY(dev_alloc(&sel, s->nsino));
  hipLaunchKernelGGL(k_select, dim3(grid_1d(s->nsino)), dim3(256), 0, s->st, s->g, s->e, s->w,
                     s->nsino, sel);
  hipError_t e = hipStreamSynchronize(s->st);
  if (e == hipSuccess)
    e = hipMemcpy(&S->ProjSelect[0][0][0], sel, s->nsino, hipMemcpyDeviceToHost);
  (void)hipFree(sel);
  if (e != hipSuccess) return set_err(XT_E_HIP, "ProjSelect download: %s", hipGetErrorString(e));
  return XT_OK;
}

extern "C" void xt_icd_destroy(xt_icd_state *s) { free_state(s); }

// ---------------------------------------------------------------------------
// Batches of independent reconstructions (slabs of any size, each created with
// xt_icd_create).  From the second iteration on, one hipGraph holds every
// slab's iteration as its own branch (each on its slab's stream, forked from
// and joined to the batch's stream), so the slabs' launches overlap on the GPU
// and one graph launch advances the whole batch.
struct xt_icd_batch {
  std::vector<xt_icd_state *> s;
  hipStream_t st = nullptr;
  hipEvent_t fork = nullptr, done = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t graph_launches = 0;
  bool graph_resyncs = false;
  // fused form: slabs grouped by kernel variant; per (pass, group) a table of
  // XtIcdSlab entries in device memory
  bool fused = false;
  struct Group {
    int ZL = 0, nt = 256;
    bool tab = false, pack = false;
    std::vector<int> slabs;
    int rows = 0;
    size_t lds = 0;
  };
  std::vector<Group> groups;
  struct Pass {
    size_t off = 0;  // first entry in tab
    int n = 0, gx_max = 0;
  };
  std::vector<Pass> passes;  // [pass][group], pass = sub_iter * 4 + (nhicd ? 2 : 0) + half
  XtIcdSlab *tab = nullptr;
  // g_opt_gen and the slabs' lines_gen the plan and the graph were made for
  uint64_t opt_gen = 0;
  std::vector<uint64_t> lines_gen;
};

// Groups of slabs that share a kernel variant, and the per-pass tables of the
// fused form.  Slabs with different sub-iteration counts or NHICD settings keep
// the branch form.
static int batch_plan(xt_icd_batch *b) {
  b->fused = false;
  if (g_opt[XT_OPT_BATCH_FUSE] == 0) return XT_OK;
  const xt_icd_state *s0 = b->s[0];
  for (xt_icd_state *s : b->s)
    if (s->K != s0->K || s->no_nhicd != s0->no_nhicd || s->g.wide || s->g.pc) return XT_OK;
  for (size_t i = 0; i < b->s.size(); i++) {
    xt_icd_state *s = b->s[i];
    XtIcdLaunch L0{};
    const bool tab = s->at_vals != nullptr, pack = icd_packable(s->g, s->ZL, L0);
    size_t k = 0;
    while (k < b->groups.size() &&
           !(b->groups[k].ZL == s->ZL && b->groups[k].tab == tab && b->groups[k].pack == pack &&
             b->groups[k].nt == s->nt))
      k++;
    if (k == b->groups.size()) {
      b->groups.emplace_back();
      b->groups[k].ZL = s->ZL;
      b->groups[k].nt = s->nt;
      b->groups[k].tab = tab;
      b->groups[k].pack = pack;
    }
    auto &G = b->groups[k];
    G.slabs.push_back((int)i);
    G.rows += s->g.N_time;
    G.lds = std::max(G.lds, icd_lds_bytes(s->g, s->ZL, tab, s->nt));
  }
  const int K = s0->K, ng = (int)b->groups.size();
  std::vector<XtIcdSlab> host;
  b->passes.assign((size_t)K * 4 * ng, xt_icd_batch::Pass());
  for (int j = 0; j < K; j++)
    for (int m = 0; m < 2; m++)
      for (int h = 0; h < 2; h++)
        for (int k = 0; k < ng; k++) {
          auto &P = b->passes[((size_t)j * 4 + m * 2 + h) * ng + k];
          P.off = host.size();
          int y0 = 0;
          for (int i : b->groups[k].slabs) {
            xt_icd_state *s = b->s[i];
            XtIcdSlab e;
            e.g = s->g;
            e.B = buffers_of(s);
            e.L.mode = m;
            e.L.half = h;
            e.L.sub_iter = j;
            e.L.K = K;
            e.L.seed = (uint32_t)(int64_t)g_opt[XT_OPT_SEED];
            e.L.count_stats = m == 0;
            e.L.zorder = 0;
            const int per = s->n_masked / K;
            const int share = (j == K - 1) ? s->n_masked - j * per : per;
            e.gx = std::max(1, m == 0 ? std::min(share, s->conc) : s->conc);
            e.y0 = y0;
            y0 += s->g.N_time;
            P.gx_max = std::max(P.gx_max, e.gx);
            host.push_back(e);
          }
          P.n = (int)(host.size() - P.off);
        }
  XT_HIP(hipMalloc((void **)&b->tab, host.size() * sizeof(XtIcdSlab)));
  XT_HIP(hipMemcpy(b->tab, host.data(), host.size() * sizeof(XtIcdSlab), hipMemcpyHostToDevice));
  b->fused = true;
  return XT_OK;
}

template <int ZL, int NT>
static void fused_launch_nt(hipStream_t st, const xt_icd_batch::Group &G,
                            const xt_icd_batch::Pass &P, const XtIcdSlab *tab) {
  const dim3 grid(P.gx_max, G.rows), block(NT);
  const XtIcdSlab *t = tab + P.off;
  if (G.tab) {
    if (G.pack)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_lines_multi<ZL, true, true, NT>), grid, block,
                         G.lds, st, t, P.n);
    else
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_lines_multi<ZL, true, false, NT>), grid, block,
                         G.lds, st, t, P.n);
  } else {
    if (G.pack)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_lines_multi<ZL, false, true, NT>), grid, block,
                         G.lds, st, t, P.n);
    else
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_icd_lines_multi<ZL, false, false, NT>), grid, block,
                         G.lds, st, t, P.n);
  }
}

template <int ZL>
static void fused_launch(hipStream_t st, const xt_icd_batch::Group &G,
                         const xt_icd_batch::Pass &P, const XtIcdSlab *tab) {
  if (G.nt == 64)
    fused_launch_nt<ZL, 64>(st, G, P, tab);
  else
    fused_launch_nt<ZL, 256>(st, G, P, tab);
}

static int fused_pass(xt_icd_batch *b, int pass) {
  const int ng = (int)b->groups.size();
  for (int k = 0; k < ng; k++) {
    const auto &G = b->groups[k];
    const auto &P = b->passes[(size_t)pass * ng + k];
    switch (G.ZL) {
      case 8: fused_launch<8>(b->st, G, P, b->tab); break;
      case 16: fused_launch<16>(b->st, G, P, b->tab); break;
      case 32: fused_launch<32>(b->st, G, P, b->tab); break;
      default: fused_launch<64>(b->st, G, P, b->tab); break;
    }
    XT_HIP(hipGetLastError());
  }
  for (xt_icd_state *s : b->s) s->launches++;
  return XT_OK;
}

// One iteration of every slab on the batch stream: the voxel-update passes
// fused across slabs, each slab's ring-offset updates and NHICD selection
// between them, the stop statistic and (resync) the FP64 resync at the end.
// The slabs' streams are pointed at the batch stream meanwhile.
static int enqueue_batch_fused(xt_icd_batch *b, bool nhicd, bool resync) {
  std::vector<hipStream_t> own;
  for (xt_icd_state *s : b->s) {
    own.push_back(s->st);
    s->st = b->st;
  }
  int rc = XT_OK;
  for (xt_icd_state *s : b->s)
    hipLaunchKernelGGL(k_begin_iteration, dim3(1), dim3(64), 0, b->st, s->iter_dev, s->stats, 8);
  const int K = b->s[0]->K;
  for (int j = 0; rc == XT_OK && j < K; j++) {
    for (int h = 0; rc == XT_OK && h < 2; h++) {
      rc = fused_pass(b, j * 4 + h);
      for (xt_icd_state *s : b->s)
        if (rc == XT_OK && s->update_offset > 1) rc = offset_update(s);
    }
    if (!nhicd) continue;
    for (xt_icd_state *s : b->s)
      if (rc == XT_OK) rc = vsc_select(s);
    for (int h = 0; rc == XT_OK && h < 2; h++) {
      rc = fused_pass(b, j * 4 + 2 + h);
      for (xt_icd_state *s : b->s)
        if (rc == XT_OK && s->update_offset > 1) rc = offset_update(s);
    }
  }
  for (xt_icd_state *s : b->s) {
    if (rc == XT_OK) rc = enqueue_stop_stats(s);
    if (rc == XT_OK && resync) rc = icd_resync(s);
  }
  for (size_t i = 0; i < b->s.size(); i++) b->s[i]->st = own[i];
  return rc;
}

extern "C" void xt_icd_batch_destroy(xt_icd_batch *b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->tab) (void)hipFree(b->tab);
  for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

static int batch_refresh(xt_icd_batch *b);

extern "C" int xt_icd_batch_create(int32_t n, xt_icd_state **states, xt_icd_batch **batch) {
  if (!batch || n <= 0 || !states) return set_err(XT_E_ARG, "bad batch arguments");
  *batch = nullptr;
  for (int i = 0; i < n; i++) {
    if (!states[i]) return set_err(XT_E_ARG, "batch slab %d is null", i);
    if (states[i]->dev != states[0]->dev) return set_err(XT_E_ARG, "batch spans devices");
    if (states[i]->iter != states[0]->iter)
      return set_err(XT_E_ARG, "batch slabs at different iterations");
    for (int k = 0; k < i; k++)
      if (states[k] == states[i]) return set_err(XT_E_ARG, "slab %d listed twice", i);
  }
  xt_icd_batch *b = new xt_icd_batch();
  b->s.assign(states, states + n);
  b->join.resize(n, nullptr);
  bool ok = hipStreamCreate(&b->st) == hipSuccess &&
            hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->done, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; ok && i < n; i++)
    ok = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    xt_icd_batch_destroy(b);
    return set_err(XT_E_HIP, "batch streams/events");
  }
  int rc = batch_refresh(b);
  if (rc != XT_OK) {
    xt_icd_batch_destroy(b);
    return rc;
  }
  *batch = b;
  return XT_OK;
}

static int batch_capture(xt_icd_batch *b, const std::vector<int> &iters) {
  XT_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  int rc = XT_OK;
  hipError_t e = hipSuccess;
  b->graph_launches = 0;
  b->graph_resyncs = (int)g_opt[XT_OPT_RESYNC_EVERY] == 1;
  if (b->fused) {
    std::vector<int64_t> l0;
    for (xt_icd_state *s : b->s) l0.push_back(s->launches);
    rc = enqueue_batch_fused(b, iters[0] > 1 && !b->s[0]->no_nhicd, b->graph_resyncs);
    for (size_t i = 0; i < b->s.size(); i++) {
      b->graph_launches += b->s[i]->launches - l0[i];
      b->s[i]->launches = l0[i];
    }
  } else {
    e = hipEventRecord(b->fork, b->st);
  }
  for (size_t i = 0; !b->fused && rc == XT_OK && e == hipSuccess && i < b->s.size(); i++) {
    xt_icd_state *s = b->s[i];
    e = hipStreamWaitEvent(s->st, b->fork, 0);
    if (e != hipSuccess) break;
    const int64_t l0 = s->launches;
    rc = enqueue_iteration(s, iters[i] > 1 && !s->no_nhicd);
    if (rc == XT_OK) rc = enqueue_stop_stats(s);
    if (rc == XT_OK && b->graph_resyncs) rc = icd_resync(s);
    b->graph_launches += s->launches - l0;
    s->launches = l0;
    if (rc == XT_OK) e = hipEventRecord(b->join[i], s->st);
    if (rc == XT_OK && e == hipSuccess) e = hipStreamWaitEvent(b->st, b->join[i], 0);
  }
  hipGraph_t gr = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->st, &gr);
  if (rc != XT_OK || e != hipSuccess || e2 != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    if (rc != XT_OK) return rc;
    return set_err(XT_E_HIP, "batch graph capture: %s",
                   hipGetErrorString(e != hipSuccess ? e : e2));
  }
  b->graph = gr;
  XT_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return XT_OK;
}

// Drop the batch's graph and fused tables when an option or a slab's line list
// changed since they were made, and plan again.
static int batch_refresh(xt_icd_batch *b) {
  bool stale = b->opt_gen != g_opt_gen || b->lines_gen.size() != b->s.size();
  for (size_t i = 0; !stale && i < b->s.size(); i++) stale = b->lines_gen[i] != b->s[i]->lines_gen;
  if (!stale) return XT_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  if (b->tab) (void)hipFree(b->tab);
  b->tab = nullptr;
  b->groups.clear();
  b->passes.clear();
  XT_TRY(batch_plan(b));
  b->opt_gen = g_opt_gen;
  b->lines_gen.resize(b->s.size());
  for (size_t i = 0; i < b->s.size(); i++) b->lines_gen[i] = b->s[i]->lines_gen;
  return XT_OK;
}

extern "C" int xt_icd_batch_iterate(xt_icd_batch *b, int32_t n_iter) {
  if (!b) return set_err(XT_E_ARG, "null batch");
  for (int it = 0; it < n_iter; it++) {
    XT_TRY(batch_refresh(b));
    std::vector<int> iters(b->s.size());
    bool graphable = g_opt[XT_OPT_GRAPH] != 0;
    for (size_t i = 0; i < b->s.size(); i++) {
      iters[i] = ++b->s[i]->iter;
      graphable = graphable && iters[i] > 1 && !b->s[i]->timing && b->s[i]->update_var != 1;
    }
    // the batch stream starts after whatever the slabs' own streams hold
    for (size_t i = 0; i < b->s.size(); i++) {
      XT_HIP(hipEventRecord(b->join[i], b->s[i]->st));
      XT_HIP(hipStreamWaitEvent(b->st, b->join[i], 0));
    }
    bool fusable = b->fused;
    for (xt_icd_state *s : b->s) fusable = fusable && !s->timing && s->update_var != 1;
    if (!graphable && fusable) {
      // first iteration (no NHICD yet), eager
      XT_TRY(enqueue_batch_fused(b, iters[0] > 1 && !b->s[0]->no_nhicd, false));
      XT_HIP(hipEventRecord(b->done, b->st));
      for (xt_icd_state *s : b->s) XT_HIP(hipStreamWaitEvent(s->st, b->done, 0));
      for (size_t i = 0; i < b->s.size(); i++) XT_TRY(iteration_host(b->s[i], iters[i], false));
      continue;
    }
    if (!graphable) {
      // a slab that needs the host mid-iteration (or the branch form's first
      // iteration): every slab's launches on its own stream, overlapping
      for (size_t i = 0; i < b->s.size(); i++) {
        xt_icd_state *s = b->s[i];
        XT_TRY(enqueue_iteration(s, iters[i] > 1 && !s->no_nhicd));
        if (s->update_var == 1) XT_TRY(variance_update(s));
        XT_TRY(enqueue_stop_stats(s));
      }
      for (size_t i = 0; i < b->s.size(); i++) XT_TRY(iteration_host(b->s[i], iters[i], false));
      continue;
    }
    if (!b->exec) XT_TRY(batch_capture(b, iters));
    XT_HIP(hipGraphLaunch(b->exec, b->st));
    XT_HIP(hipEventRecord(b->done, b->st));
    int64_t per = b->graph_launches / (int64_t)b->s.size();
    for (size_t i = 0; i < b->s.size(); i++) {
      XT_HIP(hipStreamWaitEvent(b->s[i]->st, b->done, 0));
      b->s[i]->launches += per;
    }
    for (size_t i = 0; i < b->s.size(); i++)
      XT_TRY(iteration_host(b->s[i], iters[i], b->graph_resyncs));
  }
  return XT_OK;
}

// ---------------------------------------------------------------------------
// the drop-in boundary
static int write_run_bin(const char *name, const void *data, size_t count) {
  std::string p = run_path(name);
  FILE *fp = fopen(p.c_str(), "wb");
  if (!fp) return set_err(XT_E_IO, "cannot open %s", p.c_str());
  size_t w = fwrite(data, sizeof(Real_t), count, fp);
  fclose(fp);
  if (w != count) return set_err(XT_E_IO, "short write to %s", p.c_str());
  return XT_OK;
}

static int append_run_bin(const char *name, const void *data, size_t count) {
  std::string p = run_path(name);
  FILE *fp = fopen(p.c_str(), "ab");
  if (!fp) return set_err(XT_E_IO, "cannot open %s", p.c_str());
  size_t w = fwrite(data, sizeof(Real_t), count, fp);
  fclose(fp);
  if (w != count) return set_err(XT_E_IO, "short write to %s", p.c_str());
  return XT_OK;
}

static int read_run_bin(const char *name, void *data, size_t count) {
  std::string p = run_path(name);
  FILE *fp = fopen(p.c_str(), "rb");
  if (!fp) return set_err(XT_E_IO, "cannot open %s", p.c_str());
  size_t r = fread(data, sizeof(Real_t), count, fp);
  fclose(fp);
  if (r != count) return set_err(XT_E_IO, "%s holds %zu of %zu values", p.c_str(), r, count);
  return XT_OK;
}

// initObject's upsampling of a coarse [Nzc][Nyc][Nxc] reconstruction (initICD 2:
// y and x doubled; 3: z, y and x), F/XT_ICD_update.c:316-435,484-512.
extern "C" int xt_upsample_object(int32_t init_icd, int32_t Nzc, int32_t Nyc, int32_t Nxc,
                                  const Real_t *coarse, Real_t *fine) {
  if (!coarse || !fine) return set_err(XT_E_ARG, "null argument");
  if (init_icd != 2 && init_icd != 3) return set_err(XT_E_ARG, "initICD %d is not an upsampling", init_icd);
  if (Nzc < (init_icd == 3 ? 2 : 1) || Nyc < 2 || Nxc < 2)
    return set_err(XT_E_ARG, "coarse object %dx%dx%d: every upsampled size must be >= 2", Nzc,
                   Nyc, Nxc);
  int dev = 0;
  XT_TRY(init_device(&dev));
  const int Nx = 2 * Nxc, Ny = 2 * Nyc, Nz = init_icd == 3 ? 2 * Nzc : Nzc;
  const size_t nc = (size_t)Nzc * Nyc * Nxc, nx = (size_t)Nzc * Nyc * Nx,
               ny = (size_t)Nzc * Ny * Nx, nf = (size_t)Nz * Ny * Nx;
  float *buf = nullptr;
  XT_HIP(hipMalloc((void **)&buf, (nc + nx + ny + (init_icd == 3 ? nf : 0)) * sizeof(float)));
  float *dc = buf, *dx = dc + nc, *dy = dx + nx, *dz = dy + ny;
  hipError_t e = real_to_dev(dc, coarse, nc);
  auto grid = [](size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 16384)); };
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_upsample_axis, grid(nx), dim3(256), 0, 0, dc, dx, Nzc * Nyc, Nxc, 1, 0);
    hipLaunchKernelGGL(k_upsample_axis, grid(ny), dim3(256), 0, 0, dx, dy, Nzc, Nyc, Nx, 1);
    if (init_icd == 3)
      hipLaunchKernelGGL(k_upsample_axis, grid(nf), dim3(256), 0, 0, dy, dz, 1, Nzc, Ny * Nx, 0);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = dev_to_real(fine, init_icd == 3 ? dz : dy, nf);
  (void)hipFree(buf);
  if (e != hipSuccess) return set_err(XT_E_HIP, "upsampling: %s", hipGetErrorString(e));
  return XT_OK;
}

// gen_projection_4m_HDF's arithmetic (F/XT_HDFIO.c:160-199) on the GPU.
extern "C" int xt_counts_to_sinogram(int32_t N_p, int32_t T_in, int32_t R_in, int32_t N_t,
                                     int32_t N_r, const uint16_t *counts, int32_t n_white,
                                     const uint16_t *white, int32_t n_dark, const uint16_t *dark,
                                     double bh_quad_coef, Real_t *projection, Real_t *weight,
                                     Real_t *bright) {
  if (!counts || !white || !dark || !projection || !weight || !bright)
    return set_err(XT_E_ARG, "null argument");
  if (N_p <= 0 || N_t <= 0 || N_r <= 0 || T_in < N_t || R_in < N_r || T_in % N_t || R_in % N_r ||
      n_white <= 0 || n_dark <= 0)
    return set_err(XT_E_ARG, "counts %dx%dx%d do not decimate to %d bins x %d rows (or no flat "
                   "frames: %d white, %d dark)", N_p, T_in, R_in, N_r, N_t, n_white, n_dark);
  int dev = 0;
  XT_TRY(init_device(&dev));
  const size_t plane = (size_t)T_in * R_in, ns = (size_t)N_p * N_r * N_t, nb = (size_t)N_r * N_t;
  uint16_t *dc = nullptr, *dw = nullptr, *dd = nullptr;
  float *w2 = nullptr, *d2 = nullptr, *db = nullptr, *dp = nullptr, *dwt = nullptr;
  hipError_t e = hipMalloc((void **)&dc, (size_t)N_p * plane * sizeof(uint16_t));
  if (e == hipSuccess) e = hipMalloc((void **)&dw, (size_t)n_white * plane * sizeof(uint16_t));
  if (e == hipSuccess) e = hipMalloc((void **)&dd, (size_t)n_dark * plane * sizeof(uint16_t));
  if (e == hipSuccess) e = hipMalloc((void **)&w2, 2 * plane * sizeof(float) + nb * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&dp, 2 * ns * sizeof(float));
  if (e == hipSuccess) {
    d2 = w2 + plane;
    db = d2 + plane;
    dwt = dp + ns;
    e = hipMemcpy(dc, counts, (size_t)N_p * plane * sizeof(uint16_t), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess)
    e = hipMemcpy(dw, white, (size_t)n_white * plane * sizeof(uint16_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dd, dark, (size_t)n_dark * plane * sizeof(uint16_t), hipMemcpyHostToDevice);
  auto grid = [](size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 16384)); };
  if (e == hipSuccess) {
    const int rr = R_in / N_r, rt = T_in / N_t;
    hipLaunchKernelGGL(k_flat_mean, grid(plane), dim3(256), 0, 0, dw, n_white, T_in, R_in, w2);
    hipLaunchKernelGGL(k_flat_mean, grid(plane), dim3(256), 0, 0, dd, n_dark, T_in, R_in, d2);
    hipLaunchKernelGGL(k_bright, grid(nb), dim3(256), 0, 0, w2, d2, T_in, N_r, N_t, rr, rt, db);
    hipLaunchKernelGGL(k_counts_sino, grid(ns), dim3(256), 0, 0, dc, d2, db, N_p, T_in, R_in, N_r,
                       N_t, rr, rt, bh_quad_coef, dp, dwt);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = dev_to_real(projection, dp, ns);
  if (e == hipSuccess) e = dev_to_real(weight, dwt, ns);
  if (e == hipSuccess) e = dev_to_real(bright, db, nb);
  for (void *p : {(void *)dc, (void *)dw, (void *)dd, (void *)w2, (void *)dp})
    if (p) (void)hipFree(p);
  if (e != hipSuccess) return set_err(XT_E_HIP, "counts to sinogram: %s", hipGetErrorString(e));
  return XT_OK;
}

// object_n<rank>_time_<i>.bin and proj_offset_n<rank>.bin of the current
// structs, the binary part of write_ObjectProjOff2TiffBinPerIter
// (F/XT_IOMisc.c:224-254)
static int write_object_offset(Sinogram *S, ScannedObject *O, TomoInputs *T) {
  char name[128];
  for (int i = 0; i < O->N_time; i++) {
    snprintf(name, sizeof name, "object_n%d_time_%d.bin", T->node_rank, i);
    XT_TRY(write_run_bin(name, &O->Object[i][1][0][0], (size_t)O->N_z * O->N_y * O->N_x));
  }
  snprintf(name, sizeof name, "proj_offset_n%d.bin", T->node_rank);
  return write_run_bin(name, &S->ProjOffset[0][0], (size_t)S->N_r * S->N_t);
}

extern "C" int ICD_BackProject(Sinogram *S, ScannedObject *O, TomoInputs *T) {
  if (!S || !O || !T) return set_err(XT_E_ARG, "null argument");
  char name[256];
  // initObject, F/XT_ICD_update.c:443-536: zero, the previous object (1), or the
  // previous stage's coarser object upsampled in y, x (2) or z, y, x (3); the
  // magnitude-update map is not read (every line's entry is rewritten in the
  // first iteration before NHICD uses it)
  for (int i = 0; i < O->N_time; i++)
    memset(&O->Object[i][0][0][0], 0, (size_t)(O->N_z + 2) * O->N_y * O->N_x * sizeof(Real_t));
  if (T->initICD == 1) {
    for (int i = 0; i < O->N_time; i++) {
      snprintf(name, sizeof name, "object_n%d_time_%d.bin", T->node_rank, i);
      XT_TRY(read_run_bin(name, &O->Object[i][1][0][0], (size_t)O->N_z * O->N_y * O->N_x));
    }
  } else if (T->initICD == 2 || T->initICD == 3) {
    if (O->N_y % 2 || O->N_x % 2 || (T->initICD == 3 && O->N_z % 2))
      return set_err(XT_E_ARG, "initICD=%d needs even object sizes (%dx%dx%d)", (int)T->initICD,
                     O->N_z, O->N_y, O->N_x);
    const int Nzc = T->initICD == 3 ? O->N_z / 2 : O->N_z;
    std::vector<Real_t> coarse((size_t)Nzc * (O->N_y / 2) * (O->N_x / 2));
    for (int i = 0; i < O->N_time; i++) {
      snprintf(name, sizeof name, "object_n%d_time_%d.bin", T->node_rank, i);
      XT_TRY(read_run_bin(name, coarse.data(), coarse.size()));
      XT_TRY(xt_upsample_object(T->initICD, Nzc, O->N_y / 2, O->N_x / 2, coarse.data(),
                                &O->Object[i][1][0][0]));
    }
    dbg(T, "ICD_BackProject: initial object upsampled from the previous stage (initICD %d)\n",
        (int)T->initICD);
  } else if (T->initICD != 0) {
    return set_err(XT_E_ARG, "initICD=%d is not recognized", (int)T->initICD);
  }
  xt_icd_state *s = nullptr;
  XT_TRY(create_impl(S, O, T, &s));
  dbg(T, "ICD_BackProject: %s, %d masked lines, %d updated at once per frame, z lanes %d, "
      "A table %s\n", xt_version(), s->n_masked, s->conc, s->ZL, s->at_vals ? "on" : "off");
  if (T->initICD == 0 && g_opt[XT_OPT_INIT_FBP] != 0) {
    int rf = icd_fbp(s);
    if (rf != XT_OK) {
      xt_icd_destroy(s);
      return rf;
    }
    dbg(T, "ICD_BackProject: initial object from filtered backprojection\n");
  }
  // the error sinogram now lives on the GPU; Projection is released as the
  // reference does (F/XT_ICD_update.c:999)
  xt_multifree(S->Projection, 3);
  S->Projection = nullptr;
  int rc = XT_OK;
  // the MAP cost before ICD and after every iteration into cost_n<rank>.bin, one
  // Real_t each (F/XT_ICD_update.c:1001-1007,1024-1027; the reference compiles this
  // in without NO_COST_CALCULATE)
  char costfile[64];
  snprintf(costfile, sizeof costfile, "cost_n%d.bin", T->node_rank);
  if (g_opt[XT_OPT_COST] != 0) {
    rc = xt_icd_cost(s, &s->last_cost);
    dbg(T, "ICD_BackProject: cost before ICD %.9g\n", s->last_cost);
    const Real_t c0 = (Real_t)s->last_cost;
    if (rc == XT_OK) rc = write_run_bin(costfile, &c0, 1);
  }
  for (int Iter = 1; rc == XT_OK && Iter <= T->NumIter; Iter++) {
    rc = iterate_one(s);
    // WritePerIter: the object and the ring offsets after every iteration
    // (write_ObjectProjOff2TiffBinPerIter, F/XT_ICD_update.c:1019-1020)
    if (rc == XT_OK && T->WritePerIter == 1) {
      rc = xt_icd_download(s, S, O, T);
      if (rc == XT_OK) rc = write_object_offset(S, O, T);
    }
    if (rc == XT_OK && g_opt[XT_OPT_COST] != 0) {
      const Real_t c = (Real_t)s->last_cost;
      rc = append_run_bin(costfile, &c, 1);
    }
    if (rc == XT_OK && s->stopped && Iter > 1) {
      dbg(T, "ICD_BackProject: Convergence criteria is met\n");
      break;
    }
  }
  if (rc == XT_OK) rc = xt_icd_download(s, S, O, T);
  if (rc == XT_OK) {
    snprintf(name, sizeof name, "variance_estimate_n%d.bin", T->node_rank);
    rc = write_run_bin(name, &T->var_est, 1);
  }
  if (rc == XT_OK) {
    size_t n = (size_t)s->g.N_time * s->g.nb * s->NYX;
    std::vector<Real_t> mag(n);
    if (dev_to_real(mag.data(), s->mag_blk, n) != hipSuccess)
      rc = set_err(XT_E_HIP, "magnitude map download");
    snprintf(name, sizeof name, "mag_update_map_n%d.bin", T->node_rank);
    if (rc == XT_OK) rc = write_run_bin(name, mag.data(), n);
  }
  dbg(T, "ICD_BackProject: Finished (%d iterations)\n", s->iter);
  free_state(s);
  return rc;
}

// ---------------------------------------------------------------------------
// test seams
struct SeamCtx {
  xt_icd_state *s = nullptr;
  ~SeamCtx() { free_state(s); }
};

static int seam_state(SeamCtx &c, Sinogram *S, ScannedObject *O, TomoInputs *T) {
  c.s = new xt_icd_state();
  XT_TRY(make_geom(S, O, T, &c.s->g));
  XT_TRY(init_device(&c.s->dev));
  XT_HIP(hipStreamCreate(&c.s->st));
  c.s->ZL = zl_for(c.s->g.N_z);
  c.s->nt = ((int)g_opt[XT_OPT_LINE_THREADS] == 64 && !c.s->g.pc) ? 64 : 256;
  c.s->nsino = (size_t)S->N_p * S->N_r * S->N_t;
  c.s->NYX = (size_t)O->N_y * O->N_x;
  c.s->nobj = (size_t)O->N_time * c.s->NYX * O->N_z;
  c.s->T = T;
  if (c.s->g.pc) XT_TRY(dev_alloc(&c.s->pc_tmp, c.s->nsino));
  if (T) {
    c.s->var_est = T->var_est;
    c.s->zero_mean = T->EnforceZeroMeanOffset;
  }
  return XT_OK;
}

static int seam_upload_sino(xt_icd_state *s, float **dst, const Real_t *src) {
  XT_TRY(dev_alloc(dst, s->nsino));
  XT_HIP(real_to_dev(*dst, src, s->nsino));
  return XT_OK;
}

static int seam_download_sel(xt_icd_state *s, Sinogram *S) {
  uint8_t *sel = nullptr;
  XT_TRY(dev_alloc(&sel, s->nsino));
  hipLaunchKernelGGL(k_select, dim3(grid_1d(s->nsino)), dim3(256), 0, s->st, s->g, s->e, s->w,
                     s->nsino, sel);
  hipError_t e = hipStreamSynchronize(s->st);
  if (e == hipSuccess) e = hipMemcpy(&S->ProjSelect[0][0][0], sel, s->nsino, hipMemcpyDeviceToHost);
  (void)hipFree(sel);
  if (e != hipSuccess) return set_err(XT_E_HIP, "%s", hipGetErrorString(e));
  return XT_OK;
}

extern "C" int xt_edge_detect(ScannedObject *O, TomoInputs *T, uint8_t ****Mask) {
  if (!O || !T || !Mask) return set_err(XT_E_ARG, "null argument");
  if (O->N_time < 1 || O->N_y < 1 || O->N_x < 1 || O->N_z < 1)
    return set_err(XT_E_ARG, "empty object");
  SeamCtx c;
  c.s = new xt_icd_state();
  xt_icd_state *s = c.s;
  XT_TRY(init_device(&s->dev));
  XT_HIP(hipStreamCreate(&s->st));
  s->g.N_time = O->N_time;
  s->g.N_y = O->N_y;
  s->g.N_x = O->N_x;
  s->g.N_z = O->N_z;
  s->NYX = (size_t)O->N_y * O->N_x;
  s->nobj = (size_t)O->N_time * s->NYX * O->N_z;
  s->edge_par = edge_params(O, T);
  XT_TRY(dev_alloc(&s->x, s->nobj));
  XT_TRY(dev_alloc(&s->vmask, s->nobj));
  XT_TRY(upload_object(s, O, s->x));
  XT_TRY(launch_edge_mask(s, s->x, s->vmask));
  XT_HIP(hipMemcpy(&Mask[0][0][0][0], s->vmask, s->nobj, hipMemcpyDeviceToHost));
  return XT_OK;
}

extern "C" int xt_detector_response(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t **H_r) {
  if (!S || !O || !H_r) return set_err(XT_E_ARG, "null argument");
  SeamCtx c;
  XT_TRY(seam_state(c, S, O, T));
  xt_icd_state *s = c.s;
  XT_TRY(check_angles(S));
  XT_TRY(dev_alloc(&s->viewang, S->N_p));
  XT_TRY(dev_alloc(&s->Hr, (size_t)S->N_p * XT_HR_STRIDE));
  XT_TRY(upload_view_tables(nullptr, nullptr, s->viewang, S));
  XT_TRY(launch_detector_response(s, O->BeamWidth));
  XT_HIP(hipStreamSynchronize(s->st));
  XT_HIP(dev_to_real(&H_r[0][0], s->Hr, (size_t)S->N_p * XT_HR_STRIDE));
  return XT_OK;
}

__global__ void k_amatrix_batch(XtGeom g, const float *__restrict__ Hr, const float *__restrict__ c,
                                const float *__restrict__ s, int n, const int *__restrict__ rows,
                                const int *__restrict__ cols, const int *__restrict__ views,
                                int max_entries, uint8_t *__restrict__ count,
                                int *__restrict__ index, float *__restrict__ values) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = views[i];
  int first;
  float vals[XT_MAXA_WIDE];
  int cnt = xt_amatrix_column(g, Hr + (size_t)v * XT_HR_STRIDE, c[v], s[v], rows[i], cols[i], &first,
                              vals);
  int k = 0;
  for (int b = 0; b < XT_MAXA_WIDE; b++)
    if (vals[b] > 0.f && k < max_entries) {
      index[(size_t)i * max_entries + k] = first + b;
      values[(size_t)i * max_entries + k] = vals[b];
      k++;
    }
  count[i] = (uint8_t)cnt;
}

extern "C" int xt_amatrix_columns(Sinogram *S, ScannedObject *O, Real_t **H_r, int32_t n,
                                  const int32_t *rows, const int32_t *cols, const int32_t *views,
                                  int32_t max_entries, uint8_t *count, int32_t *index,
                                  Real_t *values) {
  if (!S || !O || !H_r || n < 0 || max_entries < 1) return set_err(XT_E_ARG, "bad argument");
  if (n == 0) return XT_OK;
  for (int32_t i = 0; i < n; i++)
    if (views[i] < 0 || views[i] >= S->N_p || rows[i] < 0 || rows[i] >= O->N_y || cols[i] < 0 ||
        cols[i] >= O->N_x)
      return set_err(XT_E_ARG, "triple %d out of range", i);
  SeamCtx c;
  XT_TRY(seam_state(c, S, O, nullptr));
  xt_icd_state *s = c.s;
  int *dr = nullptr, *dc = nullptr, *dv = nullptr, *di = nullptr;
  uint8_t *dn = nullptr;
  float *dval = nullptr;
  XT_TRY(dev_alloc(&s->Hr, (size_t)S->N_p * XT_HR_STRIDE));
  XT_TRY(dev_alloc(&s->cosv, S->N_p));
  XT_TRY(dev_alloc(&s->sinv, S->N_p));
  XT_HIP(real_to_dev(s->Hr, &H_r[0][0], (size_t)S->N_p * XT_HR_STRIDE));
  XT_TRY(upload_view_tables(s->cosv, s->sinv, nullptr, S));
  int rc = XT_OK;
  if ((rc = dev_alloc(&dr, n)) || (rc = dev_alloc(&dc, n)) || (rc = dev_alloc(&dv, n)) ||
      (rc = dev_alloc(&dn, n)) || (rc = dev_alloc(&di, (size_t)n * max_entries)) ||
      (rc = dev_alloc(&dval, (size_t)n * max_entries))) {
  } else {
    hipError_t e = hipMemcpy(dr, rows, n * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dc, cols, n * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dv, views, n * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(di, 0, (size_t)n * max_entries * sizeof(int));
    if (e == hipSuccess) e = hipMemset(dval, 0, (size_t)n * max_entries * sizeof(float));
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_amatrix_batch, dim3((n + 255) / 256), dim3(256), 0, s->st, s->g, s->Hr,
                         s->cosv, s->sinv, n, dr, dc, dv, max_entries, dn, di, dval);
      e = hipStreamSynchronize(s->st);
    }
    if (e == hipSuccess) e = hipMemcpy(count, dn, n, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = hipMemcpy(index, di, (size_t)n * max_entries * sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = dev_to_real(values, dval, (size_t)n * max_entries);
    if (e != hipSuccess) rc = set_err(XT_E_HIP, "amatrix seam: %s", hipGetErrorString(e));
  }
  void *b[] = {dr, dc, dv, dn, di, dval};
  for (void *p : b)
    if (p) (void)hipFree(p);
  return rc;
}

static int seam_geometry_upload(xt_icd_state *s, Sinogram *S, ScannedObject *O, Real_t **H_r) {
  XT_TRY(upload_views(s, S, O));
  XT_TRY(dev_alloc(&s->Hr, (size_t)S->N_p * XT_HR_STRIDE));
  XT_HIP(real_to_dev(s->Hr, &H_r[0][0], (size_t)S->N_p * XT_HR_STRIDE));
  XT_TRY(dev_alloc(&s->x, s->nobj));
  XT_TRY(upload_object(s, O, s->x));
  return XT_OK;
}

extern "C" int xt_forward_project(Sinogram *S, ScannedObject *O, Real_t **H_r, double *AX64) {
  if (!S || !O || !H_r || !AX64) return set_err(XT_E_ARG, "null argument");
  SeamCtx c;
  XT_TRY(seam_state(c, S, O, nullptr));
  xt_icd_state *s = c.s;
  XT_TRY(seam_geometry_upload(s, S, O, H_r));
  XT_TRY(build_atable(s));  // the table the ICD state projects with (XT_OPT_ATABLE)
  double *out = nullptr;
  XT_TRY(dev_alloc(&out, s->nsino));
  int rc = launch_forward(s->st, s->ZL, s->g, s->x, s->Hr, s->cosv, s->sinv, s->view_frame, 0, out,
                          nullptr, nullptr, nullptr, s->pc_tmp, fp_table(s));
  if (rc == XT_OK) {
    hipError_t e = hipStreamSynchronize(s->st);
    if (e == hipSuccess) e = hipMemcpy(AX64, out, s->nsino * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = set_err(XT_E_HIP, "forward projection: %s", hipGetErrorString(e));
  }
  (void)hipFree(out);
  return rc;
}

extern "C" int xt_init_error_sino(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t **H_r,
                                  Real_t ***E) {
  if (!S || !O || !T || !H_r || !E) return set_err(XT_E_ARG, "null argument");
  SeamCtx c;
  XT_TRY(seam_state(c, S, O, T));
  xt_icd_state *s = c.s;
  XT_TRY(seam_geometry_upload(s, S, O, H_r));
  XT_TRY(seam_upload_sino(s, &s->y, &S->Projection[0][0][0]));
  XT_TRY(seam_upload_sino(s, &s->w, &T->Weight[0][0][0]));
  XT_TRY(dev_alloc(&s->e, s->nsino));
  XT_TRY(dev_alloc(&s->d, (size_t)S->N_r * S->N_t));
  XT_HIP(real_to_dev(s->d, &S->ProjOffset[0][0], (size_t)S->N_r * S->N_t));
  XT_TRY(icd_resync(s));
  XT_HIP(hipStreamSynchronize(s->st));
  XT_HIP(dev_to_real(&E[0][0][0], s->e, s->nsino));
  return seam_download_sel(s, S);
}

extern "C" int xt_compute_cost(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E,
                               double *cost) {
  if (!S || !O || !T || !E || !cost) return set_err(XT_E_ARG, "null argument");
  SeamCtx c;
  XT_TRY(seam_state(c, S, O, T));
  xt_icd_state *s = c.s;
  XT_TRY(seam_upload_sino(s, &s->e, &E[0][0][0]));
  XT_TRY(seam_upload_sino(s, &s->w, &T->Weight[0][0][0]));
  XT_TRY(dev_alloc(&s->x, s->nobj));
  XT_TRY(upload_object(s, O, s->x));
  XT_TRY(dev_alloc(&s->red_wg, 8192));
  XT_TRY(dev_alloc(&s->scal, 8));
  return xt_icd_cost(s, cost);
}

extern "C" int xt_offset_update(Sinogram *S, TomoInputs *T, Real_t ***E) {
  if (!S || !T || !E) return set_err(XT_E_ARG, "null argument");
  ScannedObject O;
  memset(&O, 0, sizeof O);
  O.N_x = O.N_y = 1;
  O.N_z = S->N_t;
  O.N_time = 1;
  int32_t pn = 0;
  O.ProjNum = &pn;
  O.delta_xy = S->delta_r;
  S->z_overlap_num = S->z_overlap_num ? S->z_overlap_num : 1;
  SeamCtx c;
  XT_TRY(seam_state(c, S, &O, T));
  xt_icd_state *s = c.s;
  XT_TRY(seam_upload_sino(s, &s->e, &E[0][0][0]));
  XT_TRY(seam_upload_sino(s, &s->w, &T->Weight[0][0][0]));
  const int NRT = S->N_r * S->N_t;
  XT_TRY(dev_alloc(&s->d, NRT));
  XT_HIP(real_to_dev(s->d, &S->ProjOffset[0][0], NRT));
  int nch = (262144 + NRT - 1) / NRT;
  if (nch > S->N_p) nch = S->N_p;
  s->off_chunk = (S->N_p + nch - 1) / nch;
  s->off_nch = (S->N_p + s->off_chunk - 1) / s->off_chunk;
  s->off_nblk = (NRT + 255) / 256;
  XT_TRY(dev_alloc(&s->off_part, (size_t)s->off_nch * 2 * NRT));
  XT_TRY(dev_alloc(&s->off_numden, (size_t)2 * NRT));
  XT_TRY(dev_alloc(&s->off_wg, (size_t)2 * s->off_nblk));
  XT_TRY(dev_alloc(&s->scal, 8));
  XT_TRY(offset_update(s));
  XT_HIP(hipStreamSynchronize(s->st));
  XT_HIP(dev_to_real(&E[0][0][0], s->e, s->nsino));
  XT_HIP(dev_to_real(&S->ProjOffset[0][0], s->d, NRT));
  return seam_download_sel(s, S);
}

// ---------------------------------------------------------------------------
// Ring-offset update with no device allocation.  Its FP64 scratch (per
// view-chunk partial sums, per-pixel sums, per-workgroup sums, gamma) lives in
// one file-scope device array; each GPU holds its own copy at its own address,
// looked up once per device with hipGetSymbolAddress and kept here.  Every
// scratch word a call reads is written earlier in the same call (nothing has to
// start at zero); calls that overlap on one GPU share the array and race.
#define XT_NOALLOC_SCRATCH_DOUBLES (1 << 20)  // 8 MiB per GPU
static __device__ __align__(256) double scratch[XT_NOALLOC_SCRATCH_DOUBLES];
#define XT_MAX_DEVICES 64
static double *g_scratch_ptr[XT_MAX_DEVICES];

static int noalloc_scratch(double **out) {
  int dev = -1;
  XT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= XT_MAX_DEVICES) return set_err(XT_E_ARG, "device %d", dev);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    XT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double *)ptr;
  }
  *out = g_scratch_ptr[dev];
  return XT_OK;
}

extern "C" int xt_offset_update_noalloc(int32_t N_p, int32_t N_r, int32_t N_t, float zinger_T,
                                        float zinger_delta, int32_t zero_mean, float *E,
                                        const float *W, float *ProjOffset) {
  if (!E || !W || !ProjOffset) return set_err(XT_E_ARG, "null argument");
  if (N_p <= 0 || N_r <= 0 || N_t <= 0)
    return set_err(XT_E_ARG, "empty sinogram %d x %d x %d", N_p, N_r, N_t);
  if ((size_t)N_p * N_r * N_t * sizeof(float) >= ((size_t)1 << 31))
    return set_err(XT_E_ARG, "sinogram of %d x %d x %d entries: at most 2^29 - 1", N_p, N_r, N_t);
  // the view chunking of the allocating form (bit-identical results)
  const int NRT = N_r * N_t;
  int nch = (262144 + NRT - 1) / NRT;
  if (nch > N_p) nch = N_p;
  const int chunk = (N_p + nch - 1) / nch;
  nch = (N_p + chunk - 1) / chunk;
  const int nblk = (NRT + 255) / 256;
  const size_t need = (size_t)2 * nch * NRT + (size_t)2 * NRT + (size_t)2 * nblk + 8;
  if (need > XT_NOALLOC_SCRATCH_DOUBLES)
    return set_err(XT_E_ARG, "%d x %d detector pixels need %zu B of offset scratch; the "
                   "allocation-free form holds %zu B (at most about 131072 pixels)", N_r, N_t,
                   need * sizeof(double), (size_t)XT_NOALLOC_SCRATCH_DOUBLES * sizeof(double));
  double *base = nullptr;
  XT_TRY(noalloc_scratch(&base));
  double *part = base, *numden = part + (size_t)2 * nch * NRT, *wg = numden + (size_t)2 * NRT,
         *gamma = wg + (size_t)2 * nblk;
  XtGeom g;
  memset(&g, 0, sizeof g);
  g.N_p = N_p;
  g.N_r = N_r;
  g.N_t = N_t;
  g.T = zinger_T;
  g.delta = zinger_delta;
  const dim3 g2((NRT + 255) / 256, nch);
  hipLaunchKernelGGL(k_offset_partial, g2, dim3(256), 0, 0, g, E, W, chunk, part);
  hipLaunchKernelGGL(k_offset_combine, dim3(nblk), dim3(256), 0, 0, g, nch, part, ProjOffset,
                     numden, wg);
  hipLaunchKernelGGL(k_offset_gamma, dim3(1), dim3(256), 0, 0, nblk, wg, zero_mean ? 1 : 0, gamma);
  hipLaunchKernelGGL(k_offset_apply, g2, dim3(256), 0, 0, g, chunk, numden, gamma, ProjOffset, E);
  XT_HIP(hipGetLastError());
  XT_HIP(hipStreamSynchronize(0));
  return XT_OK;
}

extern "C" int xt_variance_update(Sinogram *S, TomoInputs *T, Real_t ***E) {
  if (!S || !T || !E) return set_err(XT_E_ARG, "null argument");
  ScannedObject O;
  memset(&O, 0, sizeof O);
  O.N_x = O.N_y = 1;
  O.N_z = S->N_t;
  O.N_time = 1;
  int32_t pn = 0;
  O.ProjNum = &pn;
  O.delta_xy = S->delta_r;
  SeamCtx c;
  XT_TRY(seam_state(c, S, &O, T));
  xt_icd_state *s = c.s;
  XT_TRY(seam_upload_sino(s, &s->e, &E[0][0][0]));
  XT_TRY(seam_upload_sino(s, &s->w, &T->Weight[0][0][0]));
  XT_TRY(dev_alloc(&s->red_wg, 8192));
  XT_TRY(dev_alloc(&s->scal, 8));
  XT_TRY(variance_update(s));
  XT_HIP(hipStreamSynchronize(s->st));
  XT_HIP(dev_to_real(&T->Weight[0][0][0], s->w, s->nsino));
  T->var_est = s->var_est;
  return seam_download_sel(s, S);
}

extern "C" int xt_vsc_select(ScannedObject *O, TomoInputs *T, Real_t ****MagUpdateMap) {
  if (!O || !T || !MagUpdateMap) return set_err(XT_E_ARG, "null argument");
  int dev = 0;
  XT_TRY(init_device(&dev));
  const int NY = O->N_y, NX = O->N_x, nb = T->num_z_blocks, M = O->N_time * nb;
  const size_t NYX = (size_t)NY * NX;
  std::vector<int> ns(M);
  for (int i = 0; i < O->N_time; i++)
    for (int b = 0; b < nb; b++)
      ns[i * nb + b] = (int)floor((double)(T->UpdateSelectNum[i][b] / O->NHICD_Iterations));
  const int nch = (int)((NYX + XT_COMPACT_CHUNK - 1) / XT_COMPACT_CHUNK);
  float *mag = nullptr, *filt = nullptr;
  int *nsel = nullptr, *nout = nullptr, *cnt = nullptr, *list = nullptr, *count = nullptr;
  uint8_t *flag = nullptr, *all = nullptr;
  unsigned *vst = nullptr, *vhist = nullptr;
  int rc = XT_OK;
  std::vector<int> hlist((size_t)M * NYX), hcount(M);
  if ((rc = dev_alloc(&mag, M * NYX)) || (rc = dev_alloc(&filt, M * NYX)) ||
      (rc = dev_alloc(&nsel, M)) || (rc = dev_alloc(&nout, M)) || (rc = dev_alloc(&flag, M * NYX)) ||
      (rc = dev_alloc(&vst, (size_t)4 * M)) || (rc = dev_alloc(&vhist, (size_t)256 * M)) ||
      (rc = dev_alloc(&all, NYX)) || (rc = dev_alloc(&cnt, (size_t)M * nch)) ||
      (rc = dev_alloc(&list, M * NYX)) || (rc = dev_alloc(&count, M))) {
  } else {
    hipError_t e = real_to_dev(mag, &MagUpdateMap[0][0][0][0], M * NYX);
    if (e == hipSuccess) e = hipMemcpy(nsel, ns.data(), M * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_vsc_filter, dim3((unsigned)((NYX + 255) / 256), M), dim3(256), 0, 0, NY,
                         NX, mag, filt);
      launch_vsc_select(0, (int)NYX, M, filt, nsel, flag, nout, vst, vhist);
      e = hipMemset(all, 1, NYX);
    }
    if (e == hipSuccess) {
      // the selected lines of every map in raster order, by the kernels that
      // build the NHICD line lists of a reconstruction (mask: every line)
      const dim3 gc((unsigned)nch, (unsigned)M);
      hipLaunchKernelGGL(k_nhicd_chunk_count, gc, dim3(256), 0, 0, (int)NYX, nch, flag, all, cnt);
      hipLaunchKernelGGL(k_nhicd_chunk_write, gc, dim3(256), 0, 0, (int)NYX, nch, flag, all, cnt,
                         list, count);
      e = hipDeviceSynchronize();
    }
    if (e == hipSuccess)
      e = hipMemcpy(hlist.data(), list, M * NYX * sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hcount.data(), count, M * sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = set_err(XT_E_HIP, "vsc seam: %s", hipGetErrorString(e));
  }
  void *b[] = {mag, filt, nsel, nout, flag, vst, vhist, all, cnt, list, count};
  for (void *p : b)
    if (p) (void)hipFree(p);
  if (rc) return rc;
  for (int m = 0; m < M; m++)
    if (hcount[m] < 0 || (size_t)hcount[m] > NYX)
      return set_err(XT_E_HIP, "vsc seam: line count %d of map %d", hcount[m], m);
  for (int i = 0; i < O->N_time; i++)
    for (int bb = 0; bb < nb; bb++) {
      const int m = i * nb + bb;
      const int *L = hlist.data() + (size_t)m * NYX;
      for (int idx = 0; idx < hcount[m]; idx++) {
        T->x_NHICD_select[i][bb][idx] = L[idx] % NX;
        T->y_NHICD_select[i][bb][idx] = L[idx] / NX;
      }
      T->NHICDSelectNum[i][bb] = hcount[m];
    }
  return XT_OK;
}

// The listed lines in list order by one workgroup: every slice, ascending
// (parity < 0, launch mode 2), or the slices of one z-block parity at once with
// the production form of the pass (parity 0/1, mode 3: the sinogram split by
// parity and the parity's slices packed into half the lanes when the blocks
// allow it, as xt_icd_create sets up).
static int update_lines_impl(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E,
                             int32_t frame, int32_t n_lines, const int32_t *x_list,
                             const int32_t *y_list, int32_t Iter, Real_t *MagUpdate,
                             double *total_vox_mag, const uint8_t *Mask, int parity) {
  if (!S || !O || !T || !E || n_lines < 0 || (n_lines && (!x_list || !y_list)))
    return set_err(XT_E_ARG, "bad argument");
  if (frame < 0 || frame >= O->N_time) return set_err(XT_E_ARG, "frame %d", frame);
  if (n_lines > 256) return set_err(XT_E_ARG, "at most 256 lines per call");
  SeamCtx c;
  XT_TRY(seam_state(c, S, O, T));
  xt_icd_state *s = c.s;
  XtGeom &g = s->g;
  if (parity >= 0 && g_opt[XT_OPT_PACK] != 0 && !g.pc && g.zo == 1 && g.nb % 2 == 0 &&
      g.N_z % g.nb == 0 && g.N_t % 2 == 0) {
    g.pm = 1;
    g.bs = g.N_z / g.nb;
    g.H = g.N_t / 2;
  }
  if (icd_lds_bytes(g, s->ZL, false) > 64 * 1024) return set_err(XT_E_ARG, "too many views");
  std::vector<int> cols(n_lines);
  for (int i = 0; i < n_lines; i++) {
    if (x_list[i] < 0 || x_list[i] >= g.N_x || y_list[i] < 0 || y_list[i] >= g.N_y)
      return set_err(XT_E_ARG, "line %d out of range", i);
    cols[i] = y_list[i] * g.N_x + x_list[i];
  }
  // the lines in list order, one workgroup (mode 2)
  s->n_masked = n_lines;
  XT_TRY(dev_alloc(&s->lines, n_lines ? n_lines : 1));
  if (n_lines)
    XT_HIP(hipMemcpy(s->lines, cols.data(), n_lines * sizeof(int), hipMemcpyHostToDevice));
  XT_TRY(upload_views(s, S, O));
  XT_TRY(dev_alloc(&s->Hr, (size_t)S->N_p * XT_HR_STRIDE));
  XT_TRY(upload_view_tables(nullptr, nullptr, s->viewang, S));
  XT_TRY(launch_detector_response(s, O->BeamWidth));
  XT_TRY(build_atable(s));
  XT_TRY(dev_alloc(&s->x, s->nobj));
  XT_TRY(upload_object(s, O, s->x));
  XT_TRY(seam_upload_sino(s, &s->e, &E[0][0][0]));
  XT_TRY(seam_upload_sino(s, &s->w, &T->Weight[0][0][0]));
  if (g.pm) {
    // the split layout of w and e: upload_sino stages each host array through e
    float *w_split = nullptr, *e_split = nullptr;
    XT_TRY(dev_alloc(&w_split, s->nsino));
    int rc = upload_sino(s, w_split, &T->Weight[0][0][0]);
    if (rc == XT_OK) rc = dev_alloc(&e_split, s->nsino);
    if (rc == XT_OK) rc = upload_sino(s, e_split, &E[0][0][0]);
    (void)hipFree(s->w);
    (void)hipFree(s->e);
    s->w = w_split;
    s->e = e_split;
    XT_TRY(rc);
  }
  if (Mask) {
    XT_TRY(dev_alloc(&s->vmask, s->nobj));
    XT_HIP(hipMemcpy(s->vmask, Mask, s->nobj, hipMemcpyHostToDevice));
  }
  XT_TRY(dev_alloc(&s->nhicd_flag, (size_t)g.N_time * s->NYX));
  XT_HIP(hipMemset(s->nhicd_flag, 1, (size_t)g.N_time * s->NYX));
  XT_TRY(dev_alloc(&s->nhicd_blk, (size_t)g.N_time * g.nb * s->NYX));
  XT_HIP(hipMemset(s->nhicd_blk, 1, (size_t)g.N_time * g.nb * s->NYX));
  XT_TRY(dev_alloc(&s->mag_blk, (size_t)g.N_time * g.nb * s->NYX));
  XT_HIP(hipMemset(s->mag_blk, 0, (size_t)g.N_time * g.nb * s->NYX * sizeof(float)));
  XT_TRY(dev_alloc(&s->magsum, (size_t)g.N_time * s->NYX));
  XT_HIP(hipMemset(s->magsum, 0, (size_t)g.N_time * s->NYX * sizeof(float)));
  XT_TRY(dev_alloc(&s->stats, 8));
  XT_HIP(hipMemset(s->stats, 0, 8 * sizeof(double)));
  XT_TRY(dev_alloc(&s->iter_dev, 1));
  XT_HIP(hipMemcpy(s->iter_dev, &Iter, sizeof(int), hipMemcpyHostToDevice));
  XtIcdLaunch L;
  L.mode = parity < 0 ? 2 : 3;
  L.half = parity < 0 ? 0 : parity;
  L.sub_iter = 0;
  L.K = 1;
  L.seed = 0;
  L.count_stats = 1;
  L.zorder = parity < 0 ? 1 : 0;
  XT_TRY(launch_icd(s, L, 1));
  XT_HIP(hipStreamSynchronize(s->st));
  if (g.pm) {
    float *tmp = nullptr;
    XT_TRY(dev_alloc(&tmp, s->nsino));
    hipLaunchKernelGGL(k_sino_layout, dim3(grid_1d(s->nsino)), dim3(256), 0, s->st, s->g, s->e,
                       tmp, s->nsino, 0);
    hipError_t e = hipStreamSynchronize(s->st);
    if (e == hipSuccess)
      e = dev_to_real(&E[0][0][0], tmp, s->nsino);
    (void)hipFree(tmp);
    if (e != hipSuccess) return set_err(XT_E_HIP, "E download: %s", hipGetErrorString(e));
  } else {
    XT_HIP(dev_to_real(&E[0][0][0], s->e, s->nsino));
  }
  XT_TRY(download_object(s, O));
  if (MagUpdate)
    XT_HIP(dev_to_real(MagUpdate, s->magsum + (size_t)frame * s->NYX, s->NYX));
  double st[3];
  XT_HIP(hipMemcpy(st, s->stats, sizeof st, hipMemcpyDeviceToHost));
  if (total_vox_mag) *total_vox_mag = st[0];
  return XT_OK;
}

extern "C" int xt_update_lines(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E,
                               int32_t frame, int32_t n_lines, const int32_t *x_list,
                               const int32_t *y_list, int32_t Iter, Real_t *MagUpdate,
                               double *total_vox_mag, const uint8_t *Mask) {
  return update_lines_impl(S, O, T, E, frame, n_lines, x_list, y_list, Iter, MagUpdate,
                           total_vox_mag, Mask, -1);
}

extern "C" int xt_update_lines_parity(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E,
                                      int32_t n_lines, const int32_t *x_list,
                                      const int32_t *y_list, int32_t Iter, int32_t parity,
                                      Real_t *MagUpdate, double *total_vox_mag) {
  if (parity != 0 && parity != 1) return set_err(XT_E_ARG, "parity %d", parity);
  if (O && O->N_time != 1) return set_err(XT_E_ARG, "one frame (N_time = 1)");
  return update_lines_impl(S, O, T, E, 0, n_lines, x_list, y_list, Iter, MagUpdate,
                           total_vox_mag, nullptr, parity);
}
