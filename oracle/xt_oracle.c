/*
 * xt_oracle.c -- CPU oracle: a plain-C restatement of TIMBIR's ICD path.
 *
 * TEST INFRASTRUCTURE ONLY (see xt_oracle.h for what loads it and how each
 * function is pinned).  Every function cites the reference code it restates;
 * paths are relative to the reference root, F/ = Source_Code_4D_Fast/.
 *
 * Arithmetic deliberately follows the reference's C types: Real_t locals,
 * double-precision libm calls (pow, sqrt, cos, floor) on promoted operands, and
 * the same loop and accumulation order, so the float build reproduces the
 * reference's float build wherever that build can be checked.
 */
#include "xt_oracle.h"

#include <ctype.h>
#include <stddef.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* System-matrix precision.  The reference's float and double builds compute
 * the detector table in Real_t, and the two tables differ by ~3e-3 (float
 * rounding moves the trapezoid's sample boundaries).  ORACLE_FLOAT_AMATRIX
 * (the liboracle_f64a.so build) keeps every other computation in double but
 * builds H_r and the A columns exactly as the float build does, so it isolates
 * arithmetic precision from that model difference. */
#ifdef ORACLE_FLOAT_AMATRIX
typedef float amreal;
#define AM_COS(S, p) ((float)cos((double)(float)(S)->ViewPtr[p]))
#define AM_SIN(S, p) ((float)sin((double)(float)(S)->ViewPtr[p]))
#else
typedef Real_t amreal;
#define AM_COS(S, p) ((S)->cosine[p])
#define AM_SIN(S, p) ((S)->sine[p])
#endif

#define O_PI 3.14159265358979323846
#define O_PI_2 1.57079632679489661923132169163975144
#define O_PI_4 0.785398163397448309615660845819875721

/* HU conversion constants, F/XT_Constants.h:93-100 */
#define O_HFIELD_UNIT_CONV_CONST 0.0001
#define O_AIR_MASS_ATT_COEFF 0.496372335005353
#define O_WATER_MASS_ATT_COEFF 0.521225397034623
#define O_WATER_DENSITY 1.0
#define O_AIR_DENSITY 0.001205
#define O_HOUNSFIELD_WATER_MAP 1000
#define O_HOUNSFIELD_AIR_MAP 0

/* ------------------------------------------------------------------------ */
/* Park-Miller minimal standard generator, F/randlib.c:119-141,168-174.
 * The reference splits the 31-bit state into 15+16-bit halves so that the
 * product fits 32-bit longs; the value it computes is (16807*s) mod (2^31-1),
 * which 64-bit arithmetic gives directly. */
static int64_t g_rng_state = 1;

void oracle_srandom2(uint32_t seed) { g_rng_state = (int64_t)seed; }

double oracle_random2(void) {
  g_rng_state = (g_rng_state * 16807) % 2147483647;
  return ((double)g_rng_state) / 2147483647;
}

void oracle_random2_seq(uint32_t seed, int32_t n, double *out) {
  oracle_srandom2(seed);
  for (int32_t i = 0; i < n; i++) out[i] = oracle_random2();
}

/* ------------------------------------------------------------------------ */
/* Pointer-tree allocation with the layout of F/allocate.c:91-171: one block
 * per pointer level plus one data block, so &X[0]..[0] is contiguous and
 * multifree walks first-pointers (F/allocate.c:177-188). */
static void *o_calloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    exit(1);
  }
  return p;
}

void *oracle_multialloc(size_t s, int32_t d, const int64_t *dims) {
  if (d == 1) return o_calloc((size_t)dims[0], s);
  size_t count = 1;
  char **level_base = NULL;
  void *top = NULL;
  void **prev_level = NULL;
  size_t prev_count = 0;
  for (int32_t lvl = 0; lvl < d - 1; lvl++) {
    count *= (size_t)dims[lvl];
    char **arr = (char **)o_calloc(count, sizeof(char *));
    if (lvl == 0)
      top = arr;
    else {
      for (size_t j = 0; j < prev_count; j++) prev_level[j] = arr + j * (size_t)dims[lvl];
    }
    prev_level = (void **)arr;
    prev_count = count;
    level_base = arr;
  }
  char *data = (char *)o_calloc(count * (size_t)dims[d - 1], s);
  for (size_t j = 0; j < count; j++) level_base[j] = data + j * (size_t)dims[d - 1] * s;
  return top;
}

void oracle_multifree(void *r, int32_t d) {
  void **p = (void **)r;
  for (int32_t i = 0; i < d && p != NULL; i++) {
    void *next = *p;
    free(p);
    p = (void **)next;
  }
}

static void *alloc2(size_t s, int64_t a, int64_t b) {
  int64_t d[2] = {a, b};
  return oracle_multialloc(s, 2, d);
}
static void *alloc3(size_t s, int64_t a, int64_t b, int64_t c) {
  int64_t d[3] = {a, b, c};
  return oracle_multialloc(s, 3, d);
}
static void *alloc4(size_t s, int64_t a, int64_t b, int64_t c, int64_t e) {
  int64_t d[4] = {a, b, c, e};
  return oracle_multialloc(s, 4, d);
}

static void dbg(TomoInputs *T, const char *fmt, ...) {
  if (!T || !T->debug_file_ptr) return;
  va_list ap;
  va_start(ap, fmt);
  vfprintf(T->debug_file_ptr, fmt, ap);
  va_end(ap);
}

/* ------------------------------------------------------------------------ */
/* Detector response H_r[N_p][65]: trapezoidal voxel profile (F/XT_Profile.c:46-103)
 * convolved with a unit-area Hamming beam (F/XT_Profile.c:107-123) and sampled
 * at distances i*OffsetR (F/XT_Profile.c:128-173). */
void oracle_detector_response(Real_t **H_r, Sinogram *S, ScannedObject *O, TomoInputs *T) {
  (void)T;
  int32_t i, j, k, p, ProfileIndex;
  amreal angle, MaxValLineIntegral, temp, dist1, dist2, LeftCorner, LeftNear, RightNear,
      RightCorner, t;
  const amreal delta_xy = (amreal)O->delta_xy, OffsetR = (amreal)S->OffsetR;
  const amreal BeamWidth = (amreal)O->BeamWidth;
  amreal **VoxProfile = (amreal **)alloc2(sizeof(amreal), S->N_p, XT_PROFILE_RESOLUTION);
  for (i = 0; i < (int32_t)S->N_p; i++) {
    angle = (amreal)S->ViewPtr[i];
    while (angle > O_PI_2) angle -= O_PI_2;
    while (angle < 0) angle += O_PI_2;
    if (angle <= O_PI_4)
      MaxValLineIntegral = delta_xy / cos(angle);
    else
      MaxValLineIntegral = delta_xy / cos(O_PI_2 - angle);
    temp = cos(O_PI_4);
    dist1 = temp * cos((O_PI_4 - angle));
    dist2 = temp * fabs((cos((O_PI_4 + angle))));
    LeftCorner = 1 - dist1;
    LeftNear = 1 - dist2;
    RightNear = 1 + dist2;
    RightCorner = 1 + dist1;
    for (j = 0; j < XT_PROFILE_RESOLUTION; j++) {
      t = 2.0 * j / XT_PROFILE_RESOLUTION;
      if (t <= LeftCorner || t >= RightCorner)
        VoxProfile[i][j] = 0;
      else if (t > RightNear)
        VoxProfile[i][j] = MaxValLineIntegral * (RightCorner - t) / (RightCorner - RightNear);
      else if (t >= LeftNear)
        VoxProfile[i][j] = MaxValLineIntegral;
      else
        VoxProfile[i][j] = MaxValLineIntegral * (t - LeftCorner) / (LeftNear - LeftCorner);
    }
  }
  amreal *BeamProfile = (amreal *)o_calloc(XT_BEAM_RESOLUTION, sizeof(amreal));
  amreal sum = 0;
  for (i = 0; i < XT_BEAM_RESOLUTION; i++) {
    BeamProfile[i] = 0.54 - 0.46 * cos((2 * O_PI / XT_BEAM_RESOLUTION) * i);
    sum = sum + BeamProfile[i];
  }
  for (i = 0; i < XT_BEAM_RESOLUTION; i++) BeamProfile[i] /= sum;

  amreal r, rmin, ProfileCenterR, TempConst;
  amreal r0 = -(BeamWidth) / 2;
  amreal StepSize = (BeamWidth) / XT_BEAM_RESOLUTION;
  TempConst = (XT_PROFILE_RESOLUTION) / (2 * delta_xy);
  for (k = 0; k < (int32_t)S->N_p; k++) {
    for (i = 0; i < XT_DETECTOR_RESPONSE_BINS; i++) {
      ProfileCenterR = i * OffsetR;
      rmin = ProfileCenterR - delta_xy;
      sum = 0;
      for (p = 0; p < XT_BEAM_RESOLUTION; p++) {
        r = r0 + p * StepSize;
        ProfileIndex = (int32_t)floor((r - rmin) * TempConst);
        if (ProfileIndex >= 0 && ProfileIndex <= XT_PROFILE_RESOLUTION - 1)
          sum += (VoxProfile[k][ProfileIndex] * BeamProfile[p]);
      }
      H_r[k][i] = sum;
    }
    H_r[k][i] = 0.0;
  }
  free(BeamProfile);
  oracle_multifree(VoxProfile, 2);
}

/* ------------------------------------------------------------------------ */
/* Sparse A column for voxel (row=y, col=x) and one view, F/XT_AMatrix.c:50-110. */
void oracle_amatrix_column(Sinogram *S, ScannedObject *O, Real_t **H_r, AMatrixCol *Ai,
                           int32_t row, int32_t col, int32_t proj_idx) {
  int32_t j;
  amreal x, y, r, rmax, rmin, R_Center, delta_r, w1, w2, f1;
  const amreal x0 = (amreal)O->x0, y0 = (amreal)O->y0, delta_xy = (amreal)O->delta_xy;
  const amreal R0 = (amreal)S->R0, RMax = (amreal)S->RMax, dr = (amreal)S->delta_r;
  const amreal OffsetR = (amreal)S->OffsetR;
  int32_t index_min, index_max, index_delta_r;
  int32_t count = 0;
  x = x0 + ((amreal)col + 0.5) * delta_xy;
  y = y0 + ((amreal)row + 0.5) * delta_xy;
  r = x * AM_COS(S, proj_idx) - y * AM_SIN(S, proj_idx);
  rmin = r - delta_xy;
  rmax = r + delta_xy;
  if (rmax < R0 || rmin > RMax) {
    Ai->count = 0;
    return;
  }
  index_min = (int32_t)(floor(((rmin - R0) / dr)));
  index_max = (int32_t)(floor((rmax - R0) / dr));
  if (index_max >= S->N_r) index_max = S->N_r - 1;
  if (index_min < 0) index_min = 0;
  for (j = index_min; j <= index_max; j++) {
    R_Center = (R0 + (((amreal)j) + 0.5) * (dr));
    delta_r = fabs(r - R_Center);
    index_delta_r = (int32_t)(floor((delta_r / OffsetR)));
    if (index_delta_r >= 0 && index_delta_r < XT_DETECTOR_RESPONSE_BINS) {
      w1 = delta_r - index_delta_r * OffsetR;
      w2 = OffsetR - w1;
      f1 = (w2 * (amreal)H_r[proj_idx][index_delta_r] +
            w1 * (amreal)H_r[proj_idx][index_delta_r + 1]) / OffsetR;
      if (f1 > 0) {
        Ai->values[count] = f1;
        Ai->index[count] = j;
        count++;
      }
    }
  }
  Ai->count = (uint8_t)count;
}

/* Phase-contrast tomography (the reference's -DPHASE_CONTRAST_TOMOGRAPHY build):
 * every voxel's system-matrix footprint is the detector Laplacian of its
 * absorption footprint. */
static int32_t g_phase_contrast = 0;
void oracle_set_phase_contrast(int32_t on) { g_phase_contrast = on != 0; }

#define PC_MAXR 24
#define PC_MAXT 4
/* compute_2DAMatrixLine + compute_LapMatrix_4m_AMatrix, F/XT_AMatrix.c:113-187:
 * the column's entries (times delta_t) as a line over bins index[0].., then
 * L[i][j] = sum over m, n of Lap_Kernel[m+1][n+1] * line[i + m - 1] for the
 * line's rows (col j + n - 1 inside the voxel's z_overlap rows), on the grid
 * grown by one bin and one row on each side, trimmed to the detector.  Lm is
 * [r_num][t_num] for bins r_start.. and rows t_start.. */
static void pc_footprint(Sinogram *S, const AMatrixCol *A, int32_t slice, int32_t *r_start,
                         int32_t *r_num, int32_t *t_start, int32_t *t_num,
                         Real_t Lm[PC_MAXR][PC_MAXT]) {
  int32_t i, j, m, n, row, col;
  Real_t line[PC_MAXR];
  *t_start = slice * S->z_overlap_num;
  *t_num = S->z_overlap_num;
  *r_num = A->count;
  *r_start = A->count ? A->index[0] : 0;
  if (*r_num == 0 || *t_num == 0) return;
  memset(line, 0, sizeof line);
  for (i = 0; i < A->count; i++) line[A->index[i] - *r_start] = A->values[i] * S->delta_t;
  memset(Lm, 0, sizeof(Real_t) * PC_MAXR * PC_MAXT);
  for (i = 0; i < *r_num + 2; i++)
    for (j = 0; j < *t_num + 2; j++)
      for (m = -1; m <= 1; m++)
        for (n = -1; n <= 1; n++) {
          row = i + m - 1;
          col = j + n - 1;
          if (row >= 0 && col >= 0 && row < *r_num && col < *t_num)
            Lm[i][j] += S->Lap_Kernel[m + 1][n + 1] * line[row];
        }
  *r_num += 2;
  *t_num += 2;
  *r_start -= 1;
  *t_start -= 1;
  if (*r_start < 0) {
    for (i = 0; i < *r_num - 1; i++)
      for (j = 0; j < *t_num; j++) Lm[i][j] = Lm[i + 1][j];
    *r_start = 0;
    *r_num -= 1;
  }
  if (*t_start < 0) {
    for (i = 0; i < *r_num; i++)
      for (j = 0; j < *t_num - 1; j++) Lm[i][j] = Lm[i][j + 1];
    *t_start = 0;
    *t_num -= 1;
  }
  if (*r_start + *r_num > S->N_r) *r_num -= 1;
  if (*t_start + *t_num > S->N_t) *t_num -= 1;
}

static int32_t max_entries_per_column(Sinogram *S, ScannedObject *O) {
  /* AvgNumXElements, F/XT_VoxUpdate.c:109 */
  return (uint8_t)ceil(3 * O->delta_xy / S->delta_r);
}

/* ------------------------------------------------------------------------ */
/* Ax by scattering every voxel of every frame, F/XT_ICD_update.c:551-569 with
 * F/XT_ForwardProject.c:8-25.  Frames own disjoint views, so the per-frame
 * parallel loop is race-free (F/XT_Init.c:98-125). */
void oracle_forward_project(Sinogram *S, ScannedObject *O, Real_t **H_r, Real_t ***AX) {
  int32_t nmax = max_entries_per_column(S, O);
  memset(&(AX[0][0][0]), 0, (size_t)S->N_p * S->N_r * S->N_t * sizeof(Real_t));
  int32_t zo = S->z_overlap_num;
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < O->N_time; i++) {
    AMatrixCol A;
    A.values = (Real_t *)o_calloc(nmax, sizeof(Real_t));
    A.index = (int32_t *)o_calloc(nmax, sizeof(int32_t));
    for (int32_t j = 0; j < O->N_y; j++)
      for (int32_t k = 0; k < O->N_x; k++)
        for (int32_t p = 0; p < O->ProjNum[i]; p++) {
          int32_t sino_idx = O->ProjIdxPtr[i][p];
          oracle_amatrix_column(S, O, H_r, &A, j, k, sino_idx);
          for (int32_t slice = 0; slice < O->N_z; slice++) {
            Real_t pixel = O->Object[i][slice + 1][j][k];
            if (g_phase_contrast) {
              /* forward_project_voxel_AMat2D, F/XT_ForwardProject.c:27-44 */
              Real_t Lm[PC_MAXR][PC_MAXT];
              int32_t rs, rn, ts, tn;
              pc_footprint(S, &A, slice, &rs, &rn, &ts, &tn, Lm);
              for (int32_t m = 0; m < rn; m++)
                for (int32_t n = 0; n < tn; n++) AX[sino_idx][rs + m][ts + n] += pixel * Lm[m][n];
              continue;
            }
            for (int32_t m = 0; m < A.count; m++) {
              int32_t idx = A.index[m];
              Real_t val = A.values[m];
              for (int32_t n = 0; n < zo; n++) AX[sino_idx][idx][slice * zo + n] += pixel * val;
            }
          }
        }
    free(A.values);
    free(A.index);
  }
}

/* e = y - Ax - d and ProjSelect, F/XT_ICD_update.c:538-597. */
void oracle_init_error_sino(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t **H_r,
                            Real_t ***E) {
  oracle_forward_project(S, O, H_r, E);
  Real_t avg = 0;
  for (int32_t i = 0; i < S->N_p; i++)
    for (int32_t j = 0; j < S->N_r; j++)
      for (int32_t k = 0; k < S->N_t; k++) {
        E[i][j][k] = S->Projection[i][j][k] - E[i][j][k] - S->ProjOffset[j][k];
        if (fabs(E[i][j][k] * sqrt(T->Weight[i][j][k])) < T->ErrorSinoThresh)
          S->ProjSelect[i][j][k] = true;
        else
          S->ProjSelect[i][j][k] = false;
        avg += E[i][j][k];
      }
  avg = avg / (S->N_r * S->N_t * S->N_p);
  dbg(T, "oracle: average of error sinogram is %f\n", avg);
}

/* ------------------------------------------------------------------------ */
/* qGGMRF potential and its derivatives, F/XT_ICD_update.c:79-125. */
static Real_t qggmrf_spatial_value(Real_t delta, ScannedObject *O, TomoInputs *T) {
  return ((pow(fabs(delta), XT_MRF_Q) / T->Sigma_S_Q) /
          (O->C_S + pow(fabs(delta), XT_MRF_Q - O->MRF_P) / T->Sigma_S_Q_P));
}
static Real_t qggmrf_temporal_value(Real_t delta, ScannedObject *O, TomoInputs *T) {
  return ((pow(fabs(delta), XT_MRF_Q) / T->Sigma_T_Q) /
          (O->C_T + pow(fabs(delta), XT_MRF_Q - O->MRF_P) / T->Sigma_T_Q_P));
}
static Real_t qggmrf_spatial_derivative(Real_t delta, ScannedObject *O, TomoInputs *T) {
  Real_t temp1, temp2, temp3;
  temp1 = pow(fabs(delta), XT_MRF_Q - O->MRF_P) / (T->Sigma_S_Q_P);
  temp2 = pow(fabs(delta), XT_MRF_Q - 1);
  temp3 = O->C_S + temp1;
  if (delta < 0)
    return ((-1 * temp2 / (temp3 * T->Sigma_S_Q)) *
            (XT_MRF_Q - ((XT_MRF_Q - O->MRF_P) * temp1) / (temp3)));
  return ((temp2 / (temp3 * T->Sigma_S_Q)) *
          (XT_MRF_Q - ((XT_MRF_Q - O->MRF_P) * temp1) / (temp3)));
}
static Real_t qggmrf_temporal_derivative(Real_t delta, ScannedObject *O, TomoInputs *T) {
  Real_t temp1, temp2, temp3;
  temp1 = pow(fabs(delta), XT_MRF_Q - O->MRF_P) / (T->Sigma_T_Q_P);
  temp2 = pow(fabs(delta), XT_MRF_Q - 1);
  temp3 = O->C_T + temp1;
  if (delta < 0)
    return ((-1 * temp2 / (temp3 * T->Sigma_T_Q)) *
            (XT_MRF_Q - ((XT_MRF_Q - O->MRF_P) * temp1) / (temp3)));
  return ((temp2 / (temp3 * T->Sigma_T_Q)) *
          (XT_MRF_Q - ((XT_MRF_Q - O->MRF_P) * temp1) / (temp3)));
}
static Real_t qggmrf_spatial_second_derivative(ScannedObject *O, TomoInputs *T) {
  return XT_MRF_Q / (T->Sigma_S_Q * O->C_S);
}
static Real_t qggmrf_temporal_second_derivative(ScannedObject *O, TomoInputs *T) {
  return XT_MRF_Q / (T->Sigma_T_Q * O->C_T);
}

/* One-dimensional surrogate minimiser with over-relaxation,
 * F/XT_ICD_update.c:130-179.  nhood27/flag27 are [3][3][3] row-major. */
Real_t oracle_functional_substitution(Real_t V, Real_t THETA1, Real_t THETA2, ScannedObject *O,
                                      TomoInputs *T, const Real_t *nhood27, const uint8_t *flag27,
                                      const Real_t *time_nhood2, const uint8_t *time_flag2,
                                      int32_t positivity) {
  Real_t u, temp1 = 0, temp2 = 0, temp_const, RefValue = 0, Delta0, QGGMRF_Params;
  int32_t i, j, k;
  RefValue = V;
  for (i = 0; i < XT_NHOOD_Y_MAXDIM; i++)
    for (j = 0; j < XT_NHOOD_X_MAXDIM; j++)
      for (k = 0; k < XT_NHOOD_Z_MAXDIM; k++) {
        int32_t l = (i * 3 + j) * 3 + k;
        if (flag27[l] && (i != 1 || j != 1 || k != 1)) {
          Delta0 = (RefValue - nhood27[l]);
          if (Delta0 != 0)
            QGGMRF_Params = qggmrf_spatial_derivative(Delta0, O, T) / (Delta0);
          else
            QGGMRF_Params = qggmrf_spatial_second_derivative(O, T);
          temp_const = T->Spatial_Filter[i][j][k] * QGGMRF_Params;
          temp1 += temp_const * nhood27[l];
          temp2 += temp_const;
        }
      }
  for (i = 0; i < XT_NHOOD_TIME_MAXDIM - 1; i++) {
    if (time_flag2[i]) {
      Delta0 = (RefValue - time_nhood2[i]);
      if (Delta0 != 0)
        QGGMRF_Params = qggmrf_temporal_derivative(Delta0, O, T) / (Delta0);
      else
        QGGMRF_Params = qggmrf_temporal_second_derivative(O, T);
      temp_const = T->Time_Filter[0] * QGGMRF_Params;
      temp1 += temp_const * time_nhood2[i];
      temp2 += temp_const;
    }
  }
  u = (temp1 + (THETA2 * V) - THETA1) / (temp2 + THETA2);
  RefValue = RefValue + T->alpha * (u - RefValue);
  if (positivity && RefValue <= 0) RefValue = 0;
  return RefValue;
}

/* ------------------------------------------------------------------------ */
/* Voxel-line update loop, F/XT_VoxUpdate.c:85-254 (live path updateVoxels_Aditya),
 * with the per-voxel step of F/XT_VoxUpdate.c:11-83. */
static void compute_voxel_update(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E,
                                 AMatrixCol *A, const Real_t *nhood27, const uint8_t *flag27,
                                 const Real_t *tn, const uint8_t *tf, int32_t i_new, int32_t slice,
                                 int32_t j_new, int32_t k_new, int32_t positivity) {
  int32_t p, q, r, sino_view, z_overlap_num, i_r, i_t;
  Real_t V, THETA1, THETA2, THETASelTemp, UpdatedVoxelValue, ProjectionEntry;
  V = O->Object[i_new][slice + 1][j_new][k_new];
  z_overlap_num = S->z_overlap_num;
  THETA1 = 0.0;
  THETA2 = 0.0;
  if (g_phase_contrast) {
    /* compute_voxel_update_AMat2D, F/XT_VoxUpdate.c:590-690 */
    Real_t Lm[PC_MAXR][PC_MAXT];
    int32_t rs, rnum, ts, tnum;
    for (p = 0; p < O->ProjNum[i_new]; p++) {
      sino_view = O->ProjIdxPtr[i_new][p];
      pc_footprint(S, &A[p], slice, &rs, &rnum, &ts, &tnum, Lm);
      for (q = 0; q < rnum; q++) {
        i_r = rs + q;
        for (r = 0; r < tnum; r++) {
          ProjectionEntry = Lm[q][r];
          i_t = ts + r;
          if (S->ProjSelect[sino_view][i_r][i_t] == true) {
            THETA2 += (ProjectionEntry * ProjectionEntry * T->Weight[sino_view][i_r][i_t]);
            THETA1 += -(E[sino_view][i_r][i_t] * ProjectionEntry * T->Weight[sino_view][i_r][i_t]);
          } else {
            THETASelTemp = T->ErrorSinoThresh * T->ErrorSinoDelta *
                           sqrt(T->Weight[sino_view][i_r][i_t]) / fabs(E[sino_view][i_r][i_t]);
            THETA2 += (ProjectionEntry * ProjectionEntry * THETASelTemp);
            THETA1 += -(E[sino_view][i_r][i_t] * ProjectionEntry * THETASelTemp);
          }
        }
      }
    }
    UpdatedVoxelValue =
        oracle_functional_substitution(V, THETA1, THETA2, O, T, nhood27, flag27, tn, tf, positivity);
    O->Object[i_new][slice + 1][j_new][k_new] = UpdatedVoxelValue;
    for (p = 0; p < O->ProjNum[i_new]; p++) {
      sino_view = O->ProjIdxPtr[i_new][p];
      pc_footprint(S, &A[p], slice, &rs, &rnum, &ts, &tnum, Lm);
      for (q = 0; q < rnum; q++) {
        i_r = rs + q;
        for (r = 0; r < tnum; r++) {
          i_t = ts + r;
          E[sino_view][i_r][i_t] -= (Lm[q][r] * (O->Object[i_new][slice + 1][j_new][k_new] - V));
          if (fabs(E[sino_view][i_r][i_t] * sqrt(T->Weight[sino_view][i_r][i_t])) <
              T->ErrorSinoThresh)
            S->ProjSelect[sino_view][i_r][i_t] = true;
          else
            S->ProjSelect[sino_view][i_r][i_t] = false;
        }
      }
    }
    return;
  }
  for (p = 0; p < O->ProjNum[i_new]; p++) {
    sino_view = O->ProjIdxPtr[i_new][p];
    for (q = 0; q < A[p].count; q++) {
      i_r = (A[p].index[q]);
      ProjectionEntry = (A[p].values[q] * S->delta_t);
      for (r = 0; r < z_overlap_num; r++) {
        i_t = slice * z_overlap_num + r;
        if (S->ProjSelect[sino_view][i_r][i_t] == true) {
          THETA2 += (ProjectionEntry * ProjectionEntry * T->Weight[sino_view][i_r][i_t]);
          THETA1 += -(E[sino_view][i_r][i_t] * ProjectionEntry * T->Weight[sino_view][i_r][i_t]);
        } else {
          THETASelTemp = T->ErrorSinoThresh * T->ErrorSinoDelta *
                         sqrt(T->Weight[sino_view][i_r][i_t]) / fabs(E[sino_view][i_r][i_t]);
          THETA2 += (ProjectionEntry * ProjectionEntry * THETASelTemp);
          THETA1 += -(E[sino_view][i_r][i_t] * ProjectionEntry * THETASelTemp);
        }
      }
    }
  }
  UpdatedVoxelValue =
      oracle_functional_substitution(V, THETA1, THETA2, O, T, nhood27, flag27, tn, tf, positivity);
  O->Object[i_new][slice + 1][j_new][k_new] = UpdatedVoxelValue;
  for (p = 0; p < O->ProjNum[i_new]; p++) {
    sino_view = O->ProjIdxPtr[i_new][p];
    for (q = 0; q < A[p].count; q++) {
      i_r = (A[p].index[q]);
      ProjectionEntry = (A[p].values[q] * S->delta_t);
      for (r = 0; r < z_overlap_num; r++) {
        i_t = slice * z_overlap_num + r;
        E[sino_view][i_r][i_t] -=
            (ProjectionEntry * (O->Object[i_new][slice + 1][j_new][k_new] - V));
        if (fabs(E[sino_view][i_r][i_t] * sqrt(T->Weight[sino_view][i_r][i_t])) <
            T->ErrorSinoThresh)
          S->ProjSelect[sino_view][i_r][i_t] = true;
        else
          S->ProjSelect[sino_view][i_r][i_t] = false;
      }
    }
  }
}

static int32_t max_proj_num(ScannedObject *O) {
  int32_t m = O->ProjNum[0];
  for (int32_t i = 1; i < O->N_time; i++)
    if (O->ProjNum[i] > m) m = O->ProjNum[i];
  return m;
}

static Real_t update_lines_masked(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E,
                                  Real_t **H_r, int32_t i_new, int32_t slice_begin,
                                  int32_t slice_end, int32_t n_lines, const int32_t *x_rand_select,
                                  const int32_t *y_rand_select, int32_t Iter, int64_t *zero_count,
                                  Real_t **MagUpdateMap, uint8_t ***Mask, int32_t positivity) {
  int32_t p, q, r, slice, j_new, k_new, idxr, idxq, idxp, index_xy;
  Real_t V;
  bool ZSFlag;
  Real_t total_vox_mag = 0.0;
  /* single rank: halo slices 0 and N_z+1 are outside (F/XT_VoxUpdate.c:94-99) */
  int32_t z_min = 1, z_max = O->N_z;
  Real_t nhood[27], tn[2];
  uint8_t flag[27], tf[2];
  int32_t maxview = max_proj_num(O);
  int32_t nmax = max_entries_per_column(S, O);
  AMatrixCol *A = (AMatrixCol *)o_calloc(maxview, sizeof(AMatrixCol));
  for (p = 0; p < maxview; p++) {
    A[p].values = (Real_t *)o_calloc(nmax, sizeof(Real_t));
    A[p].index = (int32_t *)o_calloc(nmax, sizeof(int32_t));
  }
  for (index_xy = 0; index_xy < n_lines; index_xy++) {
    k_new = x_rand_select[index_xy];
    j_new = y_rand_select[index_xy];
    MagUpdateMap[j_new][k_new] = 0;
    for (p = 0; p < O->ProjNum[i_new]; p++)
      oracle_amatrix_column(S, O, H_r, &A[p], j_new, k_new, O->ProjIdxPtr[i_new][p]);
    for (slice = slice_begin; slice <= slice_end; slice++) {
      if (Mask && Mask[j_new][k_new][slice] != 1) continue;
      if (i_new - 1 >= 0) {
        tn[0] = O->Object[i_new - 1][slice + 1][j_new][k_new];
        tf[0] = 1;
      } else {
        tn[0] = 0.0;
        tf[0] = 0;
      }
      if (i_new + 1 < O->N_time) {
        tn[1] = O->Object[i_new + 1][slice + 1][j_new][k_new];
        tf[1] = 1;
      } else {
        tn[1] = 0.0;
        tf[1] = 0;
      }
      for (p = 0; p < 3; p++) {
        idxp = slice + p;
        for (q = 0; q < 3; q++) {
          idxq = j_new + q - 1;
          for (r = 0; r < 3; r++) {
            idxr = k_new + r - 1;
            int32_t l = (p * 3 + q) * 3 + r;
            if (idxp >= z_min && idxp <= z_max && idxq >= 0 && idxq < O->N_y && idxr >= 0 &&
                idxr < O->N_x) {
              nhood[l] = O->Object[i_new][idxp][idxq][idxr];
              flag[l] = 1;
            } else {
              nhood[l] = 0.0;
              flag[l] = 0;
            }
          }
        }
      }
      nhood[13] = 0.0;
      V = O->Object[i_new][slice + 1][j_new][k_new];
      /* zero skipping, F/XT_VoxUpdate.c:208-232 */
      ZSFlag = true;
      if (V == 0.0 && Iter > 1) {
        if (tn[0] > 0.0 || tn[1] > 0.0) ZSFlag = false;
        for (p = 0; p < 27; p++)
          if (nhood[p] > 0.0) {
            ZSFlag = false;
            break;
          }
      } else {
        ZSFlag = false;
      }
      if (ZSFlag == false) {
        compute_voxel_update(S, O, T, E, A, nhood, flag, tn, tf, i_new, slice, j_new, k_new,
                             positivity);
        MagUpdateMap[j_new][k_new] += fabs(O->Object[i_new][slice + 1][j_new][k_new] - V);
        total_vox_mag += fabs(O->Object[i_new][slice + 1][j_new][k_new]);
      } else if (zero_count) {
        (*zero_count)++;
      }
    }
  }
  for (p = 0; p < maxview; p++) {
    free(A[p].values);
    free(A[p].index);
  }
  free(A);
  return total_vox_mag;
}

Real_t oracle_update_lines(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E, Real_t **H_r,
                           int32_t frame, int32_t z0, int32_t z1, int32_t n_lines,
                           const int32_t *x_list, const int32_t *y_list, int32_t Iter,
                           int64_t *zero_count, Real_t **MagUpdateMap, int32_t positivity) {
  return update_lines_masked(S, O, T, E, H_r, frame, z0, z1, n_lines, x_list, y_list, Iter,
                             zero_count, MagUpdateMap, NULL, positivity);
}

/* ------------------------------------------------------------------------ */
/* Random line order per (frame, z-block), F/XT_ICD_update.c:599-631. */
void oracle_random_select(ScannedObject *O, TomoInputs *T, uint8_t ****Mask) {
  int32_t i, j, num, n, Index, col, row, ArraySize, block;
  ArraySize = O->N_y * O->N_x;
  int32_t *Counter = (int32_t *)o_calloc(ArraySize, sizeof(int32_t));
  for (i = 0; i < O->N_time; i++)
    for (block = 0; block < T->num_z_blocks; block++) {
      ArraySize = O->N_y * O->N_x;
      for (Index = 0; Index < ArraySize; Index++) Counter[Index] = Index;
      T->UpdateSelectNum[i][block] = 0;
      for (j = 0; j < O->N_x * O->N_y; j++) {
        Index = floor(oracle_random2() * ArraySize);
        Index = (Index == ArraySize) ? ArraySize - 1 : Index;
        col = Counter[Index] % O->N_x;
        row = Counter[Index] / O->N_x;
        for (n = block * (O->N_z / T->num_z_blocks); n < (block + 1) * (O->N_z / T->num_z_blocks);
             n++)
          if (Mask[i][row][col][n] == 1) {
            num = T->UpdateSelectNum[i][block];
            T->x_rand_select[i][block][num] = col;
            T->y_rand_select[i][block][num] = row;
            (T->UpdateSelectNum[i][block])++;
            break;
          }
        Counter[Index] = Counter[ArraySize - 1];
        ArraySize--;
      }
    }
  free(Counter);
}

/* ------------------------------------------------------------------------ */
/* NHICD voxel-line selection, F/XT_NHICD.c:63-216. */
static uint32_t o_partition(Real_t *A, uint32_t p, uint32_t r) {
  Real_t x = A[r], temp;
  uint32_t i = p - 1, j;
  for (j = p; j < r; j++)
    if (A[j] <= x) {
      i++;
      temp = A[i];
      A[i] = A[j];
      A[j] = temp;
    }
  temp = A[i + 1];
  A[i + 1] = A[r];
  A[r] = temp;
  return i + 1;
}
static uint32_t o_randomized_partition(Real_t *A, uint32_t p, uint32_t r) {
  uint32_t rnum = floor(oracle_random2() * (r - p + 1));
  rnum = (rnum == r - p + 1) ? r - p : rnum;
  uint32_t j = p + rnum;
  Real_t temp = A[r];
  A[r] = A[j];
  A[j] = temp;
  return o_partition(A, p, r);
}
static Real_t o_randomized_select(Real_t *A, uint32_t p, uint32_t r, uint32_t i) {
  for (;;) {
    if (p == r) return A[p];
    uint32_t q = o_randomized_partition(A, p, r);
    uint32_t k = q - p + 1;
    if (i == k) return A[q];
    if (i < k)
      r = q - 1;
    else {
      p = q + 1;
      i = i - k;
    }
  }
}

static void o_compute_vsc(ScannedObject *O, Real_t **mag, Real_t **filt, Real_t **filt_copy) {
  static const Real_t HammingWindow[5][5] = {{0.0064, 0.0432, 0.0800, 0.0432, 0.0064},
                                             {0.0432, 0.2916, 0.5400, 0.2916, 0.0432},
                                             {0.0800, 0.5400, 1.0000, 0.5400, 0.0800},
                                             {0.0432, 0.2916, 0.5400, 0.2916, 0.0432},
                                             {0.0064, 0.0432, 0.0800, 0.0432, 0.0064}};
  for (int32_t i = 0; i < O->N_y; i++)
    for (int32_t j = 0; j < O->N_x; j++) {
      Real_t filter_op = 0;
      for (int32_t p = -2; p <= 2; p++)
        for (int32_t q = -2; q <= 2; q++)
          if (i + p >= 0 && i + p < O->N_y && j + q >= 0 && j + q < O->N_x)
            filter_op += HammingWindow[p + 2][q + 2] * mag[i + p][j + q];
      filt[i][j] = filter_op;
      filt_copy[i][j] = filter_op;
    }
}

void oracle_vsc_select(ScannedObject *O, TomoInputs *T, Real_t ****MagUpdateMap) {
  Real_t **filt = (Real_t **)alloc2(sizeof(Real_t), O->N_y, O->N_x);
  Real_t **filt_copy = (Real_t **)alloc2(sizeof(Real_t), O->N_y, O->N_x);
  for (int32_t i = 0; i < O->N_time; i++)
    for (int32_t block = 0; block < T->num_z_blocks; block++) {
      int32_t idx = 0;
      T->NHICDSelectNum[i][block] = floor(T->UpdateSelectNum[i][block] / O->NHICD_Iterations);
      o_compute_vsc(O, MagUpdateMap[i][block], filt, filt_copy);
      Real_t thresh = o_randomized_select(&(filt[0][0]), 0, O->N_x * O->N_y - 1,
                                          O->N_x * O->N_y - T->NHICDSelectNum[i][block]);
      for (int32_t j = 0; j < O->N_y; j++)
        for (int32_t k = 0; k < O->N_x; k++)
          if (filt_copy[j][k] > thresh) {
            T->x_NHICD_select[i][block][idx] = k;
            T->y_NHICD_select[i][block][idx] = j;
            idx++;
          }
      T->NHICDSelectNum[i][block] = idx;
    }
  oracle_multifree(filt, 2);
  oracle_multifree(filt_copy, 2);
}

/* ------------------------------------------------------------------------ */
/* Variance re-estimate, F/XT_ICD_update.c:633-664 (single rank). */
void oracle_variance_update(Sinogram *S, TomoInputs *T, Real_t ***E) {
  Real_t temp_acc = 0, temp = 0;
  for (int32_t k = 0; k < S->N_p; k++)
    for (int32_t i = 0; i < S->N_r; i++)
      for (int32_t j = 0; j < S->N_t; j++) {
        T->Weight[k][i][j] = T->Weight[k][i][j] * T->var_est;
        if (S->ProjSelect[k][i][j] == true)
          temp = E[k][i][j] * E[k][i][j] * T->Weight[k][i][j];
        else
          temp = fabs(E[k][i][j]) * T->ErrorSinoDelta * T->ErrorSinoThresh *
                 sqrt(T->Weight[k][i][j] * T->var_est);
        temp_acc += temp;
      }
  temp = temp_acc;
  T->var_est = temp / ((Real_t)T->node_num * (Real_t)S->N_p * (Real_t)S->N_r * (Real_t)S->N_t);
  for (int32_t k = 0; k < S->N_p; k++)
    for (int32_t i = 0; i < S->N_r; i++)
      for (int32_t j = 0; j < S->N_t; j++) {
        T->Weight[k][i][j] /= T->var_est;
        if (fabs(E[k][i][j] * sqrt(T->Weight[k][i][j])) < T->ErrorSinoThresh)
          S->ProjSelect[k][i][j] = true;
        else
          S->ProjSelect[k][i][j] = false;
      }
}

/* Ring-offset update with optional zero-mean constraint, F/XT_ICD_update.c:666-759. */
void oracle_offset_update(Sinogram *S, TomoInputs *T, Real_t ***E) {
  Real_t **numerator = (Real_t **)alloc2(sizeof(Real_t), S->N_r, S->N_t);
  Real_t **denominator = (Real_t **)alloc2(sizeof(Real_t), S->N_r, S->N_t);
  Real_t num_sum = 0, den_sum = 0, gamma = 0;
  for (int32_t i = 0; i < S->N_r; i++)
    for (int32_t j = 0; j < S->N_t; j++) {
      numerator[i][j] = 0;
      denominator[i][j] = 0;
      for (int32_t k = 0; k < S->N_p; k++) {
        Real_t temp = T->ErrorSinoThresh * T->ErrorSinoDelta * sqrt(T->Weight[k][i][j]);
        if (S->ProjSelect[k][i][j] == true) {
          numerator[i][j] += E[k][i][j] * T->Weight[k][i][j];
          denominator[i][j] += T->Weight[k][i][j];
        } else {
          Real_t sign = (E[k][i][j] > 0) - (E[k][i][j] < 0);
          numerator[i][j] += temp * sign;
          denominator[i][j] += temp / fabs(E[k][i][j]);
        }
      }
      num_sum += S->ProjOffset[i][j] + (numerator[i][j] / denominator[i][j]);
      den_sum += 1.0 / denominator[i][j];
    }
  if (T->EnforceZeroMeanOffset == 1)
    gamma = num_sum / den_sum;
  else
    gamma = 0;
  for (int32_t i = 0; i < S->N_r; i++)
    for (int32_t j = 0; j < S->N_t; j++) {
      S->ProjOffset[i][j] = S->ProjOffset[i][j] + (numerator[i][j] - gamma) / denominator[i][j];
      for (int32_t k = 0; k < S->N_p; k++) {
        E[k][i][j] -= (numerator[i][j] - gamma) / denominator[i][j];
        if (fabs(E[k][i][j] * sqrt(T->Weight[k][i][j])) < T->ErrorSinoThresh)
          S->ProjSelect[k][i][j] = true;
        else
          S->ProjSelect[k][i][j] = false;
      }
    }
  oracle_multifree(numerator, 2);
  oracle_multifree(denominator, 2);
}

/* MAP cost, F/XT_ICD_update.c:181-314 (single rank). */
Real_t oracle_compute_cost(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E) {
  Real_t cost = 0, temp = 0, forward = 0, prior = 0, delta;
  int32_t i, j, k, p, N_z;
  for (i = 0; i < S->N_p; i++)
    for (j = 0; j < S->N_r; j++)
      for (k = 0; k < S->N_t; k++) {
        temp = E[i][j][k] * sqrt(T->Weight[i][j][k]);
        if (S->ProjSelect[i][j][k] == true)
          temp = temp * temp;
        else
          temp = 2.0 * T->ErrorSinoDelta * T->ErrorSinoThresh * fabs(temp) +
                 T->ErrorSinoThresh * T->ErrorSinoThresh * (1.0 - 2.0 * T->ErrorSinoDelta);
        cost += temp;
      }
  cost /= 2.0;
  temp = 0;
  N_z = O->N_z + 1; /* last (and only) rank, F/XT_ICD_update.c:203-205 */
  Real_t ****X = O->Object;
  Real_t (*F)[3][3] = T->Spatial_Filter;
  for (i = 0; i < O->N_time; i++)
    for (p = 1; p < O->N_z + 1; p++)
      for (j = 0; j < O->N_y; j++)
        for (k = 0; k < O->N_x; k++) {
          bool j_minus = (j - 1 >= 0), k_minus = (k - 1 >= 0);
          bool p_plus = (p + 1 < N_z), i_plus = (i + 1 < O->N_time);
          bool j_plus = (j + 1 < O->N_y), k_plus = (k + 1 < O->N_x);
          Real_t c = X[i][p][j][k];
          if (k_plus) {
            delta = c - X[i][p][j][k + 1];
            temp += F[1][1][2] * qggmrf_spatial_value(delta, O, T);
          }
          if (j_plus) {
            if (k_minus) {
              delta = c - X[i][p][j + 1][k - 1];
              temp += F[1][2][0] * qggmrf_spatial_value(delta, O, T);
            }
            delta = c - X[i][p][j + 1][k];
            temp += F[1][2][1] * qggmrf_spatial_value(delta, O, T);
            if (k_plus) {
              delta = c - X[i][p][j + 1][k + 1];
              temp += F[1][2][2] * qggmrf_spatial_value(delta, O, T);
            }
          }
          if (p_plus) {
            if (j_minus) {
              delta = c - X[i][p + 1][j - 1][k];
              temp += F[2][0][1] * qggmrf_spatial_value(delta, O, T);
            }
            delta = c - X[i][p + 1][j][k];
            temp += F[2][1][1] * qggmrf_spatial_value(delta, O, T);
            if (j_plus) {
              delta = c - X[i][p + 1][j + 1][k];
              temp += F[2][2][1] * qggmrf_spatial_value(delta, O, T);
            }
            if (j_minus) {
              if (k_minus) {
                delta = c - X[i][p + 1][j - 1][k - 1];
                temp += F[2][0][0] * qggmrf_spatial_value(delta, O, T);
              }
              if (k_plus) {
                delta = c - X[i][p + 1][j - 1][k + 1];
                temp += F[2][0][2] * qggmrf_spatial_value(delta, O, T);
              }
            }
            if (k_minus) {
              delta = c - X[i][p + 1][j][k - 1];
              temp += F[2][1][0] * qggmrf_spatial_value(delta, O, T);
            }
            if (j_plus) {
              if (k_minus) {
                delta = c - X[i][p + 1][j + 1][k - 1];
                temp += F[2][2][0] * qggmrf_spatial_value(delta, O, T);
              }
              if (k_plus) {
                delta = c - X[i][p + 1][j + 1][k + 1];
                temp += F[2][2][2] * qggmrf_spatial_value(delta, O, T);
              }
            }
            if (k_plus) {
              delta = c - X[i][p + 1][j][k + 1];
              temp += F[2][1][2] * qggmrf_spatial_value(delta, O, T);
            }
          }
          if (i_plus) {
            delta = c - X[i + 1][p][j][k];
            temp += T->Time_Filter[0] * qggmrf_temporal_value(delta, O, T);
          }
        }
  forward = cost;
  prior = temp;
  T->ErrorSino_Cost = forward;
  forward += (Real_t)T->node_num * (Real_t)S->N_p * (Real_t)S->N_r * (Real_t)S->N_t *
             log(T->var_est) / 2;
  T->Forward_Cost = forward;
  T->Prior_Cost = prior;
  return forward + prior;
}

/* ------------------------------------------------------------------------ */
/* One ICD iteration with the even/odd z-block schedule and NHICD,
 * F/XT_ICD_update.c:767-926. Returns 1 when the stop threshold is met. */
static int32_t icd_iteration(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t **H_r,
                             Real_t ***E, int32_t Iter, Real_t ****MagUpdateMap,
                             uint8_t ****Mask, int32_t positivity) {
  Real_t AverageUpdate = 0, total_vox_mag = 0.0;
  Real_t total_pix = 0;
  int32_t nb = T->num_z_blocks;
  int64_t *zero_count = (int64_t *)o_calloc((size_t)O->N_time * nb, sizeof(int64_t));
  int32_t K = O->NHICD_Iterations;
  oracle_random_select(O, T, Mask);
  dbg(T, "Number of NHICD iterations is %d\n", K);

  for (int32_t j = 0; j < K; j++) {
    for (int32_t phase = 0; phase < 4; phase++) {
      int32_t nhicd = phase >= 2;
      if (nhicd && !(Iter > 1 && T->no_NHICD == 0)) break;
      int32_t odd = phase & 1;
      Real_t mag_acc = 0.0;
#pragma omp parallel for collapse(2) schedule(dynamic) reduction(+ : mag_acc)
      for (int32_t i = 0; i < O->N_time; i++)
        for (int32_t block = 0; block < nb; block += 2) {
          int32_t idx = ((i % 2 == 0) != odd) ? block : block + 1;
          if (idx >= nb) continue;
          int32_t z_start = idx * (int32_t)floor(O->N_z / nb);
          int32_t z_stop = (idx + 1) * (int32_t)floor(O->N_z / nb) - 1;
          z_stop = (idx >= nb - 1) ? O->N_z - 1 : z_stop;
          Real_t m;
          if (!nhicd) {
            int32_t xy_start = j * (int32_t)floor(T->UpdateSelectNum[i][idx] / K);
            int32_t xy_end = (j + 1) * (int32_t)floor(T->UpdateSelectNum[i][idx] / K) - 1;
            xy_end = (j == K - 1) ? T->UpdateSelectNum[i][idx] - 1 : xy_end;
            m = update_lines_masked(S, O, T, E, H_r, i, z_start, z_stop, xy_end - xy_start + 1,
                                    &T->x_rand_select[i][idx][xy_start],
                                    &T->y_rand_select[i][idx][xy_start], Iter,
                                    &zero_count[i * nb + idx], MagUpdateMap[i][idx], Mask[i],
                                    positivity);
            mag_acc += m;
          } else {
            update_lines_masked(S, O, T, E, H_r, i, z_start, z_stop, T->NHICDSelectNum[i][idx],
                                T->x_NHICD_select[i][idx], T->y_NHICD_select[i][idx], Iter,
                                &zero_count[i * nb + idx], MagUpdateMap[i][idx], Mask[i],
                                positivity);
          }
        }
      total_vox_mag += mag_acc;
      if (T->updateProjOffset > 1) oracle_offset_update(S, T, E);
      if (phase == 1) oracle_vsc_select(O, T, MagUpdateMap);
    }
  }
  if (T->updateVar == 1) oracle_variance_update(S, T, E);

  int64_t total_zero_count = 0;
  for (int32_t i = 0; i < O->N_time; i++)
    for (int32_t block = 0; block < nb; block++) {
      total_pix += T->UpdateSelectNum[i][block] * (O->N_z / nb);
      for (int32_t j = 0; j < T->UpdateSelectNum[i][block]; j++)
        AverageUpdate +=
            MagUpdateMap[i][block][T->y_rand_select[i][block][j]][T->x_rand_select[i][block][j]];
      total_zero_count += zero_count[i * nb + block];
    }
  free(zero_count);
  Real_t avg_update_percentage = 100 * AverageUpdate / total_vox_mag;
  dbg(T, "oracle: iter %d avg update %f%% zero count %lld total voxels %f\n", Iter,
      avg_update_percentage, (long long)total_zero_count, total_pix);
  return (avg_update_percentage < T->StopThreshold) ? 1 : 0;
}

/* Timed sample of the reference's hot loop for a CPU baseline: one homogeneous
 * half-phase of updateVoxelsTimeSlices (F/XT_ICD_update.c:799-813) -- OpenMP
 * over (frame, z-block of the frame's parity) with num_z_blocks as the
 * reference derives it from the thread count -- where every task updates the
 * first lines_per_block lines of a random order, all slices of its block.
 * E must hold the error sinogram.  Returns performed voxel updates; the wall
 * time goes to *seconds and the number of concurrent tasks to *tasks. */
int64_t oracle_sample_updates(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E,
                              int32_t lines_per_block, int32_t n_threads, double *seconds,
                              int32_t *tasks) {
  int32_t nb = T->num_z_blocks;
  Real_t **H_r = (Real_t **)alloc2(sizeof(Real_t), S->N_p, XT_DETECTOR_RESPONSE_BINS + 1);
  oracle_detector_response(H_r, S, O, T);
  uint8_t ****Mask = (uint8_t ****)alloc4(sizeof(uint8_t), O->N_time, O->N_y, O->N_x, O->N_z);
  memset(&Mask[0][0][0][0], 1, (size_t)O->N_time * O->N_y * O->N_x * O->N_z);
  Real_t ****Mag = (Real_t ****)alloc4(sizeof(Real_t), O->N_time, nb, O->N_y, O->N_x);
  oracle_random_select(O, T, Mask);
  int64_t *zc = (int64_t *)o_calloc((size_t)O->N_time * nb, sizeof(int64_t));
  int32_t ntask = 0;
  for (int32_t i = 0; i < O->N_time; i++)
    for (int32_t block = 0; block < nb; block += 2) ntask++;
  *tasks = ntask;
#ifdef _OPENMP
  omp_set_num_threads(n_threads);
  double t0 = omp_get_wtime();
#else
  (void)n_threads;
  double t0 = 0;
#endif
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int32_t i = 0; i < O->N_time; i++)
    for (int32_t block = 0; block < nb; block += 2) {
      int32_t idx = (i % 2 == 0) ? block : block + 1;
      if (idx >= nb) continue;
      int32_t z_start = idx * (O->N_z / nb);
      int32_t z_stop = (idx >= nb - 1) ? O->N_z - 1 : (idx + 1) * (O->N_z / nb) - 1;
      int32_t n = lines_per_block < T->UpdateSelectNum[i][idx] ? lines_per_block
                                                                 : T->UpdateSelectNum[i][idx];
      update_lines_masked(S, O, T, E, H_r, i, z_start, z_stop, n, T->x_rand_select[i][idx],
                          T->y_rand_select[i][idx], 1, &zc[i * nb + idx], Mag[i][idx], Mask[i], 0);
    }
#ifdef _OPENMP
  *seconds = omp_get_wtime() - t0;
#else
  *seconds = 0;
#endif
  int64_t upd = 0;
  for (int32_t i = 0; i < O->N_time; i++)
    for (int32_t block = 0; block < nb; block += 2) {
      int32_t idx = (i % 2 == 0) ? block : block + 1;
      if (idx >= nb) continue;
      int32_t z_start = idx * (O->N_z / nb);
      int32_t z_stop = (idx >= nb - 1) ? O->N_z - 1 : (idx + 1) * (O->N_z / nb) - 1;
      int32_t n = lines_per_block < T->UpdateSelectNum[i][idx] ? lines_per_block
                                                                 : T->UpdateSelectNum[i][idx];
      upd += (int64_t)n * (z_stop - z_start + 1) - zc[i * nb + idx];
    }
  free(zc);
  oracle_multifree(Mag, 4);
  oracle_multifree(Mask, 4);
  oracle_multifree(H_r, 2);
  return upd;
}

static int32_t write_bin(const char *dir, const char *name, const void *data, size_t n) {
  char path[1024];
  snprintf(path, sizeof path, "%s%s%s.bin", dir ? dir : "", (dir && dir[0]) ? "/" : "", name);
  FILE *fp = fopen(path, "wb");
  if (!fp) return -1;
  size_t w = fwrite(data, sizeof(Real_t), n, fp);
  fclose(fp);
  return w == n ? 0 : -1;
}

static int32_t read_bin(const char *dir, const char *name, void *data, size_t n, TomoInputs *T) {
  char path[1024];
  snprintf(path, sizeof path, "%s%s%s.bin", dir ? dir : "", (dir && dir[0]) ? "/" : "", name);
  FILE *fp = fopen(path, "rb");
  if (!fp) {
    dbg(T, "ERROR: error in reading file %s\n", path);
    return -1;
  }
  size_t r = fread(data, sizeof(Real_t), n, fp);
  fclose(fp);
  if (r != n)
    dbg(T, "ERROR: Number of elements read does not match required, required = %zu, read = %zu\n",
        n, r);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Edge-only updates, F/XT_ImageProc.c:40-170.  filter_image (:40-55): out[i][j]
 * = sum over k, then l, of h[k + r][l + r] * in[i + k][j + l] inside the image,
 * r = (len - 1) / 2, in Real_t from 0. */
static void edge_filter(Real_t *in, Real_t *out, Real_t h[5][5], int32_t rows, int32_t cols,
                        int32_t len) {
  int32_t i, j, k, l, r = (len - 1) / 2;
  for (i = 0; i < rows; i++)
    for (j = 0; j < cols; j++) {
      Real_t acc = 0;
      for (k = -r; k <= r; k++)
        for (l = -r; l <= r; l++)
          if (i + k >= 0 && j + l >= 0 && i + k < rows && j + l < cols)
            acc += h[k + r][l + r] * in[(i + k) * cols + j + l];
      out[i * cols + j] = acc;
    }
}

/* edge_detect, F/XT_ImageProc.c:83-149: per frame and slice, a 5x5 Gaussian
 * blur (variance 3), the 3x3 gradient pair hx, hy, |gradient| > 0.0002 inside
 * radius_obj (centre from -Length/2), then a 5x5 dilation (image_dilation,
 * :57-81); Mask[n][row][col][slice] = the dilated map.  Two departures make it
 * a defined function: the blur's normaliser starts at 0 (the reference adds
 * into an uninitialised `sum`, :95-100) and slice m is read as
 * Object[n][m + 1][row][col] (the reference reads Object[n][row][col][m],
 * :117, which leaves the object's N_z + 2 planes for N_y > N_z + 2). */
void oracle_edge_detect(ScannedObject *O, TomoInputs *T, uint8_t ****Mask) {
  Real_t hx[5][5] = {{3, 0, -3, 0, 0}, {10, 0, -10, 0, 0}, {3, 0, -3, 0, 0}};
  Real_t hy[5][5] = {{3, 10, 3, 0, 0}, {0, 0, 0, 0, 0}, {-3, -10, -3, 0, 0}};
  Real_t hblur[5][5], var = 3.0, sum = 0, thresh = 0.0002;
  int32_t i, j, n, m, rows = O->N_y, cols = O->N_x;
  for (i = -2; i <= 2; i++)
    for (j = -2; j <= 2; j++) {
      hblur[i + 2][j + 2] = exp(-(i * i + j * j) / (2 * var));
      sum += hblur[i + 2][j + 2];
    }
  for (i = 0; i < 5; i++)
    for (j = 0; j < 5; j++) hblur[i][j] /= sum;
  size_t np = (size_t)rows * cols;
  Real_t *img = (Real_t *)o_calloc(np, sizeof(Real_t));
  Real_t *blur = (Real_t *)o_calloc(np, sizeof(Real_t));
  Real_t *gx = (Real_t *)o_calloc(np, sizeof(Real_t));
  Real_t *gy = (Real_t *)o_calloc(np, sizeof(Real_t));
  uint8_t *edge = (uint8_t *)o_calloc(np, 1);
  for (n = 0; n < O->N_time; n++)
    for (m = 0; m < O->N_z; m++) {
      for (i = 0; i < rows; i++)
        for (j = 0; j < cols; j++) img[i * cols + j] = O->Object[n][m + 1][i][j];
      edge_filter(img, blur, hblur, rows, cols, 5);
      edge_filter(blur, gx, hx, rows, cols, 3);
      edge_filter(blur, gy, hy, rows, cols, 3);
      for (i = 0; i < rows; i++)
        for (j = 0; j < cols; j++) {
          Real_t x = -O->Length_X / 2 + ((Real_t)j + 0.5) * O->delta_xy;
          Real_t y = -O->Length_Y / 2 + ((Real_t)i + 0.5) * O->delta_xy;
          edge[i * cols + j] = 0;
          if (x * x + y * y < T->radius_obj * T->radius_obj) {
            Real_t gxv = gx[i * cols + j], gyv = gy[i * cols + j];
            Real_t mag = sqrt(gxv * gxv + gyv * gyv);
            if (mag > thresh) edge[i * cols + j] = 1;
          }
        }
      for (i = 0; i < rows; i++)
        for (j = 0; j < cols; j++) {
          uint8_t v = 0;
          for (int32_t k = -2; k <= 2; k++)
            for (int32_t l = -2; l <= 2; l++)
              if (i + k >= 0 && j + l >= 0 && i + k < rows && j + l < cols)
                v |= edge[(i + k) * cols + j + l];
          Mask[n][i][j][m] = v;
        }
    }
  free(img);
  free(blur);
  free(gx);
  free(gy);
  free(edge);
}

/* select_edge_voxels, F/XT_ImageProc.c:151-170: the mask, and MagUpdateMap
 * cleared for every (frame, z-block, line) without a masked voxel. */
static void oracle_select_edge_voxels(ScannedObject *O, TomoInputs *T, Real_t ****MagUpdateMap,
                                      uint8_t ****Mask) {
  int32_t i, j, k, l, n, flag, bs = O->N_z / T->num_z_blocks;
  oracle_edge_detect(O, T, Mask);
  for (i = 0; i < O->N_time; i++)
    for (k = 0; k < O->N_y; k++)
      for (l = 0; l < O->N_x; l++)
        for (j = 0; j < T->num_z_blocks; j++) {
          flag = 0;
          for (n = j * bs; n < (j + 1) * bs; n++)
            if (Mask[i][k][l][n] == 1) flag = 1;
          if (flag == 0) MagUpdateMap[i][j][k][l] = 0;
        }
}

/* oracle_update_lines with the voxel mask of the frame ([N_y][N_x][N_z]). */
Real_t oracle_update_lines_mask(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E,
                                Real_t **H_r, int32_t frame, int32_t z0, int32_t z1,
                                int32_t n_lines, const int32_t *x_list, const int32_t *y_list,
                                int32_t Iter, int64_t *zero_count, Real_t **MagUpdateMap,
                                uint8_t ***Mask, int32_t positivity) {
  return update_lines_masked(S, O, T, E, H_r, frame, z0, z1, n_lines, x_list, y_list, Iter,
                             zero_count, MagUpdateMap, Mask, positivity);
}

/* ICD driver, F/XT_ICD_update.c:928-1094.  initICD 0 (zero start), 1 (read
 * object_n0_time_<i>.bin), 2 and 3 (the previous stage's half-size object
 * upsampled, oracle_upsample_object). */
int32_t oracle_icd_backproject(Sinogram *S, ScannedObject *O, TomoInputs *T, const char *dir,
                               int32_t flags, double *cost_out) {
  int32_t i, j, k, l, Iter, flag = 0, iters = 0;
  int32_t do_cost = flags & 1, positivity = (flags >> 1) & 1;
  int32_t nb = T->num_z_blocks;
  Real_t ****MagUpdateMap = (Real_t ****)alloc4(sizeof(Real_t), O->N_time, nb, O->N_y, O->N_x);
  uint8_t ****Mask = (uint8_t ****)alloc4(sizeof(uint8_t), O->N_time, O->N_y, O->N_x, O->N_z);
#ifdef _OPENMP
  /* The z-block count follows num_threads (initStructures); the OpenMP team may
   * be capped separately (ORACLE_OMP_THREADS) so that a host with fewer cores
   * than num_threads runs the same blocks without oversubscription. */
  {
    int32_t nthr = T->num_threads;
    const char *cap = getenv("ORACLE_OMP_THREADS");
    if (cap && atoi(cap) > 0 && atoi(cap) < nthr) nthr = atoi(cap);
    if (nthr > 0) omp_set_num_threads(nthr);
  }
#endif
  for (i = 0; i < O->N_time; i++)
    for (j = 0; j < O->N_y; j++)
      for (k = 0; k < O->N_x; k++) {
        Real_t x = O->x0 + ((Real_t)k + 0.5) * O->delta_xy;
        Real_t y = O->y0 + ((Real_t)j + 0.5) * O->delta_xy;
        uint8_t in = (x * x + y * y < T->radius_obj * T->radius_obj) ? 1 : 0;
        for (l = 0; l < O->N_z; l++) Mask[i][j][k][l] = in;
      }
  Real_t **H_r = (Real_t **)alloc2(sizeof(Real_t), S->N_p, XT_DETECTOR_RESPONSE_BINS + 1);
  oracle_detector_response(H_r, S, O, T);
  Real_t ***E = (Real_t ***)alloc3(sizeof(Real_t), S->N_p, S->N_r, S->N_t);

  /* initObject, F/XT_ICD_update.c:443-475 */
  for (i = 0; i < O->N_time; i++)
    for (j = 0; j < O->N_z; j++)
      for (k = 0; k < O->N_y; k++)
        for (l = 0; l < O->N_x; l++) O->Object[i][j][k][l] = 0;
  if (T->initICD == 1) {
    for (i = 0; i < O->N_time; i++) {
      char name[128];
      snprintf(name, sizeof name, "object_n%d_time_%d", T->node_rank, i);
      read_bin(dir, name, &(O->Object[i][1][0][0]), (size_t)O->N_z * O->N_y * O->N_x, T);
    }
  } else if (T->initICD == 2 || T->initICD == 3) {
    /* F/XT_ICD_update.c:476-512: the previous stage's half-size object, upsampled */
    int32_t nzc = T->initICD == 3 ? O->N_z / 2 : O->N_z;
    size_t nc = (size_t)nzc * (O->N_y / 2) * (O->N_x / 2);
    Real_t *coarse = (Real_t *)o_calloc(nc, sizeof(Real_t));
    for (i = 0; i < O->N_time; i++) {
      char name[128];
      snprintf(name, sizeof name, "object_n%d_time_%d", T->node_rank, i);
      read_bin(dir, name, coarse, nc, T);
      oracle_upsample_object(T->initICD, nzc, O->N_y / 2, O->N_x / 2, coarse,
                             &(O->Object[i][1][0][0]));
    }
    free(coarse);
  } else if (T->initICD != 0) {
    dbg(T, "ERROR: initObject: initICD value not recognized\n");
    return -1;
  }
  /* F/XT_ICD_update.c:996-997 */
  if (T->only_Edge_Updates == 1) oracle_select_edge_voxels(O, T, MagUpdateMap, Mask);
  oracle_init_error_sino(S, O, T, H_r, E);
  oracle_multifree(S->Projection, 3);
  S->Projection = NULL;
  if (cost_out)
    for (i = 0; i <= T->NumIter; i++) cost_out[i] = NAN;
  if (do_cost) {
    Real_t c = oracle_compute_cost(S, O, T, E);
    if (cost_out) cost_out[0] = c;
    dbg(T, "oracle: cost before ICD %f\n", c);
  }
  for (Iter = 1; Iter <= T->NumIter; Iter++) {
    flag = icd_iteration(S, O, T, H_r, E, Iter, MagUpdateMap, Mask, positivity);
    iters = Iter;
    if (do_cost) {
      Real_t c = oracle_compute_cost(S, O, T, E);
      if (cost_out) cost_out[Iter] = c;
      dbg(T, "oracle: ICD iter %d cost %f\n", Iter, c);
    }
    if (flags & 4) { /* snapshot of every iteration's object */
      for (i = 0; i < O->N_time; i++) {
        char name[128];
        snprintf(name, sizeof name, "object_n%d_time_%d_iter_%d", T->node_rank, i, Iter);
        write_bin(dir, name, &(O->Object[i][1][0][0]), (size_t)O->N_z * O->N_y * O->N_x);
      }
      {
        char name[128];
        snprintf(name, sizeof name, "proj_offset_n%d_iter_%d", T->node_rank, Iter);
        write_bin(dir, name, &(S->ProjOffset[0][0]), (size_t)S->N_r * S->N_t);
      }
      if (T->debug_file_ptr) fflush(T->debug_file_ptr);
    }
    if (flag != 0 && Iter > 1) break;
  }
  {
    char name[128];
    snprintf(name, sizeof name, "variance_estimate_n%d", T->node_rank);
    write_bin(dir, name, &T->var_est, 1);
    snprintf(name, sizeof name, "mag_update_map_n%d", T->node_rank);
    write_bin(dir, name, &(MagUpdateMap[0][0][0][0]), (size_t)O->N_time * nb * O->N_y * O->N_x);
  }
  oracle_multifree(E, 3);
  oracle_multifree(H_r, 2);
  oracle_multifree(Mask, 4);
  oracle_multifree(MagUpdateMap, 4);
  return iters;
}

/* ------------------------------------------------------------------------ */
/* view_info.txt tokenizer and parser, F/XT_Init.c:56-75,130-162: a token ended
 * by '-' is a time, a token ended by ',' is an angle in degrees. */
int32_t oracle_parse_view_info(const char *path, int32_t max_views, Real_t *views_rad,
                               Real_t *times) {
  FILE *fp = fopen(path, "r");
  if (!fp) return -1;
  char str[128];
  Real_t time = 0;
  int32_t index = 0;
  for (;;) {
    int32_t n = 0, c, last = EOF;
    while ((c = fgetc(fp)) != EOF) {
      if (isspace(c)) continue;
      if (!isalnum(c) && c != '.') {
        last = c;
        break;
      }
      if (n < (int32_t)sizeof(str) - 1) str[n++] = (char)c;
    }
    str[n] = '\0';
    if (last == EOF) break;
    if (last == '-')
      time = (Real_t)atof(str);
    else if (last == ',') {
      if (str[0] != '\0') {
        if (index < max_views) {
          views_rad[index] = (Real_t)atof(str) * O_PI / 180.0;
          times[index] = time;
        }
        index++;
      }
    } else {
      fclose(fp);
      return -1;
    }
  }
  fclose(fp);
  return index;
}

/* Views to frames by acquisition time, F/XT_Init.c:98-125. */
void oracle_bin_views(const Real_t *times, int32_t N_p, int32_t N_time, Real_t Rtime0,
                      Real_t delta_Rtime, int32_t *proj_idx_flat, int32_t *proj_num) {
  int32_t obj = 0, num = 0;
  Real_t object_time2 = Rtime0 + delta_Rtime;
  for (int32_t t = 0; t < N_time; t++) proj_num[t] = 0;
  for (int32_t i = 0; i < N_p; i++) {
    Real_t projection_time = times[i];
    while (projection_time >= object_time2 && obj + 1 < N_time) {
      proj_num[obj] = num;
      object_time2 += delta_Rtime;
      obj++;
      num = 0;
    }
    proj_idx_flat[(size_t)obj * N_p + num] = i;
    num++;
  }
  proj_num[obj] = num;
}

/* Interlaced view order and clipping, 4D_Recon_Launcher_for_Real_Data/
 * XT_Interlaced_Angles.py:6-27 (clip), 30-34 (bit reversal), 107-139 (0..inf). */
int32_t oracle_interlaced_views(int32_t K, int32_t N_theta, int32_t N_p, double min_time,
                                double rotation_speed, double *views, double *times,
                                double *del_views, double *del_times, int32_t *n_del) {
  int32_t L = N_theta / K, maxbits = 0;
  while ((1 << (maxbits + 1)) <= K) maxbits++;
  double delta_theta = 180.0 / N_theta;
  double *v = (double *)o_calloc(N_p, sizeof(double));
  double *t = (double *)o_calloc(N_p, sizeof(double));
  for (int32_t i = 0; i < N_p; i++) {
    int32_t sub = (i / L) % K, rev = 0;
    for (int32_t b = 0; b < maxbits; b++)
      if (sub & (1 << b)) rev |= 1 << (maxbits - 1 - b);
    double idx = (double)i * K + rev;
    t[i] = idx / K;
    v[i] = idx * delta_theta;
  }
  double min_dTheta = rotation_speed * min_time;
  int32_t keep = 1, rem = 0;
  views[0] = v[0];
  times[0] = 0;
  for (int32_t i = 1; i < N_p; i++) {
    if (v[i] - v[i - 1] > min_dTheta) {
      views[keep] = v[i];
      times[keep] = t[i];
      keep++;
    } else {
      del_views[rem] = v[i];
      del_times[rem] = t[i];
      rem++;
    }
  }
  *n_del = rem;
  free(v);
  free(t);
  return keep;
}

/* ------------------------------------------------------------------------ */
/* Command line (F/XT_Init.c:452-629), geometry (F/XT_Init.c:277-448), prior
 * weights (F/XT_Init.c:199-256), sines/cosines (F/XT_Init.c:259-272) and the
 * binary inputs (F/XT_genSinogram.c:264-300, F/XT_Engine.c:120-123). */
typedef struct {
  const char *name;
  int has_arg;
  char code;
} o_opt;
static const o_opt g_opts[] = {
    {"p", 1, 'a'},           {"sigma_s", 1, 'b'},          {"c_s", 1, 'c'},
    {"sigma_t", 1, 'd'},     {"c_t", 1, 'e'},              {"delta_xy", 1, 'f'},
    {"delta_z", 1, 'z'},     {"length_r", 1, 'g'},         {"num_threads", 1, 'h'},
    {"length_t", 1, '3'},    {"voxel_thresh", 1, 'j'},     {"iter", 1, 'k'},
    {"rotation_center", 1, 'l'}, {"alpha", 1, 'm'},        {"time_reg", 0, 'n'},
    {"sinobin", 1, 'o'},     {"initICD", 1, 'p'},          {"writeTiff", 1, 'q'},
    {"NoNoise", 0, 'r'},     {"Rtime0", 1, 's'},           {"Rtime_delta", 1, 't'},
    {"Rtime_num", 1, 'u'},   {"num_projections", 1, 'v'},  {"N_r", 1, 'w'},
    {"dont_reconstruct", 0, 'x'}, {"cost_thresh", 1, 'y'}, {"radius_obj", 1, 'i'},
    {"detector_slice_begin", 1, '1'}, {"detector_slice_num", 1, '2'},
    {"phantom_N_xy", 1, '4'}, {"phantom_N_z", 1, '5'},     {"N_t", 1, '6'},
    {"updateProjOffset", 1, '7'}, {"EnforceZeroMeanOffset", 0, '#'},
    {"no_NHICD", 0, '8'},    {"WritePerIter", 0, '+'},     {"only_Edge_Updates", 0, '-'},
    {"zingerT", 1, '*'},     {"zingerDel", 1, '^'},        {"initMagUpMap", 0, '&'},
    {"readSino4mHDF", 0, '>'}, {"do_VarEst", 0, '%'},      {"Est_of_Var", 1, '('},
    {"RMSE_converged", 0, ';'}, {"x_start", 1, '['},       {"x_width", 1, ']'},
    {NULL, 0, 0}};

static int32_t parse_args(int32_t argc, char **argv, Sinogram *S, ScannedObject *O,
                          TomoInputs *T) {
  Real_t slope_HU, c_HU;
  slope_HU = (O_HOUNSFIELD_WATER_MAP - O_HOUNSFIELD_AIR_MAP) /
             (O_WATER_MASS_ATT_COEFF * O_WATER_DENSITY - O_AIR_MASS_ATT_COEFF * O_AIR_DENSITY) /
             O_HFIELD_UNIT_CONV_CONST;
  c_HU = -slope_HU * (O_AIR_MASS_ATT_COEFF * O_AIR_DENSITY * O_HFIELD_UNIT_CONV_CONST);
  T->sinobin = 0;
  T->initICD = 0;
  T->Write2Tiff = 0;
  T->time_reg = 0;
  T->reconstruct = 1;
  T->No_Projection_Noise = 0;
  T->cost_thresh = 100;
  T->updateProjOffset = 0;
  T->EnforceZeroMeanOffset = 0;
  T->no_NHICD = 0;
  T->WritePerIter = 0;
  T->only_Edge_Updates = 0;
  T->initMagUpMap = 0;
  T->updateVar = 0;
  T->var_est = 1;
  T->readSino4mHDF = 0;
  T->RMSE_converged = 0;
  for (int32_t a = 0; a < argc; a++) {
    const char *s = argv[a];
    if (s[0] != '-' || s[1] != '-') return -1;
    s += 2;
    const char *eq = strchr(s, '=');
    size_t len = eq ? (size_t)(eq - s) : strlen(s);
    const o_opt *o = NULL;
    for (const o_opt *q = g_opts; q->name; q++)
      if (strlen(q->name) == len && strncmp(q->name, s, len) == 0) o = q;
    if (!o) return -1;
    const char *v = NULL;
    if (o->has_arg) {
      if (eq)
        v = eq + 1;
      else if (a + 1 < argc)
        v = argv[++a];
      else
        return -1;
    }
    switch (o->code) {
      case 'a': O->MRF_P = (Real_t)atof(v); break;
      case 'b': O->Sigma_S = (Real_t)(atof(v) - c_HU) / slope_HU; break;
      case 'c': O->C_S = (Real_t)atof(v); break;
      case 'd': O->Sigma_T = (Real_t)(atof(v) - c_HU) / slope_HU; break;
      case 'e': O->C_T = (Real_t)atof(v); break;
      case 'f': O->mult_xy = (Real_t)atof(v); break;
      case 'z': O->mult_z = (Real_t)atof(v); break;
      case 'g': S->Length_R = (Real_t)atof(v); break;
      case 'h': T->num_threads = (int32_t)atoi(v); break;
      case '3': S->Length_T = (Real_t)atof(v); break;
      case 'j': T->StopThreshold = (Real_t)atof(v); break;
      case 'k': T->NumIter = (int32_t)atoi(v); break;
      case 'l': T->RotCenter = (Real_t)atof(v); break;
      case 'm': T->alpha = (Real_t)atof(v); break;
      case 'n': T->time_reg = 1; break;
      case 'o': T->sinobin = (uint8_t)atoi(v); break;
      case 'p': T->initICD = (uint8_t)atoi(v); break;
      case 'q': T->Write2Tiff = (uint8_t)atoi(v); break;
      case 'r': T->No_Projection_Noise = 1; break;
      case 's': O->Rtime0 = (Real_t)atof(v); break;
      case 't': O->delta_Rtime = (Real_t)atof(v); break;
      case 'u': O->N_time = (int32_t)atoi(v); break;
      case 'v': S->N_p = (int32_t)atoi(v); break;
      case 'w': S->N_r = (int32_t)atoi(v); break;
      case 'x': T->reconstruct = 0; break;
      case 'y': T->cost_thresh = (Real_t)atof(v); break;
      case 'i': T->radius_obj = (Real_t)atof(v); break;
      case '1': S->slice_begin = (int32_t)atoi(v); break;
      case '2': S->slice_num = (int32_t)atoi(v); break;
      case '4': T->phantom_N_xy = (int32_t)atoi(v); break;
      case '5': T->phantom_N_z = (int32_t)atoi(v); break;
      case '6': S->total_t_slices = (int32_t)atoi(v); break;
      case '7': T->updateProjOffset = (int32_t)atoi(v); break;
      case '#': T->EnforceZeroMeanOffset = 1; break;
      case '8': T->no_NHICD = 1; break;
      case '+': T->WritePerIter = 1; break;
      case '-': T->only_Edge_Updates = 1; break;
      case '*': T->ErrorSinoThresh = (Real_t)atof(v); break;
      case '^': T->ErrorSinoDelta = (Real_t)atof(v); break;
      case '&': T->initMagUpMap = 1; break;
      case '>': T->readSino4mHDF = 1; break;
      case '%': T->updateVar = 1; break;
      case '(': T->var_est = (Real_t)atof(v); break;
      case ';': T->RMSE_converged = 1; break;
      case '[': S->x_start = (int32_t)atoi(v); break;
      case ']': S->x_width = (int32_t)atoi(v); break;
    }
  }
  /* sinobin 1: F/XT_Init.c:617-622 */
  O->delta_xy = O->mult_xy;
  O->delta_z = O->mult_z;
  S->z_overlap_num = O->mult_z;
  return 0;
}

static Real_t distance2node(uint8_t i, uint8_t j, uint8_t k, uint8_t l) {
  return (sqrt(pow((Real_t)i - 1.0, 2.0) + pow((Real_t)j - 1.0, 2.0) + pow((Real_t)k - 1.0, 2.0) +
               pow((Real_t)l - 1.0, 2.0)));
}

static void init_filter(ScannedObject *O, TomoInputs *T) {
  uint8_t i, j, k;
  Real_t temp1, sum = 0, prior_const = 0;
  prior_const = O->mult_xy * O->mult_xy * O->mult_xy * O->delta_Rtime;
  T->Time_Filter[0] = 1.0 / distance2node(0, 1, 1, 1);
  sum += 2.0 * T->Time_Filter[0];
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++)
      for (k = 0; k < 3; k++) {
        if (i != 1 || j != 1 || k != 1) {
          temp1 = 1.0 / distance2node(1, i, j, k);
          T->Spatial_Filter[i][j][k] = temp1;
          sum = sum + temp1;
        } else
          T->Spatial_Filter[i][j][k] = 0;
      }
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++)
      for (k = 0; k < 3; k++)
        T->Spatial_Filter[i][j][k] = prior_const * T->Spatial_Filter[i][j][k] / sum;
  T->Time_Filter[0] = prior_const * T->Time_Filter[0] / sum;
}

static int32_t init_structures(const char *dir, Sinogram *S, ScannedObject *O, TomoInputs *T) {
  char name[1024];
  if (T->initICD != 0) {
    snprintf(name, sizeof name, "variance_estimate_n%d", T->node_rank);
    if (read_bin(dir, name, &T->var_est, 1, T)) return -1;
  }
  S->Length_T = S->Length_T / T->node_num;
  S->N_t = S->total_t_slices / T->node_num;
  if (S->N_t % (int32_t)O->delta_z != 0) return -1;
  if (S->N_r % (int32_t)O->delta_xy != 0) return -1;
  S->delta_r = S->Length_R / (S->N_r);
  S->delta_t = S->Length_T / (S->N_t);
  {
    /* F/XT_Init.c:284,439-441 */
    Real_t lap[3][3] = {{-0.5, -1, -0.5}, {-1, 6, -1}, {-0.5, -1, -0.5}};
    for (int32_t i = 0; i < 3; i++)
      for (int32_t j = 0; j < 3; j++) S->Lap_Kernel[i][j] = lap[i][j] / (S->delta_r * S->delta_r);
  }
  S->R0 = -T->RotCenter * S->delta_r;
  S->RMax = (S->N_r - T->RotCenter) * S->delta_r;
  S->T0 = -S->Length_T / 2.0;
  S->TMax = S->Length_T / 2.0;
  O->Length_X = S->Length_R;
  O->Length_Y = S->Length_R;
  O->Length_Z = S->Length_T;
  O->N_x = (int32_t)(S->N_r / O->delta_xy);
  O->N_y = (int32_t)(S->N_r / O->delta_xy);
  O->N_z = (int32_t)(S->N_t / O->delta_z);
  O->delta_xy = O->delta_xy * S->delta_r;
  O->delta_z = O->delta_z * S->delta_t;
  S->Projection = (Real_t ***)alloc3(sizeof(Real_t), S->N_p, S->N_r, S->N_t);
  S->ProjSelect = (bool ***)alloc3(sizeof(bool), S->N_p, S->N_r, S->N_t);
  memset(&(S->ProjSelect[0][0][0]), 1, (size_t)S->N_p * S->N_r * S->N_t * sizeof(bool));
  S->ProjOffset = (Real_t **)alloc2(sizeof(Real_t), S->N_r, S->N_t);
  T->Weight = (Real_t ***)alloc3(sizeof(Real_t), S->N_p, S->N_r, S->N_t);
  if (T->updateProjOffset == 1 || T->updateProjOffset == 3) {
    snprintf(name, sizeof name, "proj_offset_n%d", T->node_rank);
    if (read_bin(dir, name, &(S->ProjOffset[0][0]), (size_t)S->N_r * S->N_t, T)) return -1;
  }
  O->x0 = S->R0;
  O->z0 = S->T0;
  O->y0 = -O->Length_Y / 2.0;
  O->BeamWidth = S->delta_r;
  O->Object = (Real_t ****)o_calloc(O->N_time, sizeof(Real_t ***));
  for (int32_t i = 0; i < O->N_time; i++)
    O->Object[i] = (Real_t ***)alloc3(sizeof(Real_t), O->N_z + 2, O->N_y, O->N_x);
  S->OffsetR = (O->delta_xy / sqrt(2.0) + S->delta_r / 2.0) / XT_DETECTOR_RESPONSE_BINS;
  S->OffsetT = ((O->delta_z / 2) + S->delta_t / 2) / XT_DETECTOR_RESPONSE_BINS;
  T->Sigma_S_Q = pow((O->Sigma_S * O->mult_xy), XT_MRF_Q);
  T->Sigma_S_Q_P = pow(O->Sigma_S * O->mult_xy, XT_MRF_Q - O->MRF_P);
  T->Sigma_T_Q = pow((O->Sigma_T * O->delta_Rtime), XT_MRF_Q);
  T->Sigma_T_Q_P = pow(O->Sigma_T * O->delta_Rtime, XT_MRF_Q - O->MRF_P);
  T->num_z_blocks = 2 * floor(T->num_threads / O->N_time);
  if (T->num_z_blocks < 2)
    T->num_z_blocks = 2;
  else if (T->num_z_blocks > O->N_z)
    T->num_z_blocks = O->N_z;
  int32_t nb = T->num_z_blocks, nxy = O->N_y * O->N_x;
  T->x_rand_select = (int32_t ***)alloc3(sizeof(int32_t), O->N_time, nb, nxy);
  T->y_rand_select = (int32_t ***)alloc3(sizeof(int32_t), O->N_time, nb, nxy);
  T->x_NHICD_select = (int32_t ***)alloc3(sizeof(int32_t), O->N_time, nb, nxy);
  T->y_NHICD_select = (int32_t ***)alloc3(sizeof(int32_t), O->N_time, nb, nxy);
  T->UpdateSelectNum = (int32_t **)alloc2(sizeof(int32_t), O->N_time, nb);
  T->NHICDSelectNum = (int32_t **)alloc2(sizeof(int32_t), O->N_time, nb);
  O->NHICD_Iterations = 10;
  init_filter(O, T);
  /* initRandomAngles, F/XT_Init.c:165-197 */
  S->ViewPtr = (Real_t *)o_calloc(S->N_p, sizeof(Real_t));
  S->TimePtr = (Real_t *)o_calloc(S->N_p, sizeof(Real_t));
  O->ProjIdxPtr = (int32_t **)o_calloc(O->N_time, sizeof(int32_t *));
  O->ProjNum = (int32_t *)o_calloc(O->N_time, sizeof(int32_t));
  snprintf(name, sizeof name, "%s%sview_info.txt", dir ? dir : "", (dir && dir[0]) ? "/" : "");
  int32_t nv = oracle_parse_view_info(name, S->N_p, S->ViewPtr, S->TimePtr);
  if (nv != S->N_p) {
    dbg(T, "ERROR: Number of projections read is %d and expected is %d\n", nv, S->N_p);
    return -1;
  }
  int32_t *flat = (int32_t *)o_calloc((size_t)O->N_time * S->N_p, sizeof(int32_t));
  oracle_bin_views(S->TimePtr, S->N_p, O->N_time, O->Rtime0, O->delta_Rtime, flat, O->ProjNum);
  for (int32_t i = 0; i < O->N_time; i++) {
    O->ProjIdxPtr[i] = (int32_t *)o_calloc(O->ProjNum[i] ? O->ProjNum[i] : 1, sizeof(int32_t));
    memcpy(O->ProjIdxPtr[i], flat + (size_t)i * S->N_p, O->ProjNum[i] * sizeof(int32_t));
  }
  free(flat);
  S->cosine = (Real_t *)o_calloc(S->N_p, sizeof(Real_t));
  S->sine = (Real_t *)o_calloc(S->N_p, sizeof(Real_t));
  for (int32_t i = 0; i < S->N_p; i++) {
    S->cosine[i] = cos(S->ViewPtr[i]);
    S->sine[i] = sin(S->ViewPtr[i]);
  }
  return 0;
}

static void free_structures(Sinogram *S, ScannedObject *O, TomoInputs *T) {
  for (int32_t i = 0; i < O->N_time; i++) {
    free(O->ProjIdxPtr[i]);
    oracle_multifree(O->Object[i], 3);
  }
  free(O->ProjIdxPtr);
  free(O->ProjNum);
  free(O->Object);
  oracle_multifree(T->x_rand_select, 3);
  oracle_multifree(T->y_rand_select, 3);
  oracle_multifree(T->x_NHICD_select, 3);
  oracle_multifree(T->y_NHICD_select, 3);
  oracle_multifree(T->UpdateSelectNum, 2);
  oracle_multifree(T->NHICDSelectNum, 2);
  oracle_multifree(S->ProjOffset, 2);
  oracle_multifree(S->ProjSelect, 3);
  oracle_multifree(T->Weight, 3);
  if (S->Projection) oracle_multifree(S->Projection, 3);
  free(S->ViewPtr);
  free(S->TimePtr);
  free(S->cosine);
  free(S->sine);
}

int32_t oracle_engine_setup(const char *dir, int32_t argc, char **argv, Sinogram *S,
                            ScannedObject *O, TomoInputs *T) {
  memset(S, 0, sizeof *S);
  memset(O, 0, sizeof *O);
  memset(T, 0, sizeof *T);
  oracle_srandom2(761521); /* F/XT_Engine.c:137 */
  T->node_num = 1;
  T->node_rank = 0;
  if (parse_args(argc, argv, S, O, T) != 0) return 1;
  if (T->sinobin != 1) return 1;
  if (init_structures(dir, S, O, T) != 0) return 1;
  return 0;
}

int32_t oracle_engine_read_sinogram(const char *dir, Sinogram *S, TomoInputs *T) {
  size_t nsino = (size_t)S->N_p * S->N_r * S->N_t;
  if (read_bin(dir, "projection_n0", &(S->Projection[0][0][0]), nsino, T) ||
      read_bin(dir, "weight_n0", &(T->Weight[0][0][0]), nsino, T))
    return 1;
  for (size_t q = 0; q < nsino; q++) (&(T->Weight[0][0][0]))[q] /= T->var_est;
  return 0;
}

void oracle_engine_free(Sinogram *S, ScannedObject *O, TomoInputs *T) {
  if (O->Object) free_structures(S, O, T);
  O->Object = NULL;
}

int32_t oracle_engine_run(const char *dir, int32_t argc, char **argv, int32_t flags,
                          double *cost_out, int32_t *iters_out) {
  Sinogram S;
  ScannedObject O;
  TomoInputs T;
  int32_t rc = oracle_engine_setup(dir, argc, argv, &S, &O, &T);
  char logpath[1024];
  snprintf(logpath, sizeof logpath, "%s%sDEBUG_oracle.log", dir ? dir : "",
           (dir && dir[0]) ? "/" : "");
  T.debug_file_ptr = fopen(logpath, "w");
  if (rc == 0) rc = oracle_engine_read_sinogram(dir, &S, &T);
  if (rc == 0 && T.reconstruct == 1) {
    int32_t it = oracle_icd_backproject(&S, &O, &T, dir, flags, cost_out);
    if (it < 0) rc = 1;
    if (iters_out) *iters_out = it;
    if (rc == 0) {
      for (int32_t i = 0; i < O.N_time; i++) {
        char name[128];
        snprintf(name, sizeof name, "object_n%d_time_%d", T.node_rank, i);
        write_bin(dir, name, &(O.Object[i][1][0][0]), (size_t)O.N_z * O.N_y * O.N_x);
      }
      write_bin(dir, "proj_offset_n0", &(S.ProjOffset[0][0]), (size_t)S.N_r * S.N_t);
    }
  }
  oracle_engine_free(&S, &O, &T);
  if (T.debug_file_ptr) fclose(T.debug_file_ptr);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Multi-resolution upsampling, Source_Code_4D_Fast/XT_ICD_update.c:316-435.
 * Along one axis of n coarse samples c[0..n-1] the 2n fine samples are
 *   f[0] = c[0], f[1] = (3 c[0] + c[1]) / 4, f[2n-2] = (c[n-2] + 3 c[n-1]) / 4,
 *   f[2n-1] = c[n-1], and for 0 < k < n-1: f[2k] = (c[k-1] + 3 c[k]) / 4,
 *   f[2k+1] = (3 c[k] + c[k+1]) / 4.
 * Every weighted sum is formed in double (the reference's 3.0 literals) and
 * rounded to Real_t once, except the interior odd rows of the y pass, whose
 * reference text multiplies by the integer 3 and so sums in Real_t before the
 * division by 4.0 (:339,:411). */
static Real_t up_even(Real_t prev, Real_t cur) { return (Real_t)((prev + 3.0 * cur) / 4.0); }
static Real_t up_odd(Real_t cur, Real_t next) { return (Real_t)((3.0 * cur + next) / 4.0); }
static Real_t up_odd_real(Real_t cur, Real_t next) {
  Real_t t = (Real_t)3 * cur;
  t = t + next;
  return (Real_t)(t / 4.0);
}

/* one axis: n coarse samples at stride `cs` -> 2n fine samples at stride `fs`;
 * `real_odd` selects the Real_t form of the interior odd samples */
static void up_axis(const Real_t *c, ptrdiff_t cs, int32_t n, Real_t *f, ptrdiff_t fs, int real_odd) {
  f[0] = c[0];
  f[fs] = up_odd(c[0], c[cs]);
  f[(2 * n - 1) * fs] = c[(n - 1) * cs];
  f[(2 * n - 2) * fs] = up_even(c[(n - 2) * cs], c[(n - 1) * cs]);
  for (int32_t k = 1; k < n - 1; k++) {
    f[2 * k * fs] = up_even(c[(k - 1) * cs], c[k * cs]);
    f[(2 * k + 1) * fs] =
        real_odd ? up_odd_real(c[k * cs], c[(k + 1) * cs]) : up_odd(c[k * cs], c[(k + 1) * cs]);
  }
}

void oracle_upsample_object(int32_t init_icd, int32_t Nzc, int32_t Nyc, int32_t Nxc,
                            const Real_t *coarse, Real_t *fine) {
  const int32_t Nx = 2 * Nxc, Ny = 2 * Nyc;
  Real_t *bx = malloc(sizeof(Real_t) * (size_t)Nzc * Nyc * Nx);
  Real_t *by = init_icd == 3 ? malloc(sizeof(Real_t) * (size_t)Nzc * Ny * Nx) : fine;
  for (int32_t z = 0; z < Nzc; z++)
    for (int32_t j = 0; j < Nyc; j++)  /* x pass */
      up_axis(coarse + ((size_t)z * Nyc + j) * Nxc, 1, Nxc, bx + ((size_t)z * Nyc + j) * Nx, 1, 0);
  for (int32_t z = 0; z < Nzc; z++)
    for (int32_t k = 0; k < Nx; k++)  /* y pass */
      up_axis(bx + (size_t)z * Nyc * Nx + k, Nx, Nyc, by + (size_t)z * Ny * Nx + k, Nx, 1);
  if (init_icd == 3) {
    for (int32_t j = 0; j < Ny; j++)  /* z pass */
      for (int32_t k = 0; k < Nx; k++)
        up_axis(by + (size_t)j * Nx + k, (ptrdiff_t)Ny * Nx, Nzc, fine + (size_t)j * Nx + k,
                (ptrdiff_t)Ny * Nx, 0);
    free(by);
  }
  free(bx);
}

/* ------------------------------------------------------------------------ */
/* Counts to sinogram, Source_Code_4D_Fast/XT_HDFIO.c:160-199: the white and
 * dark frames averaged per detector pixel in Real_t, then per (view, bin,
 * row) the absolute count-minus-dark summed over the ratio_r x ratio_t pixels
 * of the bin (weight), the same for white-minus-dark (bright), and
 * temp = log(bright / weight) rounded to Real_t, projection =
 * BH * temp * temp + temp in double (the macro's double literal). */
void oracle_counts_to_sinogram(int32_t N_p, int32_t T_in, int32_t R_in, int32_t N_t, int32_t N_r,
                               const uint16_t *counts, int32_t n_white, const uint16_t *white,
                               int32_t n_dark, const uint16_t *dark, double bh_quad_coef,
                               Real_t *projection, Real_t *weight, Real_t *bright) {
  const int32_t rr = R_in / N_r, rt = T_in / N_t;
  const size_t plane = (size_t)T_in * R_in;
  Real_t *w2 = (Real_t *)o_calloc(plane, sizeof(Real_t));  /* [R_in][T_in] as the reference */
  Real_t *d2 = (Real_t *)o_calloc(plane, sizeof(Real_t));
  for (int32_t j = 0; j < R_in; j++)
    for (int32_t k = 0; k < T_in; k++) {
      Real_t sw = 0, sd = 0;
      for (int32_t i = 0; i < n_white; i++) sw += white[(size_t)i * plane + (size_t)k * R_in + j];
      for (int32_t i = 0; i < n_dark; i++) sd += dark[(size_t)i * plane + (size_t)k * R_in + j];
      sw /= n_white;
      sd /= n_dark;
      w2[(size_t)j * T_in + k] = sw;
      d2[(size_t)j * T_in + k] = sd;
    }
  for (int32_t j = 0; j < N_r; j++)
    for (int32_t k = 0; k < N_t; k++) {
      Real_t b = 0;
      for (int32_t m = 0; m < rr; m++)
        for (int32_t n = 0; n < rt; n++) {
          const size_t q = (size_t)(j * rr + m) * T_in + (k * rt + n);
          b += fabs(w2[q] - d2[q]);
        }
      bright[(size_t)j * N_t + k] = b;
    }
  for (int32_t i = 0; i < N_p; i++)
    for (int32_t j = 0; j < N_r; j++)
      for (int32_t k = 0; k < N_t; k++) {
        Real_t w = 0;
        for (int32_t m = 0; m < rr; m++)
          for (int32_t n = 0; n < rt; n++) {
            const int32_t r = j * rr + m, t = k * rt + n;
            w += fabs(counts[(size_t)i * plane + (size_t)t * R_in + r] - d2[(size_t)r * T_in + t]);
          }
        const size_t o = ((size_t)i * N_r + j) * N_t + k;
        Real_t temp = log(bright[(size_t)j * N_t + k] / w);
        weight[o] = w;
        projection[o] = bh_quad_coef * temp * temp + temp;
      }
  free(w2);
  free(d2);
}
