/*
 * xt_structures.h -- the engine API structs of the TIMBIR ICD path, restated.
 *
 * These three structs are the ABI of the drop-in boundary: a host that was
 * built against the reference's Source_Code_4D_Fast/XT_Structures.h:61-203
 * passes pointers to them into ICD_BackProject().  Field order, field types
 * and the Real_t switch therefore have to be identical to the reference;
 * everything else about this library is its own design.
 *
 *   Real_t  float   (default, as Source_Code_4D_Fast/XT_Constants.h:49-53)
 *           double  when XT_REAL_IS_DOUBLE is defined (REAL_IS_DOUBLE there)
 *
 * Array fields are pointer trees over one contiguous block (the layout
 * Source_Code_4D_Fast/allocate.c:91-171 builds), so &X[0][0][0] addresses
 * the whole array in row-major order:
 *   Projection, Weight, ProjSelect, ErrorSino   [N_p][N_r][N_t]   (t fastest)
 *   ProjOffset                                  [N_r][N_t]
 *   Object                                      [N_time] -> [N_z+2][N_y][N_x]
 *                                               (z slices 0 and N_z+1 are halo)
 */
#ifndef XT_STRUCTURES_H_AMD
#define XT_STRUCTURES_H_AMD

#include <stdint.h>
#include <stdio.h>
#include <stdbool.h>

#ifdef XT_REAL_IS_DOUBLE
typedef double Real_t;
#else
typedef float Real_t;
#endif

/* Neighbourhood and table sizes (Source_Code_4D_Fast/XT_Constants.h:123-132). */
#define XT_PROFILE_RESOLUTION 1536
#define XT_BEAM_RESOLUTION 512
#define XT_DETECTOR_RESPONSE_BINS 64
#define XT_NHOOD_Z_MAXDIM 3
#define XT_NHOOD_Y_MAXDIM 3
#define XT_NHOOD_X_MAXDIM 3
#define XT_NHOOD_TIME_MAXDIM 3
#define XT_MRF_Q 2.0

#ifdef __cplusplus
extern "C" {
#endif

/* Sinogram and detector (Source_Code_4D_Fast/XT_Structures.h:61-89). */
typedef struct {
  Real_t ***Projection;    /* y = log(I0/counts) */
  Real_t **ProjOffset;     /* additive ring offset d[r][t] */
  bool ***ProjSelect;      /* |e*sqrt(w)| < T classification */
  int32_t x_start;
  int32_t x_width;
  int32_t N_r;
  int32_t N_t;
  int32_t N_p;
  int32_t total_t_slices;
  Real_t delta_r;
  Real_t delta_t;
  Real_t R0, RMax;
  Real_t T0, TMax;
  Real_t *cosine, *sine;
  Real_t Length_R;
  Real_t Length_T;
  Real_t OffsetR;
  Real_t OffsetT;
  Real_t *ViewPtr;         /* view angles in radians */
  Real_t *TimePtr;         /* acquisition times */
  int32_t slice_begin;
  int32_t slice_num;
  int32_t z_overlap_num;
  Real_t Lap_Kernel[3][3];
} Sinogram;

/* Reconstructed object and prior (Source_Code_4D_Fast/XT_Structures.h:91-127). */
typedef struct {
  Real_t ****Object;
  Real_t ****Conv_Object;
  Real_t Length_X;
  Real_t Length_Y;
  Real_t Length_Z;
  int32_t N_x;
  int32_t N_z;
  int32_t N_y;
  int32_t N_time;
  Real_t x0;
  Real_t z0;
  Real_t y0;
  Real_t delta_xy;
  Real_t delta_z;
  Real_t mult_xy;
  Real_t mult_z;
  Real_t BeamWidth;
  Real_t Sigma_S;
  Real_t Sigma_T;
  Real_t MRF_P;
  Real_t C_S;
  Real_t C_T;
  int32_t **ProjIdxPtr;    /* [N_time][ProjNum[i]] view indices of each frame */
  int32_t *ProjNum;
  int32_t NHICD_Iterations;
  Real_t Rtime0;
  Real_t delta_Rtime;
} ScannedObject;

/* One sparse system-matrix column (Source_Code_4D_Fast/XT_Structures.h:130-135). */
typedef struct {
  Real_t *values;
  uint8_t count;
  int32_t *index;
} AMatrixCol;

/* Algorithm inputs and scratch handles (Source_Code_4D_Fast/XT_Structures.h:137-203). */
typedef struct {
  int32_t NumIter;
  Real_t StopThreshold;
  Real_t RotCenter;
  int32_t phantom_N_xy;
  int32_t phantom_N_z;
  Real_t radius_obj;
  Real_t Sigma_S_Q;
  Real_t Sigma_T_Q;
  Real_t Sigma_S_Q_P;
  Real_t Sigma_T_Q_P;
  Real_t Spatial_Filter[XT_NHOOD_Z_MAXDIM][XT_NHOOD_Y_MAXDIM][XT_NHOOD_X_MAXDIM];
  Real_t Time_Filter[(XT_NHOOD_TIME_MAXDIM - 1) / 2];
  Real_t ***Weight;        /* w (already divided by var_est) */
  Real_t var_est;
  Real_t alpha;
  Real_t cost_thresh;
  uint8_t time_reg;
  uint8_t sinobin;
  uint8_t initICD;
  uint8_t Write2Tiff;
  uint8_t No_Projection_Noise;
  uint8_t reconstruct;
  int32_t num_threads;
  uint8_t updateProjOffset;
  uint8_t EnforceZeroMeanOffset;
  uint8_t no_NHICD;
  uint8_t WritePerIter;
  int32_t num_z_blocks;
  int32_t ***x_NHICD_select;
  int32_t ***y_NHICD_select;
  int32_t ***x_rand_select;
  int32_t ***y_rand_select;
  int32_t **UpdateSelectNum;
  int32_t **NHICDSelectNum;
  Real_t ErrorSinoThresh;
  Real_t ErrorSinoDelta;
  uint8_t only_Edge_Updates;
  int32_t node_num;
  int32_t node_rank;
  uint8_t updateVar;
  uint8_t initMagUpMap;
  uint8_t readSino4mHDF;
  FILE *debug_file_ptr;
  Real_t ErrorSino_Cost;
  Real_t Forward_Cost;
  Real_t Prior_Cost;
  uint8_t RMSE_converged;
} TomoInputs;

#ifdef __cplusplus
}
#endif

#endif /* XT_STRUCTURES_H_AMD */
