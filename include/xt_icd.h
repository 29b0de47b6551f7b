/*
 * xt_icd.h -- C-ABI of libxt_hip.so, the MI355X (gfx950) ICD engine.
 *
 * Two builds of the same sources: libxt_hip.so with Real_t = float (the
 * reference's production build) and libxt_hip_f64.so with Real_t = double
 * (its -DREAL_IS_DOUBLE build; compile the host with XT_REAL_IS_DOUBLE).  Real_t
 * is the type of every struct field and host array below; device arrays are
 * float in both.
 *
 * Plain C: structs from xt_structures.h, raw pointers and sizes; no torch or
 * HIP types.  Every entry point returns 0 on success and a non-zero XT_E*
 * code on failure (the reference exits the process instead; a library must
 * not), with the message retrievable from xt_last_error().
 *
 * Drop-in entry point
 *   ICD_BackProject       replaces Source_Code_4D_Fast/XT_ICD_update.h:40
 *                         (defined XT_ICD_update.c:928-1094)
 *
 * Engine around it (the reference's XT_Engine.c main, for hosts that drive
 * the engine through its command line and files)
 *   xt_engine_setup / xt_engine_read_sinogram / xt_engine_write_outputs /
 *   xt_engine_free / xt_engine_run
 *
 * Test seams: one entry per hot-path function, each computing on the GPU what
 * the cited reference function computes on the CPU, with host arrays in and
 * out.  They allocate device memory per call and are not meant to be fast.
 */
#ifndef XT_ICD_H_AMD
#define XT_ICD_H_AMD

#include "xt_structures.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  XT_OK = 0,
  XT_E_ARG = 1,      /* unsupported or inconsistent input */
  XT_E_HIP = 2,      /* HIP runtime error (no device, out of memory, launch failure) */
  XT_E_IO = 3,       /* file could not be read or written */
};

/* Message for the last non-zero return on this thread ("" if none). */
const char *xt_last_error(void);

/* Library build identification: "libxt_hip gfx950 Real_t=float" or "...=double". */
const char *xt_version(void);

/* ------------------------------------------------------------------------
 * Drop-in boundary.  Same contract as the reference function:
 *  - reads Sinogram {Projection, ProjOffset, cosine, sine, ViewPtr, N_p, N_r,
 *    N_t, delta_r, delta_t, R0, RMax, OffsetR, z_overlap_num}, ScannedObject
 *    {N_x, N_y, N_z, N_time, x0, y0, delta_xy, delta_z, mult_xy, BeamWidth,
 *    MRF_P, C_S, C_T, ProjIdxPtr, ProjNum, NHICD_Iterations} and TomoInputs
 *    {Weight, var_est, Sigma_*_Q*, Spatial_Filter, Time_Filter, alpha, NumIter,
 *    StopThreshold, radius_obj, ErrorSinoThresh, ErrorSinoDelta,
 *    updateProjOffset, EnforceZeroMeanOffset, updateVar, no_NHICD, initICD,
 *    num_z_blocks, node_rank, node_num};
 *  - writes Object slices 1..N_z of every frame, ProjOffset, ProjSelect,
 *    Weight and var_est (when the variance is re-estimated), and the files
 *    variance_estimate_n<rank>.bin and mag_update_map_n<rank>.bin in the
 *    current directory;
 *  - frees SinogramPtr->Projection (pointer tree, F/XT_ICD_update.c:999);
 *  - returns 0 on success.
 * One process, one GPU (device 0 unless xt_set_option(XT_OPT_DEVICE) says
 * otherwise).  Not reentrant. */
int ICD_BackProject(Sinogram *SinogramPtr, ScannedObject *ScannedObjectPtr,
                    TomoInputs *TomoInputsPtr);

/* Options the reference fixes at compile time (-D macros) or not at all. */
enum {
  XT_OPT_POSITIVITY = 1,     /* POSITIVITY_CONSTRAINT: clamp voxels at 0 (default 0) */
  XT_OPT_COST = 2,           /* compute the MAP cost each iteration, i.e. a build without
                                NO_COST_CALCULATE (default 0) */
  XT_OPT_RESYNC_EVERY = 3,   /* FP64 recompute of e = y - Ax - d every n iterations; 0 = never
                                (default 0) */
  XT_OPT_SEED = 4,           /* seed of the random line order of every iteration, frame and
                                z-block parity (default 761521) */
  XT_OPT_LINES = 5,          /* voxel lines of a frame updated at once (workgroups per frame
                                and launch); 0 = automatic: min(1024, max(16, lines / 64)) */
  XT_OPT_DEVICE = 6,         /* HIP device ordinal (default 0) */
  XT_OPT_VERBOSE = 7,        /* per-iteration lines to the debug file (default 1) */
  XT_OPT_OUT_DIR = 8,        /* reserved */
  XT_OPT_ATABLE = 9,         /* 1: keep every line's A columns in an HBM table built once
                                (when it fits and each column spans <= 4 bins); 0: build
                                them on every visit (default 1) */
  XT_OPT_GRAPH = 10,         /* 1: capture one iteration's launches as a hipGraph and replay
                                it from the second iteration on (not while per-launch timing
                                is on); 0: enqueue every launch (default 1) */
  XT_OPT_PACK = 11,          /* 1: a pass over one z-block parity packs that parity's slices
                                into half the lanes when the blocks are of equal size and
                                even in number; 0: one lane per slice (default 1) */
  XT_OPT_NHICD_LINES = 12,   /* 0: NHICD selects per z-block as the reference does
                                (F/XT_NHICD.c:178-205); 1: per z-block parity on the
                                parity's summed update map, so a selected line is updated in
                                every block of the parity: the same number of voxel updates
                                in fewer, fuller line visits (default 1) */
  XT_OPT_INIT_FBP = 13,      /* 1: with initICD 0, ICD_BackProject starts from the filtered
                                backprojection of y - d instead of zero (no counterpart in
                                the reference; default 0) */
  XT_OPT_AUTOTUNE = 14,      /* 1: when XT_OPT_LINES is 0, xt_icd_create times one
                                homogeneous sub-iteration with the automatic number of lines
                                in flight and with 1/2 and 1/4 of it, restores the state, and
                                keeps the fastest (default 1) */
  XT_OPT_BATCH_FUSE = 15,    /* 1: xt_icd_batch_iterate runs each voxel-update pass of all
                                slabs as one launch per kernel variant (slabs of the same
                                line width), the slabs' own small kernels between them;
                                0: each slab's iteration as its own graph branch (default 0:
                                one slab's lines in flight already fill the workgroup slots
                                the line kernel's registers allow, so the fused passes run the
                                slabs nearly in turn, with a barrier per pass) */
  XT_OPT_LINE_THREADS = 16,  /* threads of a voxel-update workgroup (one line at a time):
                                256, 64 (one wave: 4x more lines progress at once, for
                                short lines and few views per frame), or 0 to let the
                                create-time autotuner choose (default 0) */
  XT_OPT_PHASE_CONTRAST = 17, /* 1: phase-contrast tomography (the reference's
                                -DPHASE_CONTRAST_TOMOGRAPHY build): every footprint is the
                                detector Laplacian of the absorption one (Lap_Kernel), slices
                                of a line are updated in turn; absorption starts (filtered
                                backprojection, network) are refused (default 0) */
  XT_OPT_BH_QUAD_COEF = 18,  /* BH_QUAD_COEF of the reference build: the beam-hardening
                                quadratic coefficient of the counts read with sinobin 3
                                (default 0) */
  XT_OPT_LINE_GROUP = 19     /* lines a voxel-update workgroup takes at once.  Only 1 is
                                accepted in this build: groups of 2 (1 x 2 tiles) and 4
                                (2 x 2 tiles), updated in turn and writing the error
                                sinogram once per group, are switched off (default 1) */
};
int xt_set_option(int option, double value);
double xt_get_option(int option);

/* ------------------------------------------------------------------------
 * Stepwise form of ICD_BackProject for hosts that time or steer iterations
 * themselves (the benchmark).  xt_icd_create uploads the inputs, builds the
 * detector table and the error sinogram; xt_icd_iterate runs n iterations
 * (returns 1 in *stop when the stop threshold was met); xt_icd_download writes
 * the results into the structs as ICD_BackProject does.  Projection is not
 * freed by this form. */
typedef struct xt_icd_state xt_icd_state;

typedef struct {
  int32_t iterations;           /* iterations run so far */
  double voxel_updates;         /* performed (not zero-skipped) voxel updates */
  double footprint_entries;     /* sum over performed updates of the sinogram
                                   footprint F_v = sum_views count * z_overlap_num */
  double icd_kernel_ms;         /* summed duration of the voxel-update launches
                                   (valid when timing is on) */
  int64_t icd_launches;
  double last_cost;             /* MAP cost after the last iteration (XT_OPT_COST) */
  double stop_percentage;       /* 100 * sum|dx| / sum|x|, the reference's stop test */
  int32_t stopped;
  double line_visits;           /* workgroup visits of a voxel line (a visit updates the line's
                                   slices of one z-block parity) */
  int32_t lines_in_flight;      /* voxel lines of a frame updated at once (autotuned) */
  int32_t line_threads;         /* threads per voxel-update workgroup (autotuned) */
} xt_icd_stats;

int xt_icd_create(Sinogram *S, ScannedObject *O, TomoInputs *T, xt_icd_state **state);
/* xt_icd_create with y and w copied on the device from sino_from, a state over the
 * same sinogram (N_p, N_r, N_t, frames and HBM layout equal, same device), instead of
 * uploaded from S->Projection and T->Weight, which are not read: for a host that
 * reconstructs one sinogram at several voxel sizes (the multi-resolution stages), with
 * no counterpart in the reference.  w is copied as sino_from holds it at the call. */
int xt_icd_create_from(Sinogram *S, ScannedObject *O, TomoInputs *T,
                       const xt_icd_state *sino_from, xt_icd_state **state);
int xt_icd_iterate(xt_icd_state *state, int32_t n_iter, int32_t *stop);
int xt_icd_set_timing(xt_icd_state *state, int32_t on);
int xt_icd_get_stats(xt_icd_state *state, xt_icd_stats *stats);
int xt_icd_cost(xt_icd_state *state, double *cost);          /* FP64 MAP cost now */
int xt_icd_resync(xt_icd_state *state);                       /* FP64 e = y - Ax - d now */
/* The inverse of xt_icd_download: the object (slices 1..N_z of every frame) and
 * ProjOffset of the structs become the state's x and d, and e = y - Ax - d is
 * recomputed in FP64 (a stepwise host's restart from a volume of its own, e.g.
 * the previous stage of a multi-resolution run). */
int xt_icd_upload(xt_icd_state *state, Sinogram *S, ScannedObject *O, TomoInputs *T);
/* Replace the object by the filtered backprojection of y - d (per frame: ramp
 * filter of sample spacing delta_r after a 3-point median along the bins that
 * removes isolated zingers; backprojection with the A-matrix geometry and the
 * views' angular weights), then recompute e.  When the ring offsets are
 * estimated (updateProjOffset > 0), d is first set to a guess: the view-mean of
 * y without outliers minus its 15-bin running mean along the bins (zero-mean if
 * enforced).
 * A warm start for ICD with no counterpart in the reference. */
int xt_icd_fbp(xt_icd_state *state);
int xt_icd_download(xt_icd_state *state, Sinogram *S, ScannedObject *O, TomoInputs *T);
int xt_icd_synchronize(xt_icd_state *state);
void xt_icd_destroy(xt_icd_state *state);

/* Batches of independent reconstructions (slabs of different sizes, each from
 * xt_icd_create, all at the same iteration).  xt_icd_batch_iterate advances
 * every slab by n iterations; from the second iteration on one captured
 * hipGraph holds each slab's iteration as its own branch, so the slabs overlap
 * on the GPU.  Results per slab are those of xt_icd_iterate.  The batch does
 * not own the states.  No counterpart in the reference (one slab per process). */
typedef struct xt_icd_batch xt_icd_batch;
int xt_icd_batch_create(int32_t n, xt_icd_state **states, xt_icd_batch **batch);
int xt_icd_batch_iterate(xt_icd_batch *batch, int32_t n_iter);
void xt_icd_batch_destroy(xt_icd_batch *batch);

/* ------------------------------------------------------------------------
 * Engine around the boundary: the reference's command line
 * (F/XT_Init.c:452-629), initStructures (F/XT_Init.c:277-448), the binary
 * inputs (F/XT_genSinogram.c:264-300, F/XT_Engine.c:120-123) and outputs
 * (F/XT_IOMisc.c:224-254).  dir is the run folder (NULL or "" = cwd). */
int xt_engine_setup(const char *dir, int argc, char **argv, Sinogram *S, ScannedObject *O,
                    TomoInputs *T);
/* Reads the inputs as genSinogram_fromBin (sinobin 1: projection_n0.bin,
 * weight_n0.bin) or genSinogram_subsampleCounts (sinobin 3: bright_n0.bin and
 * weight_n0.bin counts at mult_xy x mult_z times the detector resolution, binned;
 * beam-hardening coefficient XT_OPT_BH_QUAD_COEF), then divides Weight by var_est
 * (F/XT_Engine.c:98-123). */
int xt_engine_read_sinogram(const char *dir, Sinogram *S, ScannedObject *O, TomoInputs *T);
int xt_engine_write_outputs(const char *dir, Sinogram *S, ScannedObject *O, TomoInputs *T);
void xt_engine_free(Sinogram *S, ScannedObject *O, TomoInputs *T);
/* setup + read + ICD_BackProject + write + free, as XT_Engine.c main. */
int xt_engine_run(const char *dir, int argc, char **argv);

/* Pointer-tree allocator with the reference's layout (F/allocate.c:91-188), so
 * arrays the boundary frees (Projection) can be allocated by any caller. */
void *xt_multialloc(size_t s, int d, const int64_t *dims);
void xt_multifree(void *r, int d);

/* ------------------------------------------------------------------------
 * Test seams. */

/* initObject's multi-resolution upsampling (Source_Code_4D_Fast/XT_ICD_update.c:
 * upsample_object_bilinear_2D :355-388 for initICD 2, upsample_object_bilinear_3D
 * :389-435 for initICD 3): the coarse [Nzc][Nyc][Nxc] object doubled along y, x
 * (fine [Nzc][2Nyc][2Nxc]) or along z, y, x (fine [2Nzc][2Nyc][2Nxc]) on the GPU,
 * bit-exact with the reference's float/double steps.  ICD_BackProject uses it for
 * --initICD 2/3 (the previous stage's object_n<r>_time_<i>.bin at half size). */
int xt_upsample_object(int32_t init_icd, int32_t Nzc, int32_t Nyc, int32_t Nxc,
                       const Real_t *coarse, Real_t *fine);

/* Real-data ingest, gen_projection_4m_HDF's arithmetic (Source_Code_4D_Fast/
 * XT_HDFIO.c:160-199; the launcher's generate_projections,
 * 4D_Recon_Launcher_for_Real_Data/XT_Projections.py:28-123): uint16 counts
 * [N_p][T_in][R_in] and white / dark frames [n][T_in][R_in] of the region of
 * interest (the frames the reference averages, already read from the HDF5
 * file) become projection and weight [N_p][N_r][N_t] and the bright field
 * [N_r][N_t]: per bin, ratio_r = R_in/N_r by ratio_t = T_in/N_t detector pixels
 * summed, weight = sum |counts - dark|, bright = sum |white - dark|,
 * projection = BH t^2 + t with t = log(bright / weight).  On the GPU; weights
 * and bright field bit-exact with the reference's float sums. */
int xt_counts_to_sinogram(int32_t N_p, int32_t T_in, int32_t R_in, int32_t N_t, int32_t N_r,
                          const uint16_t *counts, int32_t n_white, const uint16_t *white,
                          int32_t n_dark, const uint16_t *dark, double bh_quad_coef,
                          Real_t *projection, Real_t *weight, Real_t *bright);

/* DetectorResponseProfile, F/XT_Profile.c:128-217: H_r[N_p][65]. */
int xt_detector_response(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t **H_r);

/* calcAMatrixColumnforAngle, F/XT_AMatrix.c:50-110, for n (row, col, view)
 * triples at once: count[n], index[n][max_entries], values[n][max_entries]. */
int xt_amatrix_columns(Sinogram *S, ScannedObject *O, Real_t **H_r, int32_t n,
                       const int32_t *rows, const int32_t *cols, const int32_t *views,
                       int32_t max_entries, uint8_t *count, int32_t *index, Real_t *values);

/* Ax of every frame (initErrorSinogam's projection, F/XT_ICD_update.c:551-569,
 * F/XT_ForwardProject.c:8-25), accumulated in FP64: AX64[N_p][N_r][N_t]. */
int xt_forward_project(Sinogram *S, ScannedObject *O, Real_t **H_r, double *AX64);

/* initErrorSinogam, F/XT_ICD_update.c:538-597: E = y - Ax - d, ProjSelect. */
int xt_init_error_sino(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t **H_r,
                       Real_t ***E);

/* computeCost, F/XT_ICD_update.c:181-314, in FP64. */
int xt_compute_cost(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E, double *cost);

/* update_Sinogram_Offset, F/XT_ICD_update.c:666-759. */
int xt_offset_update(Sinogram *S, TomoInputs *T, Real_t ***E);

/* The ring-offset update of xt_offset_update on device-resident arrays, with no
 * device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  E and W are device arrays
 * [N_p][N_r][N_t] (E updated in place), ProjOffset a device array [N_r][N_t]
 * (updated in place); zinger_T / zinger_delta are ErrorSinoThresh /
 * ErrorSinoDelta, zero_mean is EnforceZeroMeanOffset.  Results are bit-identical
 * to xt_offset_update on the same values.  Its FP64 scratch is one 8 MiB array
 * per GPU, held until the process ends, which limits the detector to about
 * 131072 pixels (N_r * N_t); larger sizes return XT_E_ARG.  Runs on the null
 * stream and returns when done.  Calls that overlap on one GPU share the scratch
 * array and race. */
int xt_offset_update_noalloc(int32_t N_p, int32_t N_r, int32_t N_t, float zinger_T,
                             float zinger_delta, int32_t zero_mean, float *E, const float *W,
                             float *ProjOffset);

/* update_variance_parameter, F/XT_ICD_update.c:633-664. */
int xt_variance_update(Sinogram *S, TomoInputs *T, Real_t ***E);

/* VSC_based_Voxel_Line_Select, F/XT_NHICD.c:164-216: fills NHICDSelectNum and
 * x/y_NHICD_select from MagUpdateMap[N_time][num_z_blocks][N_y][N_x].  The
 * threshold is the same order statistic; no random numbers are drawn. */
int xt_vsc_select(ScannedObject *O, TomoInputs *T, Real_t ****MagUpdateMap);

/* updateVoxels + compute_voxel_update, F/XT_VoxUpdate.c:11-254: updates the
 * listed voxel lines of one frame, one line at a time in list order, every
 * slice of a line (slices in red-black order), with E kept in FP32.
 * MagUpdate[N_y][N_x] receives the summed |change| of each updated line.
 * Returns the summed |new value| of performed updates in *total_vox_mag.
 * Mask: null, or the voxel mask [N_time][N_y][N_x][N_z] (updateVoxels' Mask,
 * F/XT_VoxUpdate.c:138): voxels where it is 0 are left alone. */
int xt_update_lines(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E, int32_t frame,
                    int32_t n_lines, const int32_t *x_list, const int32_t *y_list, int32_t Iter,
                    Real_t *MagUpdate, double *total_vox_mag, const uint8_t *Mask);

/* The production form of the voxel-update pass on listed lines of a 3D slab
 * (N_time = 1): the same reference functions as xt_update_lines, but as the
 * ICD iterations run them -- the slices of the z-blocks of one parity (0: even
 * blocks, 1: odd, F/XT_ICD_update.c:799-881) updated at once per line, with the
 * error sinogram split by parity and those slices packed into half the lanes
 * when the blocks are of equal size and even in number.  Lines in list order,
 * one at a time.  Blocks of one parity share no sinogram entries and no prior
 * term, so the result equals the reference's per-block loop over the same lines
 * (each block of the parity in turn).  MagUpdate[N_y][N_x] receives the summed
 * |change| of each line over the parity's blocks. */
int xt_update_lines_parity(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E,
                           int32_t n_lines, const int32_t *x_list, const int32_t *y_list,
                           int32_t Iter, int32_t parity, Real_t *MagUpdate,
                           double *total_vox_mag);

/* edge_detect, F/XT_ImageProc.c:83-149: the edge-only update mask
 * Mask[N_time][N_y][N_x][N_z] (one contiguous block, as the reference's
 * multialloc) of O->Object: per frame and slice a 5x5 Gaussian blur, the 3x3
 * gradient pair, |gradient| > 0.0002 inside radius_obj, a 5x5 dilation.  Slice m
 * is read from Object[n][m + 1] and the blur is normalised by its own sum (the
 * reference indexes the object transposed and sums into an uninitialised
 * variable).  ICD_BackProject applies it when only_Edge_Updates is set. */
int xt_edge_detect(ScannedObject *O, TomoInputs *T, uint8_t ****Mask);

/* Measurement probe (no counterpart in the reference): the rate, in FP32 atomic
 * lane operations per second, of the voxel update's write pattern alone (per view
 * of a line three bins, sixteen contiguous floats each, no-return buffer atomics;
 * 1024 views x 768 bins, lines over a 512 x 512 grid) -- the ceiling of the line
 * kernel's e update on this device.  Allocates and frees its own 50 MB buffer. */
int xt_probe_footprint_atomics(int32_t n_lines, double *lane_ops_per_s);

#ifdef __cplusplus
}
#endif

#endif /* XT_ICD_H_AMD */
