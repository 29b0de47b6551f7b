/*
 * xt_oracle.h -- CPU oracle for the TIMBIR ICD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (Source_Code_4D_Fast), used as the checker for the HIP library and
 * as the timed CPU baseline.  Only tests/, the smoke check and the benchmark's
 * CPU-baseline leg may load it; the product library never links or calls it.
 *
 * Built twice from one source: liboracle_f32.so (Real_t = float, the
 * reference's production build) and liboracle_f64.so (XT_REAL_IS_DOUBLE, the
 * reference's -DREAL_IS_DOUBLE build).
 *
 * Parity status of each function (how it is pinned):
 *   oracle_random2/srandom2      known-answer table Source_Code_4D_Fast/randlib.c:69-91
 *                                and bit-exact vs the reference's randlib.c (oracle/_ref)
 *   oracle_detector_response     bit-exact vs the reference's XT_Profile.c (oracle/_ref)
 *                                plus closed-form checks (symmetry, area)
 *   oracle_amatrix_column        bit-exact vs the reference's XT_AMatrix.c (oracle/_ref)
 *   oracle_forward_project       vs the reference's forward_project_voxel (oracle/_ref)
 *                                summed over all voxels, and dense-matrix brute force
 *   oracle_vsc_select            bit-exact vs the reference's XT_NHICD.c (oracle/_ref)
 *   oracle_random_select         each (frame, block)'s masked lines selected once; N_y*N_x
 *                                draws of the pinned RNG; first and first-half positions
 *                                uniform over 4000 seeds (three planted mistakes fail)
 *   oracle_parse_view_info       the reference's view_info.txt fixtures
 *   oracle_interlaced_views      golden log Source_Code_4D_Fast/info.txt:1-175
 *   oracle_counts_to_sinogram    closed forms: flat fields give 0, Beer-Lambert ratios
 *                                give their logarithm; block sums brute-forced
 *                                (tests/test_oracle_pins.py)
 *   oracle_upsample_object       closed forms: constants and linear ramps reproduced
 *                                exactly (tests/test_oracle_pins.py)
 *   voxel update / surrogate     closed forms: quadratic prior + no zingers reduces ICD to
 *                                Gauss-Seidel on (A'WA+H)x = A'W(y-d), solved densely (one
 *                                frame, and two frames with the temporal term);
 *                                qGGMRF derivative vs finite differences of the value;
 *                                theta1/theta2 read back from updates = the data cost's
 *                                derivative along the voxel (central differences) and
 *                                its majorizer, with rejected entries
 *   offset update                each pixel's step minimises the symmetric IRLS majorizer
 *                                of the data cost; zero-mean steps = the KKT solution
 *   variance                     closed form mean(e^2 w); its fixed point is the cost's
 *                                stationary point in sigma^2
 *   cost                         closed form on a zero object; prior half = brute-force
 *                                sum over unordered 26-neighbour pairs (3D, two frames)
 *   prior filter weights         inverse-distance ratios, centre 0, sum mult_xy^3 dRtime
 *   full ICD driver              converges to the dense MAP solution (quadratic case);
 *                                the reference's own ICD driver (XT_ICD_update.c) cannot
 *                                be built here (it needs mpi.h), so end-to-end trajectory
 *                                parity with the reference binary is unpinned.
 */
#ifndef XT_ORACLE_H
#define XT_ORACLE_H

#include "../include/xt_structures.h"

#ifdef __cplusplus
extern "C" {
#endif

void oracle_srandom2(uint32_t seed);
double oracle_random2(void);
void oracle_random2_seq(uint32_t seed, int32_t n, double *out);

void *oracle_multialloc(size_t s, int32_t d, const int64_t *dims);
void oracle_multifree(void *r, int32_t d);

void oracle_detector_response(Real_t **H_r, Sinogram *S, ScannedObject *O, TomoInputs *T);
void oracle_amatrix_column(Sinogram *S, ScannedObject *O, Real_t **H_r, AMatrixCol *Ai,
                           int32_t row, int32_t col, int32_t proj_idx);
/* Ax for one frame's views, accumulated in Real_t into AX (zeroed first). */
void oracle_forward_project(Sinogram *S, ScannedObject *O, Real_t **H_r, Real_t ***AX);
/* e = y - Ax - d and ProjSelect, as initErrorSinogam. */
void oracle_init_error_sino(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t **H_r,
                            Real_t ***E);
Real_t oracle_compute_cost(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E);
void oracle_offset_update(Sinogram *S, TomoInputs *T, Real_t ***E);
void oracle_variance_update(Sinogram *S, TomoInputs *T, Real_t ***E);
Real_t oracle_functional_substitution(Real_t V, Real_t THETA1, Real_t THETA2, ScannedObject *O,
                                      TomoInputs *T, const Real_t *nhood27, const uint8_t *flag27,
                                      const Real_t *time_nhood2, const uint8_t *time_flag2,
                                      int32_t positivity);
/* Updates the listed (y,x) voxel lines of frame i, slices [z0,z1], in list order. */
Real_t oracle_update_lines(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E, Real_t **H_r,
                           int32_t frame, int32_t z0, int32_t z1, int32_t n_lines,
                           const int32_t *x_list, const int32_t *y_list, int32_t Iter,
                           int64_t *zero_count, Real_t **MagUpdateMap, int32_t positivity);
Real_t oracle_update_lines_mask(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E,
                                Real_t **H_r, int32_t frame, int32_t z0, int32_t z1,
                                int32_t n_lines, const int32_t *x_list, const int32_t *y_list,
                                int32_t Iter, int64_t *zero_count, Real_t **MagUpdateMap,
                                uint8_t ***Mask, int32_t positivity);
/* Edge-only update mask [N_time][N_y][N_x][N_z] of the current object. */
void oracle_edge_detect(ScannedObject *O, TomoInputs *T, uint8_t ****Mask);
void oracle_vsc_select(ScannedObject *O, TomoInputs *T, Real_t ****MagUpdateMap);
void oracle_random_select(ScannedObject *O, TomoInputs *T, uint8_t ****Mask);

/* Timed sample of one homogeneous half-phase of the reference's schedule (the
 * benchmark's CPU baseline); see xt_oracle.c. */
int64_t oracle_sample_updates(Sinogram *S, ScannedObject *O, TomoInputs *T, Real_t ***E,
                              int32_t lines_per_block, int32_t n_threads, double *seconds,
                              int32_t *tasks);

/* Full driver: same semantics as ICD_BackProject.  flags: bit0 compute the MAP
 * cost every iteration (the reference's build without NO_COST_CALCULATE),
 * bit1 positivity constraint, bit2 write object_n0_time_<i>_iter_<k>.bin
 * after every iteration.  cost_out (may be NULL) receives NumIter+1
 * costs (index 0 = before ICD; unused entries NaN).  Returns iterations run. */
int32_t oracle_icd_backproject(Sinogram *S, ScannedObject *O, TomoInputs *T, const char *dir,
                               int32_t flags, double *cost_out);

/* Whole engine: parse the reference CLI, read the inputs from dir, run ICD, write
 * object_n0_time_<i>.bin and proj_offset_n0.bin into dir.  Returns 0 on success. */
int32_t oracle_engine_run(const char *dir, int32_t argc, char **argv, int32_t flags,
                          double *cost_out, int32_t *iters_out);

/* Engine pieces: parse the reference CLI and run initStructures (reads
 * view_info.txt and, when asked, proj_offset_n0.bin / variance_estimate_n0.bin
 * from dir), read projection_n0.bin / weight_n0.bin, free what setup allocated. */
int32_t oracle_engine_setup(const char *dir, int32_t argc, char **argv, Sinogram *S,
                            ScannedObject *O, TomoInputs *T);
int32_t oracle_engine_read_sinogram(const char *dir, Sinogram *S, TomoInputs *T);
void oracle_engine_free(Sinogram *S, ScannedObject *O, TomoInputs *T);

/* view_info.txt parse (returns views read, or -1 on a format error). */
int32_t oracle_parse_view_info(const char *path, int32_t max_views, Real_t *views_rad,
                               Real_t *times);
/* Frame binning; proj_idx_flat is [N_time][N_p] row-major, proj_num [N_time]. */
void oracle_bin_views(const Real_t *times, int32_t N_p, int32_t N_time, Real_t Rtime0,
                      Real_t delta_Rtime, int32_t *proj_idx_flat, int32_t *proj_num);
/* Multi-resolution initial object (initObject with initICD 2 or 3,
 * Source_Code_4D_Fast/XT_ICD_update.c:316-435,484-512): the coarse
 * [Nzc][Nyc][Nxc] reconstruction upsampled 2x along y and x (init_icd 2; fine
 * [Nzc][2Nyc][2Nxc]) or along z, y and x (init_icd 3; fine [2Nzc][2Nyc][2Nxc]),
 * with the reference's mix of double and Real_t arithmetic.  Each coarse size
 * that is upsampled must be >= 2. */
void oracle_upsample_object(int32_t init_icd, int32_t Nzc, int32_t Nyc, int32_t Nxc,
                            const Real_t *coarse, Real_t *fine);

/* Phase-contrast tomography on (1) or off (0) for every later call (the
 * reference's -DPHASE_CONTRAST_TOMOGRAPHY build: Laplacian footprints,
 * F/XT_AMatrix.c:113-187, F/XT_ForwardProject.c:27-44, F/XT_VoxUpdate.c:590-845). */
void oracle_set_phase_contrast(int32_t on);

/* Counts to sinogram, gen_projection_4m_HDF's arithmetic
 * (Source_Code_4D_Fast/XT_HDFIO.c:160-199) on the region of interest already
 * read: counts [N_p][T_in][R_in], white [n_white][T_in][R_in], dark
 * [n_dark][T_in][R_in] (the frames the reference averages); outputs
 * projection and weight [N_p][N_r][N_t] and bright [N_r][N_t], with
 * ratio_r = R_in / N_r and ratio_t = T_in / N_t detector pixels summed per bin. */
void oracle_counts_to_sinogram(int32_t N_p, int32_t T_in, int32_t R_in, int32_t N_t, int32_t N_r,
                               const uint16_t *counts, int32_t n_white, const uint16_t *white,
                               int32_t n_dark, const uint16_t *dark, double bh_quad_coef,
                               Real_t *projection, Real_t *weight, Real_t *bright);

/* Interlaced bit-reversed order and clipping (4D_Recon_Launcher_for_Real_Data/
 * XT_Interlaced_Angles.py:6-27,30-34,107-139).  Returns the kept count; the
 * deleted views/times go to del_views/del_times, their count to *n_del. */
int32_t oracle_interlaced_views(int32_t K, int32_t N_theta, int32_t N_p, double min_time,
                                double rotation_speed, double *views, double *times,
                                double *del_views, double *del_times, int32_t *n_del);

#ifdef __cplusplus
}
#endif

#endif
