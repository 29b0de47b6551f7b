// xt_engine.cpp -- the host around the drop-in boundary: the reference
// engine's command line, structure set-up, binary inputs and outputs, so a run
// folder prepared for Source_Code_4D_Fast/XT_Engine can be reconstructed by
// this library unchanged.
//
//   xt_engine_setup         argsParser (F/XT_Init.c:452-629) + initStructures
//                           (F/XT_Init.c:277-448): geometry, prior weights,
//                           view_info.txt and frame binning, sines/cosines
//   xt_engine_read_sinogram genSinogram_fromBin (F/XT_genSinogram.c:264-300)
//                           and Weight /= var_est (F/XT_Engine.c:120-123)
//   xt_engine_write_outputs write_ObjectProjOff2TiffBinPerIter's binary part
//                           (F/XT_IOMisc.c:224-254)
//   xt_engine_run           main (F/XT_Engine.c:128-172)
//
// Options parse as "--name value", "--name=value" or "--flag", with the
// reference's option names.  Errors return XT_E_ARG / XT_E_IO instead of
// exiting the process.
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xt_icd.h"

extern std::string g_xt_run_dir;

namespace {

thread_local char g_msg[512];

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
  fprintf(stderr, "xt_engine: %s\n", g_msg);
  return code;
}

std::string join(const char *dir, const char *name) {
  if (!dir || !dir[0]) return name;
  return std::string(dir) + "/" + name;
}

int read_reals(const char *dir, const char *name, Real_t *dst, size_t n) {
  std::string p = join(dir, name);
  FILE *fp = fopen(p.c_str(), "rb");
  if (!fp) return fail(XT_E_IO, "cannot open %s", p.c_str());
  size_t r = fread(dst, sizeof(Real_t), n, fp);
  fclose(fp);
  if (r != n) return fail(XT_E_IO, "%s holds %zu of %zu values", p.c_str(), r, n);
  return XT_OK;
}

int write_reals(const char *dir, const char *name, const Real_t *src, size_t n) {
  std::string p = join(dir, name);
  FILE *fp = fopen(p.c_str(), "wb");
  if (!fp) return fail(XT_E_IO, "cannot create %s", p.c_str());
  size_t w = fwrite(src, sizeof(Real_t), n, fp);
  fclose(fp);
  if (w != n) return fail(XT_E_IO, "short write to %s", p.c_str());
  return XT_OK;
}

template <typename T>
T tree(int d, std::initializer_list<int64_t> dims) {
  std::vector<int64_t> v(dims);
  return (T)xt_multialloc(sizeof(typename std::remove_pointer<typename std::remove_pointer<
                                typename std::remove_pointer<T>::type>::type>::type),
                            d, v.data());
}

// HU <-> attenuation, F/XT_Constants.h:93-100 and F/XT_Init.c:510-511
const double kAirAtt = 0.496372335005353, kWaterAtt = 0.521225397034623;
const double kWaterDen = 1.0, kAirDen = 0.001205, kUnit = 0.0001;

struct Opt {
  const char *name;
  bool arg;
  int code;
};
enum {
  O_P = 1, O_SIGS, O_CS, O_SIGT, O_CT, O_DXY, O_DZ, O_LR, O_NTHR, O_LT, O_VTH, O_ITER, O_ROT,
  O_ALPHA, O_TREG, O_SINOBIN, O_INITICD, O_TIFF, O_NONOISE, O_RT0, O_RTD, O_RTN, O_NP, O_NR,
  O_NOREC, O_CTH, O_RAD, O_SB, O_SN, O_PHXY, O_PHZ, O_NT, O_UPO, O_ZMO, O_NONH, O_WPI, O_EDGE,
  O_ZT, O_ZD, O_IMU, O_HDF, O_VAR, O_VEST, O_RMSE, O_XS, O_XW
};
const Opt kOpts[] = {
    {"p", true, O_P}, {"sigma_s", true, O_SIGS}, {"c_s", true, O_CS}, {"sigma_t", true, O_SIGT},
    {"c_t", true, O_CT}, {"delta_xy", true, O_DXY}, {"delta_z", true, O_DZ},
    {"length_r", true, O_LR}, {"num_threads", true, O_NTHR}, {"length_t", true, O_LT},
    {"voxel_thresh", true, O_VTH}, {"iter", true, O_ITER}, {"rotation_center", true, O_ROT},
    {"alpha", true, O_ALPHA}, {"time_reg", false, O_TREG}, {"sinobin", true, O_SINOBIN},
    {"initICD", true, O_INITICD}, {"writeTiff", true, O_TIFF}, {"NoNoise", false, O_NONOISE},
    {"Rtime0", true, O_RT0}, {"Rtime_delta", true, O_RTD}, {"Rtime_num", true, O_RTN},
    {"num_projections", true, O_NP}, {"N_r", true, O_NR}, {"dont_reconstruct", false, O_NOREC},
    {"cost_thresh", true, O_CTH}, {"radius_obj", true, O_RAD},
    {"detector_slice_begin", true, O_SB}, {"detector_slice_num", true, O_SN},
    {"phantom_N_xy", true, O_PHXY}, {"phantom_N_z", true, O_PHZ}, {"N_t", true, O_NT},
    {"updateProjOffset", true, O_UPO}, {"EnforceZeroMeanOffset", false, O_ZMO},
    {"no_NHICD", false, O_NONH}, {"WritePerIter", false, O_WPI},
    {"only_Edge_Updates", false, O_EDGE}, {"zingerT", true, O_ZT}, {"zingerDel", true, O_ZD},
    {"initMagUpMap", false, O_IMU}, {"readSino4mHDF", false, O_HDF}, {"do_VarEst", false, O_VAR},
    {"Est_of_Var", true, O_VEST}, {"RMSE_converged", false, O_RMSE}, {"x_start", true, O_XS},
    {"x_width", true, O_XW}};

int parse_args(int argc, char **argv, Sinogram *S, ScannedObject *O, TomoInputs *T) {
  const Real_t slope_HU =
      (Real_t)((1000 - 0) / (kWaterAtt * kWaterDen - kAirAtt * kAirDen) / kUnit);
  const Real_t c_HU = (Real_t)(-(double)slope_HU * (kAirAtt * kAirDen * kUnit));
  // defaults, F/XT_Init.c:513-531
  T->var_est = 1;
  T->reconstruct = 1;
  T->cost_thresh = 100;
  for (int a = 0; a < argc; a++) {
    const char *s = argv[a];
    if (!s || s[0] != '-' || s[1] != '-') return fail(XT_E_ARG, "unexpected argument '%s'", s ? s : "");
    s += 2;
    const char *eq = strchr(s, '=');
    size_t len = eq ? (size_t)(eq - s) : strlen(s);
    const Opt *o = nullptr;
    for (const Opt &q : kOpts)
      if (strlen(q.name) == len && strncmp(q.name, s, len) == 0) o = &q;
    if (!o) return fail(XT_E_ARG, "cannot recognize argument --%.*s", (int)len, s);
    const char *v = nullptr;
    if (o->arg) {
      if (eq)
        v = eq + 1;
      else if (a + 1 < argc)
        v = argv[++a];
      else
        return fail(XT_E_ARG, "--%s needs a value", o->name);
    }
    switch (o->code) {
      case O_P: O->MRF_P = (Real_t)atof(v); break;
      case O_SIGS: O->Sigma_S = (Real_t)(atof(v) - c_HU) / slope_HU; break;
      case O_CS: O->C_S = (Real_t)atof(v); break;
      case O_SIGT: O->Sigma_T = (Real_t)(atof(v) - c_HU) / slope_HU; break;
      case O_CT: O->C_T = (Real_t)atof(v); break;
      case O_DXY: O->mult_xy = (Real_t)atof(v); break;
      case O_DZ: O->mult_z = (Real_t)atof(v); break;
      case O_LR: S->Length_R = (Real_t)atof(v); break;
      case O_NTHR: T->num_threads = atoi(v); break;
      case O_LT: S->Length_T = (Real_t)atof(v); break;
      case O_VTH: T->StopThreshold = (Real_t)atof(v); break;
      case O_ITER: T->NumIter = atoi(v); break;
      case O_ROT: T->RotCenter = (Real_t)atof(v); break;
      case O_ALPHA: T->alpha = (Real_t)atof(v); break;
      case O_TREG: T->time_reg = 1; break;
      case O_SINOBIN: T->sinobin = (uint8_t)atoi(v); break;
      case O_INITICD: T->initICD = (uint8_t)atoi(v); break;
      case O_TIFF: T->Write2Tiff = (uint8_t)atoi(v); break;
      case O_NONOISE: T->No_Projection_Noise = 1; break;
      case O_RT0: O->Rtime0 = (Real_t)atof(v); break;
      case O_RTD: O->delta_Rtime = (Real_t)atof(v); break;
      case O_RTN: O->N_time = atoi(v); break;
      case O_NP: S->N_p = atoi(v); break;
      case O_NR: S->N_r = atoi(v); break;
      case O_NOREC: T->reconstruct = 0; break;
      case O_CTH: T->cost_thresh = (Real_t)atof(v); break;
      case O_RAD: T->radius_obj = (Real_t)atof(v); break;
      case O_SB: S->slice_begin = atoi(v); break;
      case O_SN: S->slice_num = atoi(v); break;
      case O_PHXY: T->phantom_N_xy = atoi(v); break;
      case O_PHZ: T->phantom_N_z = atoi(v); break;
      case O_NT: S->total_t_slices = atoi(v); break;
      case O_UPO: T->updateProjOffset = (uint8_t)atoi(v); break;
      case O_ZMO: T->EnforceZeroMeanOffset = 1; break;
      case O_NONH: T->no_NHICD = 1; break;
      case O_WPI: T->WritePerIter = 1; break;
      case O_EDGE: T->only_Edge_Updates = 1; break;
      case O_ZT: T->ErrorSinoThresh = (Real_t)atof(v); break;
      case O_ZD: T->ErrorSinoDelta = (Real_t)atof(v); break;
      case O_IMU: T->initMagUpMap = 1; break;
      case O_HDF: T->readSino4mHDF = 1; break;
      case O_VAR: T->updateVar = 1; break;
      case O_VEST: T->var_est = (Real_t)atof(v); break;
      case O_RMSE: T->RMSE_converged = 1; break;
      case O_XS: S->x_start = atoi(v); break;
      case O_XW: S->x_width = atoi(v); break;
    }
  }
  if (T->sinobin != 1 && T->sinobin != 3)
    return fail(XT_E_ARG, "sinobin=%d: the engine reads projection/weight .bin files (1) or "
                "counts it subsamples itself (3); phantom simulation (2) is not built",
                (int)T->sinobin);
  if (T->readSino4mHDF)
    return fail(XT_E_ARG, "HDF5 input is read by the ingest module, not by the engine");
  if (T->sinobin == 3) {
    // F/XT_Init.c:608-616: the detector is binned by mult_xy x mult_z, a voxel is
    // one binned pixel (the prior keeps its mult_xy scaling), slices are one row
    if ((int)O->mult_xy < 1 || (int)O->mult_z < 1)
      return fail(XT_E_ARG, "sinobin 3 needs integer delta_xy, delta_z >= 1");
    S->total_t_slices = (int32_t)(S->total_t_slices / O->mult_z);
    S->N_r = (int32_t)(S->N_r / O->mult_xy);
    O->delta_xy = 1;
    O->delta_z = 1;
    T->RotCenter /= O->mult_xy;
    S->z_overlap_num = 1;
  } else {
    O->delta_xy = O->mult_xy;  // F/XT_Init.c:617-622
    O->delta_z = O->mult_z;
    S->z_overlap_num = (int32_t)O->mult_z;
  }
  return XT_OK;
}

// view_info.txt: tokens of alphanumerics and '.', a token ended by '-' is a
// time, one ended by ',' an angle in degrees (F/XT_Init.c:56-75,130-162).
int parse_view_info(const char *dir, Sinogram *S) {
  std::string p = join(dir, "view_info.txt");
  FILE *fp = fopen(p.c_str(), "r");
  if (!fp) return fail(XT_E_IO, "cannot open %s", p.c_str());
  std::string tok;
  Real_t time = 0;
  int index = 0, c;
  for (;;) {
    tok.clear();
    int last = EOF;
    while ((c = fgetc(fp)) != EOF) {
      if (isspace(c)) continue;
      if (!isalnum(c) && c != '.') {
        last = c;
        break;
      }
      tok.push_back((char)c);
    }
    if (last == EOF) break;
    if (last == '-') {
      time = (Real_t)atof(tok.c_str());
    } else if (last == ',') {
      if (!tok.empty()) {
        if (index < S->N_p) {
          S->ViewPtr[index] = (Real_t)((Real_t)atof(tok.c_str()) * M_PI / 180.0);
          S->TimePtr[index] = time;
        }
        index++;
      }
    } else {
      fclose(fp);
      return fail(XT_E_ARG, "%s is not correctly formatted", p.c_str());
    }
  }
  fclose(fp);
  if (index != S->N_p)
    return fail(XT_E_ARG, "number of projections read is %d and expected is %d", index, S->N_p);
  return XT_OK;
}

// F/XT_Init.c:98-125: a view belongs to the frame whose interval holds its time.
void bin_views(Sinogram *S, ScannedObject *O) {
  O->ProjIdxPtr = (int32_t **)calloc(O->N_time, sizeof(int32_t *));
  O->ProjNum = (int32_t *)calloc(O->N_time, sizeof(int32_t));
  std::vector<int32_t> idx;
  int obj = 0;
  Real_t t2 = O->Rtime0 + O->delta_Rtime;
  auto flush = [&](int f) {
    O->ProjIdxPtr[f] = (int32_t *)calloc(idx.size() ? idx.size() : 1, sizeof(int32_t));
    if (!idx.empty()) memcpy(O->ProjIdxPtr[f], idx.data(), idx.size() * sizeof(int32_t));
    O->ProjNum[f] = (int32_t)idx.size();
    idx.clear();
  };
  for (int i = 0; i < S->N_p; i++) {
    while (S->TimePtr[i] >= t2 && obj + 1 < O->N_time) {
      flush(obj);
      t2 += O->delta_Rtime;
      obj++;
    }
    idx.push_back(i);
  }
  flush(obj);
  for (int f = obj + 1; f < O->N_time; f++) flush(f);
}

// prior weights proportional to 1/distance, F/XT_Init.c:199-256
void init_filter(ScannedObject *O, TomoInputs *T) {
  auto dist = [](int i, int j, int k, int l) {
    return (Real_t)sqrt(pow((Real_t)i - 1.0, 2.0) + pow((Real_t)j - 1.0, 2.0) +
                       pow((Real_t)k - 1.0, 2.0) + pow((Real_t)l - 1.0, 2.0));
  };
  Real_t prior_const = O->mult_xy * O->mult_xy * O->mult_xy * O->delta_Rtime;
  Real_t sum = 0;
  T->Time_Filter[0] = (Real_t)(1.0 / dist(0, 1, 1, 1));
  sum += (Real_t)(2.0 * T->Time_Filter[0]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        if (i != 1 || j != 1 || k != 1) {
          Real_t t = (Real_t)(1.0 / dist(1, i, j, k));
          T->Spatial_Filter[i][j][k] = t;
          sum = sum + t;
        } else {
          T->Spatial_Filter[i][j][k] = 0;
        }
      }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        T->Spatial_Filter[i][j][k] = prior_const * T->Spatial_Filter[i][j][k] / sum;
  T->Time_Filter[0] = prior_const * T->Time_Filter[0] / sum;
}

}  // namespace

extern "C" int xt_engine_setup(const char *dir, int argc, char **argv, Sinogram *S,
                               ScannedObject *O, TomoInputs *T) {
  if (!S || !O || !T) return fail(XT_E_ARG, "null struct");
  memset(S, 0, sizeof *S);
  memset(O, 0, sizeof *O);
  memset(T, 0, sizeof *T);
  T->node_num = 1;
  T->node_rank = 0;
  int rc = parse_args(argc, argv, S, O, T);
  if (rc) return rc;
  if (O->N_time < 1) return fail(XT_E_ARG, "Rtime_num must be >= 1");
  if (T->initICD != 0) {
    rc = read_reals(dir, "variance_estimate_n0.bin", &T->var_est, 1);
    if (rc) return rc;
  }
  // initStructures, F/XT_Init.c:293-448 (single rank)
  S->N_t = S->total_t_slices;
  if ((int)O->delta_z <= 0 || (int)O->delta_xy <= 0 || S->N_t % (int)O->delta_z != 0 ||
      S->N_r % (int)O->delta_xy != 0)
    return fail(XT_E_ARG, "delta_xy/delta_z do not divide N_r/N_t");
  S->delta_r = S->Length_R / (S->N_r);
  S->delta_t = S->Length_T / (S->N_t);
  {
    // the detector Laplacian of phase-contrast footprints, F/XT_Init.c:284,439-441
    const Real_t lap[3][3] = {{-0.5f, -1.f, -0.5f}, {-1.f, 6.f, -1.f}, {-0.5f, -1.f, -0.5f}};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) S->Lap_Kernel[i][j] = lap[i][j] / (S->delta_r * S->delta_r);
  }
  S->R0 = -T->RotCenter * S->delta_r;
  S->RMax = (S->N_r - T->RotCenter) * S->delta_r;
  S->T0 = (Real_t)(-S->Length_T / 2.0);
  S->TMax = (Real_t)(S->Length_T / 2.0);
  O->Length_X = S->Length_R;
  O->Length_Y = S->Length_R;
  O->Length_Z = S->Length_T;
  O->N_x = (int32_t)(S->N_r / O->delta_xy);
  O->N_y = (int32_t)(S->N_r / O->delta_xy);
  O->N_z = (int32_t)(S->N_t / O->delta_z);
  O->delta_xy = O->delta_xy * S->delta_r;
  O->delta_z = O->delta_z * S->delta_t;
  S->Projection = tree<Real_t ***>(3, {S->N_p, S->N_r, S->N_t});
  S->ProjSelect = (bool ***)xt_multialloc(sizeof(bool), 3,
                                          std::vector<int64_t>{S->N_p, S->N_r, S->N_t}.data());
  S->ProjOffset = tree<Real_t **>(2, {S->N_r, S->N_t});
  T->Weight = tree<Real_t ***>(3, {S->N_p, S->N_r, S->N_t});
  if (!S->Projection || !S->ProjSelect || !S->ProjOffset || !T->Weight)
    return fail(XT_E_ARG, "out of host memory");
  memset(&S->ProjSelect[0][0][0], 1, (size_t)S->N_p * S->N_r * S->N_t);
  if (T->updateProjOffset == 1 || T->updateProjOffset == 3) {
    if (T->sinobin == 3) {
      // the previous stage's offsets on its coarser detector, each value repeated
      // over the pixels it covers (F/XT_Init.c:340-361)
      const int pxy = (T->initICD == 2 || T->initICD == 3) ? 2 : 1, pz = T->initICD == 3 ? 2 : 1;
      const int nr = S->N_r / pxy, nt = S->N_t / pz;
      std::vector<Real_t> c((size_t)nr * nt);
      rc = read_reals(dir, "proj_offset_n0.bin", c.data(), c.size());
      if (rc) return rc;
      for (int i = 0; i < nr; i++)
        for (int j = 0; j < nt; j++)
          for (int p = 0; p < pxy; p++)
            for (int q = 0; q < pz; q++)
              S->ProjOffset[i * pxy + p][j * pz + q] = c[(size_t)i * nt + j];
    } else {
      rc = read_reals(dir, "proj_offset_n0.bin", &S->ProjOffset[0][0], (size_t)S->N_r * S->N_t);
      if (rc) return rc;
    }
  }
  O->x0 = S->R0;
  O->z0 = S->T0;
  O->y0 = (Real_t)(-O->Length_Y / 2.0);
  O->BeamWidth = S->delta_r;
  O->Object = (Real_t ****)calloc(O->N_time, sizeof(Real_t ***));
  for (int i = 0; i < O->N_time; i++) O->Object[i] = tree<Real_t ***>(3, {O->N_z + 2, O->N_y, O->N_x});
  S->OffsetR = (Real_t)((O->delta_xy / sqrt(2.0) + S->delta_r / 2.0) / 64);
  S->OffsetT = (Real_t)(((O->delta_z / 2) + S->delta_t / 2) / 64);
  T->Sigma_S_Q = (Real_t)pow((O->Sigma_S * O->mult_xy), 2.0);
  T->Sigma_S_Q_P = (Real_t)pow(O->Sigma_S * O->mult_xy, 2.0 - O->MRF_P);
  T->Sigma_T_Q = (Real_t)pow((O->Sigma_T * O->delta_Rtime), 2.0);
  T->Sigma_T_Q_P = (Real_t)pow(O->Sigma_T * O->delta_Rtime, 2.0 - O->MRF_P);
  T->num_z_blocks = (int32_t)(2 * floor(T->num_threads / O->N_time));
  if (T->num_z_blocks < 2)
    T->num_z_blocks = 2;
  else if (T->num_z_blocks > O->N_z)
    T->num_z_blocks = O->N_z;
  const int nb = T->num_z_blocks, nxy = O->N_y * O->N_x;
  T->x_rand_select = tree<int32_t ***>(3, {O->N_time, nb, nxy});
  T->y_rand_select = tree<int32_t ***>(3, {O->N_time, nb, nxy});
  T->x_NHICD_select = tree<int32_t ***>(3, {O->N_time, nb, nxy});
  T->y_NHICD_select = tree<int32_t ***>(3, {O->N_time, nb, nxy});
  T->UpdateSelectNum = tree<int32_t **>(2, {O->N_time, nb});
  T->NHICDSelectNum = tree<int32_t **>(2, {O->N_time, nb});
  O->NHICD_Iterations = 10;
  init_filter(O, T);
  S->ViewPtr = (Real_t *)calloc(S->N_p > 0 ? S->N_p : 1, sizeof(Real_t));
  S->TimePtr = (Real_t *)calloc(S->N_p > 0 ? S->N_p : 1, sizeof(Real_t));
  rc = parse_view_info(dir, S);
  if (rc) return rc;
  bin_views(S, O);
  S->cosine = (Real_t *)calloc(S->N_p, sizeof(Real_t));
  S->sine = (Real_t *)calloc(S->N_p, sizeof(Real_t));
  for (int i = 0; i < S->N_p; i++) {
    S->cosine[i] = (Real_t)cos(S->ViewPtr[i]);
    S->sine[i] = (Real_t)sin(S->ViewPtr[i]);
  }
  return XT_OK;
}

// genSinogram_subsampleCounts, F/XT_genSinogram.c:303-370: bright-field counts
// [N_r * mult_xy][N_t * mult_z] and counts [N_p][N_r * mult_xy][N_t * mult_z]
// summed over each binned pixel's mult_xy x mult_z detector pixels (in Real_t, the
// reference's order); projection = BH t^2 + t with t = log(bright / weight) (the
// log and the quadratic in double, rounded to Real_t), weight = the summed counts.
static int read_subsampled_counts(const char *dir, Sinogram *S, ScannedObject *O, TomoInputs *T) {
  const int mxy = (int)O->mult_xy, mz = (int)O->mult_z;
  const int NR = S->N_r * mxy, NT = S->N_t * mz;
  std::vector<Real_t> bright((size_t)NR * NT), weight((size_t)S->N_p * NR * NT);
  int rc = read_reals(dir, "bright_n0.bin", bright.data(), bright.size());
  if (rc) return rc;
  rc = read_reals(dir, "weight_n0.bin", weight.data(), weight.size());
  if (rc) return rc;
  const double bh = xt_get_option(XT_OPT_BH_QUAD_COEF);
  for (int i = 0; i < S->N_p; i++)
    for (int k = 0; k < S->N_r; k++)
      for (int slice = 0; slice < S->N_t; slice++) {
        Real_t weight_temp = 0, bright_temp = 0;
        for (int p = 0; p < mxy; p++)
          for (int q = 0; q < mz; q++) {
            weight_temp += weight[((size_t)i * NR + k * mxy + p) * NT + slice * mz + q];
            bright_temp += bright[(size_t)(k * mxy + p) * NT + slice * mz + q];
          }
        const Real_t temp = (Real_t)log((double)(bright_temp / weight_temp));
        S->Projection[i][k][slice] = (Real_t)(bh * temp * temp + temp);
        T->Weight[i][k][slice] = weight_temp;
      }
  return XT_OK;
}

extern "C" int xt_engine_read_sinogram(const char *dir, Sinogram *S, ScannedObject *O,
                                       TomoInputs *T) {
  if (!S || !O || !T) return fail(XT_E_ARG, "null struct");
  size_t n = (size_t)S->N_p * S->N_r * S->N_t;
  int rc;
  if (T->sinobin == 3) {
    rc = read_subsampled_counts(dir, S, O, T);
  } else {
    rc = read_reals(dir, "projection_n0.bin", &S->Projection[0][0][0], n);
    if (rc == XT_OK) rc = read_reals(dir, "weight_n0.bin", &T->Weight[0][0][0], n);
  }
  if (rc) return rc;
  Real_t *w = &T->Weight[0][0][0];
  for (size_t i = 0; i < n; i++) w[i] /= T->var_est;  // F/XT_Engine.c:120-123
  return XT_OK;
}

extern "C" int xt_engine_write_outputs(const char *dir, Sinogram *S, ScannedObject *O,
                                       TomoInputs *T) {
  char name[128];
  for (int i = 0; i < O->N_time; i++) {
    snprintf(name, sizeof name, "object_n%d_time_%d.bin", T->node_rank, i);
    int rc = write_reals(dir, name, &O->Object[i][1][0][0], (size_t)O->N_z * O->N_y * O->N_x);
    if (rc) return rc;
  }
  snprintf(name, sizeof name, "proj_offset_n%d.bin", T->node_rank);
  return write_reals(dir, name, &S->ProjOffset[0][0], (size_t)S->N_r * S->N_t);
}

extern "C" void xt_engine_free(Sinogram *S, ScannedObject *O, TomoInputs *T) {
  if (O->ProjIdxPtr) {
    for (int i = 0; i < O->N_time; i++) free(O->ProjIdxPtr[i]);
    free(O->ProjIdxPtr);
  }
  free(O->ProjNum);
  if (O->Object) {
    for (int i = 0; i < O->N_time; i++) xt_multifree(O->Object[i], 3);
    free(O->Object);
  }
  xt_multifree(T->x_rand_select, 3);
  xt_multifree(T->y_rand_select, 3);
  xt_multifree(T->x_NHICD_select, 3);
  xt_multifree(T->y_NHICD_select, 3);
  xt_multifree(T->UpdateSelectNum, 2);
  xt_multifree(T->NHICDSelectNum, 2);
  xt_multifree(S->ProjOffset, 2);
  xt_multifree(S->ProjSelect, 3);
  xt_multifree(T->Weight, 3);
  if (S->Projection) xt_multifree(S->Projection, 3);
  free(S->ViewPtr);
  free(S->TimePtr);
  free(S->cosine);
  free(S->sine);
  memset(S, 0, sizeof *S);
  memset(O, 0, sizeof *O);
  FILE *dbgf = T->debug_file_ptr;
  memset(T, 0, sizeof *T);
  T->debug_file_ptr = dbgf;
}

extern "C" int xt_engine_run(const char *dir, int argc, char **argv) {
  Sinogram S;
  ScannedObject O;
  TomoInputs T;
  int rc = xt_engine_setup(dir, argc, argv, &S, &O, &T);
  std::string logp = join(dir, "DEBUG_n0_xt_hip.log");
  FILE *log = fopen(logp.c_str(), "w");
  T.debug_file_ptr = log;
  if (rc == XT_OK) rc = xt_engine_read_sinogram(dir, &S, &O, &T);
  if (rc == XT_OK && T.reconstruct == 1) {
    std::string saved = g_xt_run_dir;
    g_xt_run_dir = (dir && dir[0]) ? dir : "";
    rc = ICD_BackProject(&S, &O, &T);
    g_xt_run_dir = saved;
    if (rc != XT_OK) fprintf(stderr, "xt_engine: ICD_BackProject: %s\n", xt_last_error());
    if (rc == XT_OK && T.WritePerIter == 0) rc = xt_engine_write_outputs(dir, &S, &O, &T);
  }
  xt_engine_free(&S, &O, &T);
  if (log) fclose(log);
  return rc;
}
