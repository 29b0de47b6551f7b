// xt_engine -- command-line drop-in for the reference's ./XT_Engine: run it in a
// run folder (view_info.txt, projection_n0.bin, weight_n0.bin, ...) with the
// same options (Source_Code_4D_Fast/XT_Engine.c:128-172, XT_Init.c:452-629).
#include <cstdio>

#include "../../include/xt_icd.h"

int main(int argc, char **argv) {
  int rc = xt_engine_run(nullptr, argc - 1, argv + 1);
  if (rc != XT_OK) fprintf(stderr, "xt_engine: failed (%d): %s\n", rc, xt_last_error());
  return rc == XT_OK ? 0 : 1;
}
