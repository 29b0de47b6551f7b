/* link_check.c -- TEST INFRASTRUCTURE: link check of the drop-in boundary.
 *
 * Compiled with the reference's own headers (-I Source_Code_4D_Fast) into
 * oracle/_ref/libxt_host_link.so together with the reference host sources that
 * compile here and stay on the engine's link line when libxt_hip.so replaces
 * the ICD path (XT_genSinogram.c, XT_IOMisc.c, XT_ImageProc.c, XT_Filter.c,
 * XT_HDFIO.c, allocate.c, randlib.c, tiff.c, XT_AMatrix.c, XT_Profile.c), and
 * linked against libxt_hip.so with -Wl,--no-undefined: every symbol those
 * objects need must resolve.  The table below holds the functions of those
 * objects and of the library that XT_Engine.c calls (ICD_BackProject at
 * Source_Code_4D_Fast/XT_Engine.c:159, the sinogram readers at :99,:103 and the
 * result writer at :162), so the boundary symbol itself is resolved from
 * libxt_hip.so as well.  XT_Engine.c, XT_Init.c and XT_MPI.c need <mpi.h> or
 * <hdf5.h>, which this image lacks, and are not compiled.  Nothing here runs.
 */
#include <stdio.h>

#include "XT_Structures.h"
#include "XT_ICD_update.h"
#include "XT_genSinogram.h"
#include "XT_IOMisc.h"

typedef int (*xt_link_engine_fn)(Sinogram *, ScannedObject *, TomoInputs *);
typedef void (*xt_link_writer_fn)(Sinogram *, ScannedObject *, TomoInputs *);

const xt_link_engine_fn xt_link_check_entries[] = {ICD_BackProject, genSinogram_fromBin,
                                                   genSinogram_subsampleCounts};
const xt_link_writer_fn xt_link_check_writer = write_ObjectProjOff2TiffBinPerIter;
