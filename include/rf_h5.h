/*
 * rf_h5.h -- HDF5 / XDMF output of structured random fields, in the layout
 * the reference's post-processing (statistics, SEM) reads.  Host code only:
 * the fields are host arrays (copy a device field back first).  Built as
 * librf_h5.so against the HDF5 C library (1.10).
 *
 * Layout shared with the reference:
 *   <folder>/h5/<name>.h5   dataset "samples", IEEE F64, rank nDim, HDF5
 *                           dims (n_d, ..., n_1) -- the Fortran (n_1, ..., n_d)
 *                           with x fastest; realization 1 only
 *   <folder>/xmf/<name>.xmf XDMF 2.0 CoRectMesh over ../h5/<name>.h5:/samples
 *   file attributes          write_HDF5_attributes' names and types
 */
#ifndef RF_H5_H
#define RF_H5_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RF_H5_OK      0
#define RF_H5_E_ARG  (-1)   /* bad argument                         */
#define RF_H5_E_IO   (-7)   /* HDF5 or file-system failure           */

/* write_HDF5_attributes' argument list (writeResultFile_RF.f90:1758-1833);
 * every vector holds nDim entries. */
typedef struct rf_h5_attrs {
    int32_t nb_procs, nDim, Nmc, method, seedStart, corrMod, margiFirst, localizationLevel;
    int32_t kNStep[3];
    int32_t nFields[3];
    double  xMinGlob[3], xMaxGlob[3], xStep[3], corrL[3], overlap[3], procExtent[3], kMax_out[3];
} rf_h5_attrs;

/* write_Simple_pHDF5_Str (writeResultFile_RF.f90:1062-1247) on one rank:
 * creates <folder>/h5/<name>.h5 with the dataset "samples" of the global grid
 * xNStepGlob, writes the window minPos..maxPos (1-based, inclusive) of the
 * field (xNStep points, x fastest) at the global offset origin - 1, then the
 * XDMF file <folder>/xmf/<name>.xmf (nDim 2 and 3; the reference writes none
 * in 1-D).  Missing h5/ and xmf/ folders are created.  h5Path and xmfPath
 * (pathCap bytes each, may be NULL) receive the two paths. */
int rf_h5_write_simple(const char* folder, const char* name, int32_t nDim,
                       const int32_t* minPos, const int32_t* maxPos, const int32_t* origin,
                       const int32_t* xNStep, const double* xStep, const double* xMinGlob,
                       const int32_t* xNStepGlob, const double* field,
                       char* h5Path, char* xmfPath, int32_t pathCap);

/* write_HDF5_attributes (writeResultFile_RF.f90:1758-1833): opens h5Path
 * read-write and attaches the attributes to the file root: scalars as
 * H5T_STD_I32LE (SCALAR), integer vectors H5T_STD_I32LE, real vectors
 * H5T_IEEE_F64LE (write_h5attr_int / _int_vec / _real_vec, hdf5_RF.f90:42-177). */
int rf_h5_write_attributes(const char* h5Path, const rf_h5_attrs* a);

/* write_XMF_Elements (writeResultFile_RF.f90:835-938): the XDMF text alone. */
int rf_h5_write_xmf(const char* xmfPath, const char* h5FileName, const char* h5ToXmfPath,
                    int32_t nDim, const double* xMin, const double* xMax, const double* xStep,
                    const char* dsetName);

/* The read side of interpolateToMesh (sample_RF.f90:906-975): the file
 * attributes nDim, Nmc, xMinGlob, xMaxGlob, xStep, the global grid
 * xNStep = nint((xMaxGlob - xMinGlob)/xStep) + 1 and the bounding box of
 * [xMinLoc, xMaxLoc]: minPos = floor((xMinLoc - xMinGlob)/xStep) + 1,
 * maxPos = ceiling((xMaxLoc - xMinGlob)/xStep) + 1, clamped to the grid. */
typedef struct rf_h5_box {
    int32_t nDim, Nmc;
    double  xMinGlob[3], xMaxGlob[3], xStep[3];
    int32_t xNStep[3], minPos[3], maxPos[3];
} rf_h5_box;
int rf_h5_read_box_info(const char* h5Path, int32_t nDim, const double* xMinLoc,
                        const double* xMaxLoc, rf_h5_box* box);
/* The box minPos..maxPos of "samples" into out (prod(maxPos - minPos + 1)
 * doubles, x fastest). */
int rf_h5_read_box(const char* h5Path, const rf_h5_box* box, double* out);

/* Message of the last failure on the calling thread ("" if none). */
const char* rf_h5_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
