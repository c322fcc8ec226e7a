/*
 * Env.hpp -- macros, enums and the execution flag of the APPFIS API
 * (same names and meaning as the reference Appfis/Env.hpp:16-32).
 *
 * The API has one backend here: MI355X through HIP.  Compile with -DAPPFIS_HIP
 * (hipcc, gfx950).  The reference apps select their APPFIS variant with
 * `#if defined(APPFIS_CUDA) || defined(APPFIS_CPU)`, so they are built with
 * -DAPPFIS_HIP -DAPPFIS_CPU: APPFIS_CPU only picks the *_appfis.hpp sources there,
 * and APPFIS_HIP decides the backend here.  There is no CPU execution backend: a
 * build without APPFIS_HIP stops at compile time instead of running elsewhere.
 */
#ifndef APPFIS_ENV_HPP
#define APPFIS_ENV_HPP

#ifndef APPFIS_HIP
#error "Appfis/ targets MI355X through HIP: compile with hipcc -DAPPFIS_HIP (gfx950)"
#endif

#include <hip/hip_runtime.h>

namespace APPFIS
{
// user stencils run on the device; Array accessors on both sides
#define PARALLEL __device__
#define EXEC_SPACE __host__ __device__
#define STENCIL Stencil

    // true between InitializeExecution and FinalizeExecution: GetSubGrid() then
    // returns device Array views (reference Env.hpp:26, Grid.hpp:296-328)
    inline bool EXECUTION_FLAG = false;

    enum AXIS { X, Y, Z };
    enum DIMENSION { DIM_1D = 1, DIM_2D = 2, DIM_3D = 3 };
    enum DIRECTION { LEFT, RIGHT, UP, DOWN, FRONT, BACK };
    enum GRID_TYPE { SINGLE, DUAL, FLAT };
    enum REDUCTION { MAX, MIN, SUM };
}  // namespace APPFIS

#endif
