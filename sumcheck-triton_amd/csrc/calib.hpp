// Device throughput calibration (calib.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace sc {

enum CalibOp { CALIB_MAD_U64 = 0, CALIB_MUL_HI_U32 = 1, CALIB_ADD_CO = 2, CALIB_FMA_F64 = 3, CALIB_COUNT = 4 };

// rates[k] = sustained lane operations per second of op k (n <= CALIB_COUNT entries).
hipError_t calibrate(hipStream_t s, double* rates, int n);

}  // namespace sc
