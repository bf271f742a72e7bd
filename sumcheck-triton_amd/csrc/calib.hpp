// Device throughput calibration (calib.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace sc {

enum CalibOp {
  CALIB_MAD_U64 = 0,
  CALIB_MUL_HI_U32 = 1,
  CALIB_ADD_CO = 2,
  CALIB_FMA_F64 = 3,
  CALIB_HBM_READ = 4,  // bytes/s, streaming read of 2 GiB
  CALIB_HBM_COPY = 5,  // bytes/s read + written, streaming copy of 2 GiB
  CALIB_HBM_MIX21 = 6, // bytes/s read + written, 2 GiB read while 1 GiB is written (a fold's mix)
  CALIB_COUNT = 7
};

// rates[k] = sustained lane operations per second of op k (k < 4) or HBM bytes per second
// (k = 4, 5, 6); n <= CALIB_COUNT entries.
hipError_t calibrate(hipStream_t s, double* rates, int n);

}  // namespace sc
