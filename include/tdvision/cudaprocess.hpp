// Reference header LibThunderVision/src/tdvision/cudaprocess.hpp: its declarations (CUDAProcess)
// live in process.hpp.
#ifndef TDV_FWD_CUDAPROCESS_HPP
#define TDV_FWD_CUDAPROCESS_HPP
#include "process.hpp"
#endif
