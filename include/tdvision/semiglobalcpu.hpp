// Reference header LibThunderVision/src/tdvision/semiglobalcpu.hpp: its declarations (SemiGlobalCPU (device-backed))
// live in stages.hpp.
#ifndef TDV_FWD_SEMIGLOBALCPU_HPP
#define TDV_FWD_SEMIGLOBALCPU_HPP
#include "stages.hpp"
#endif
