// Reference header LibThunderVision/src/tdvision/dynamicprogcpu.hpp: its declarations (DynamicProgCPU (device-backed))
// live in stages.hpp.
#ifndef TDV_FWD_DYNAMICPROGCPU_HPP
#define TDV_FWD_DYNAMICPROGCPU_HPP
#include "stages.hpp"
#endif
