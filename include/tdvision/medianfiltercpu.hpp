// Reference header LibThunderVision/src/tdvision/medianfiltercpu.hpp: its declarations (MedianFilterCPU (device-backed))
// live in matchers.hpp.
#ifndef TDV_FWD_MEDIANFILTERCPU_HPP
#define TDV_FWD_MEDIANFILTERCPU_HPP
#include "matchers.hpp"
#endif
