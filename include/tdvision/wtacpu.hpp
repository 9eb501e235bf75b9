// Reference header LibThunderVision/src/tdvision/wtacpu.hpp: its declarations (WTACPU (device-backed))
// live in stages.hpp.
#ifndef TDV_FWD_WTACPU_HPP
#define TDV_FWD_WTACPU_HPP
#include "stages.hpp"
#endif
