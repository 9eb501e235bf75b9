// Reference header LibThunderVision/src/tdvision/dynamicprogdev.hpp: its declarations (DynamicProgDev)
// live in stages.hpp.
#ifndef TDV_FWD_DYNAMICPROGDEV_HPP
#define TDV_FWD_DYNAMICPROGDEV_HPP
#include "stages.hpp"
#endif
