// Reference header LibThunderVision/src/tdvision/wtadev.hpp: its declarations (WTADev)
// live in stages.hpp.
#ifndef TDV_FWD_WTADEV_HPP
#define TDV_FWD_WTADEV_HPP
#include "stages.hpp"
#endif
