// Reference header LibThunderVision/src/tdvision/medianfilterdev.hpp: its declarations (MedianFilterDev)
// live in matchers.hpp.
#ifndef TDV_FWD_MEDIANFILTERDEV_HPP
#define TDV_FWD_MEDIANFILTERDEV_HPP
#include "matchers.hpp"
#endif
