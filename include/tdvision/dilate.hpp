// Reference header LibThunderVision/src/tdvision/dilate.hpp: its declarations (Dilate)
// live in matchers.hpp.
#ifndef TDV_FWD_DILATE_HPP
#define TDV_FWD_DILATE_HPP
#include "matchers.hpp"
#endif
