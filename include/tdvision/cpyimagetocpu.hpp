// Reference header LibThunderVision/src/tdvision/cpyimagetocpu.hpp: its declarations (CpyImageToCPU)
// live in matchers.hpp.
#ifndef TDV_FWD_CPYIMAGETOCPU_HPP
#define TDV_FWD_CPYIMAGETOCPU_HPP
#include "matchers.hpp"
#endif
