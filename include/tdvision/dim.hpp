// Reference header LibThunderVision/src/tdvision/dim.hpp: its declarations (Dim)
// live in common.hpp.
#ifndef TDV_FWD_DIM_HPP
#define TDV_FWD_DIM_HPP
#include "common.hpp"
#endif
