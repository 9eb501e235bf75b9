// Reference header LibThunderVision/src/tdvision/optimizer.hpp: its declarations (Optimizer, AbstractOptimizer)
// live in stages.hpp.
#ifndef TDV_FWD_OPTIMIZER_HPP
#define TDV_FWD_OPTIMIZER_HPP
#include "stages.hpp"
#endif
