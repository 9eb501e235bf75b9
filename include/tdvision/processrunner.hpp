// Reference header LibThunderVision/src/tdvision/processrunner.hpp: its declarations (ProcessRunner)
// live in process.hpp.
#ifndef TDV_FWD_PROCESSRUNNER_HPP
#define TDV_FWD_PROCESSRUNNER_HPP
#include "process.hpp"
#endif
