// Reference header LibThunderVision/src/tdvision/ctrlprocess.hpp: its declarations (CtrlWork, CtrlProcess)
// live in matchers.hpp.
#ifndef TDV_FWD_CTRLPROCESS_HPP
#define TDV_FWD_CTRLPROCESS_HPP
#include "matchers.hpp"
#endif
