// Reference header LibThunderVision/src/tdvision/benchmark.hpp: its declarations (Benchmark, CudaBenchmarker)
// live in images.hpp.
#ifndef TDV_FWD_BENCHMARK_HPP
#define TDV_FWD_BENCHMARK_HPP
#include "images.hpp"
#endif
