// FP64 instances of the launchers (the FP64 oracle runs of the GPU path).
#include "ops_impl.hpp"
template struct scl::Ops<double>;
