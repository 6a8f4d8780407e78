// FP32 instances of the launchers.
#include "ops_impl.hpp"
template struct scl::Ops<float>;
