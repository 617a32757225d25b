// batch.cpp - batched persistent launches and hipGraph capture (see vkr.h).
#include "kernels.hpp"
