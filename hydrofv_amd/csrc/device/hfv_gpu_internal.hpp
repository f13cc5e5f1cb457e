// hfv_gpu_internal.hpp -- C ABI declarations shared by the device sources.
#pragma once

#include <new>

#include "hydrofv/hfv_gpu.h"
