// FP64 build of the oracle (test infrastructure; Makefile target liboracle_f64.so): included
// ahead of cpu_ref.cpp, after the record layouts, it makes every float quantity of the restated
// tick a double. The records keep the ABI's types (V_m and the ions are double there; ATP,
// temperature, transmitters and the synaptic scalars are stored as float), so the build gives
// the tick with every operation in double precision and the reference's record storage.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "../include/biocore_abi.h"
typedef float abi_float;
#define ORC_F64 1
#define float double
