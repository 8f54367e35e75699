// Class-kernel registry: one launcher per (derivative, precision, LI, LJ, LK,
// LL) instantiated in the generated translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace cx {

typedef hipError_t (*LaunchFn)(const EriArgs&, hipStream_t);

struct KernelInfo {
  LaunchFn fn;
  int tuples_per_block;  // Q
  int lanes_per_tuple;   // T
};

struct Registry {
  KernelInfo k[2][2][kMaxL + 1][kMaxL + 1][kMaxL + 1][kMaxL + 1];
};

// Filled by the generated instantiation units.
void register_class_kernels(Registry& reg);

}  // namespace cx
