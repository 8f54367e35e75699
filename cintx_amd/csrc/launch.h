// Class-kernel registry: up to kVariants launchers per (derivative,
// form, LI, LJ, LK, LL), instantiated in the generated translation units;
// variant 0 is the default, the others are autotuning candidates.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace cx {

typedef hipError_t (*LaunchFn)(const EriArgs&, hipStream_t);

constexpr int kVariants = 6;

struct KernelInfo {
  LaunchFn fn;
  int tuples_per_block;  // Q
  int lanes_per_tuple;   // T
  int prims_per_step;    // P (one lane per tuple: row chunks)
};

// forms: 0 FP64, 1 FP32 recurrences, 2 FP64 with Cartesian output
constexpr int kForms = 3;

struct Registry {
  KernelInfo k[2][kForms][kMaxL + 1][kMaxL + 1][kMaxL + 1][kMaxL + 1][kVariants];
  // int2e_sph launch shapes with the fused J/K epilogue, FP64 and FP32 forms
  KernelInfo jk[2][kMaxL + 1][kMaxL + 1][kMaxL + 1][kMaxL + 1][kVariants];
};

// Filled by the generated instantiation units.
void register_class_kernels(Registry& reg);

}  // namespace cx
