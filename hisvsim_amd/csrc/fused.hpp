// Fused-part path of libhsv: one part's gates multiplied into a dense
// 2^k x 2^k FP64 unitary on the host, applied to the state as a batched
// complex GEMM on the gfx950 matrix cores.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "gates.hpp"

namespace hsv {

// Scratch device buffer owned by the caller's handle (grown on demand).
struct DeviceScratch {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Applies the fused unitary of `gates` to the n-qubit state at `amp`
// (dtype HSV_C128 or HSV_C64) on `stream`. mode: hsv_fused_mode.
// `num_cu` sizes the persistent grid. bf_swizzle selects the row order of
// the BF16-split kernel's loads and stores (1: four consecutive rows per
// column in each instruction). Returns 0 or a negative hsv_status.
int fused_apply(void* amp, int n, int dtype, hipStream_t stream, int num_cu, DeviceScratch* scratch,
                const std::vector<Lowered>& gates, int mode, std::string* err, int bf_swizzle = 1);

// The fused unitary itself (for tests and the cost model): qubits of the
// part (sorted, padded with the lowest free qubits to at least min_qubits)
// and the row-major matrix.
int fused_unitary(int n, const std::vector<Lowered>& gates, std::vector<int>* qubits, std::vector<cd>* U,
                  std::string* err, int min_qubits = 4);

}  // namespace hsv
