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
  // a buffer the library does not own (the no-allocation entry point's
  // static array): never freed or grown; a larger need is an error
  bool fixed = false;
};

// Launch choices of the BF16-split kernels.
struct BfTune {
  int variant = 2;   // parts of 5..7 qubits: 0 / 1 column-block kernel (1: four
                     // consecutive rows per column in each load), 2 tile kernel
  int tile_tpw = 8;  // tiles of 4096 amplitudes per workgroup, tile kernel
  int tile_order = 4;  // tile kernels: m = the highest m slots outside the tile
                       // vary fastest across workgroups, 0 = natural order
  int part_tpw = 0;  // tiles per workgroup, wide-part kernel (8..14 qubits);
                     // 0: 8 where the step fragments stay in LDS (the next
                     // tile is prefetched during the steps), else 1
  int part13 = 0;    // 13-qubit tiles: 0 = 512 threads, fragments streamed
                     // through LDS, two workgroups per CU; 1 = 1024 threads,
                     // fragments resident, next tile prefetched (when the
                     // steps fit); 2 = as 0, each wave reading the fragments
                     // from L2 straight into registers
};

// Applies the fused unitary of `gates` to the n-qubit state at `amp`
// (dtype HSV_C128 or HSV_C64) on `stream`. mode: hsv_fused_mode.
// `num_cu` sizes the persistent grids of the FP64 / FP32-split kernels.
// Returns 0 or a negative hsv_status.
int fused_apply(void* amp, int n, int dtype, hipStream_t stream, int num_cu, DeviceScratch* scratch,
                const std::vector<Lowered>& gates, int mode, std::string* err, const BfTune& bf = BfTune());

// One part on the BF16-split wide-part kernel (a 12- to 14-qubit tile that
// holds slots 0-4) with a qubit layout change: logical qubit q sits in slot
// phys[q] and the pass leaves it in slot phys_after[q]; only qubits of the
// tile may move. tile_qubits: the logical qubits the tile must hold (at
// least those of the gates). live_slots: slots whose bit can be 1 in a
// nonzero amplitude (the others hold |0> on the whole state, so tiles with
// such a bit set are zero and are skipped); ~0 when nothing is known.
// pieces: BF16 pieces per value, 2 (HSV_FUSED_BF16SPLIT) or 3
// (HSV_FUSED_BF16X3, 12- and 13-qubit tiles).
int fused_part_bf_layout(void* amp, int n, int dtype, hipStream_t stream, DeviceScratch* scratch,
                         const std::vector<Lowered>& gates, uint64_t tile_qubits, const int* phys,
                         const int* phys_after, uint64_t live_slots, const BfTune& bf, std::string* err,
                         int pieces = 2);

// The fused unitary itself (for tests and the cost model): qubits of the
// part (sorted, padded with the lowest free qubits to at least min_qubits)
// and the row-major matrix.
int fused_unitary(int n, const std::vector<Lowered>& gates, std::vector<int>* qubits, std::vector<cd>* U,
                  std::string* err, int min_qubits = 4);

}  // namespace hsv
