// Encoding shared by the host symbolic layer and the device kernels.
#pragma once

namespace g2g {

constexpr int kLanes = 64;          // rows per wavefront (one lane per batch row)
constexpr int kMisWaves = 4;        // wavefronts per tile in the mismatch kernel
constexpr int kMaxC = 16;           // longest off-diagonal L column the kernels unroll for

// program slot references
constexpr int kSlotMask = 0xFFFF;   // slot index; kNoSlot = none
constexpr int kNoSlot = 0xFFFF;
constexpr int kFirst = 1 << 16;     // first update of the target: store, not accumulate
constexpr int kInA = 1 << 17;       // the value starts from the next A-stream element

// backward-program bus update ops
constexpr int kOpThetaSelf = 1;       // this x is dVa; the bus has no |V| unknown
constexpr int kOpThetaSelfVmSlot = 2; // this x is dVa; dVm from the partner slot
constexpr int kOpVmSelfThetaSlot = 3; // this x is dVm; dVa from the partner slot

// row status (gpusim2grid contingency/row_status.hpp:33-43) plus "still iterating"
constexpr int kActive = -1;
constexpr int kConverged = 0;
constexpr int kMaxIter = 1;
constexpr int kDiverged = 2;
constexpr int kNotSimulated = 3;

}  // namespace g2g
