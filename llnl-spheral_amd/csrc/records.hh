// Per-node records in the neighbour build's sorted order.
//
// Every pass gathers the state of its neighbours; packing what one side of a
// pair needs into one record keeps each gather to whole 128-byte lines.  The
// isotropic pair and density kernels gather smaller records made from these
// (pack.hh).
//
//  isotropic-H record (every H = h I, no Q multipliers): 16 doubles, one line
//    [0..2] x  [3] 1/h=H.xx  [4] m  [5] det H  [6] caller index  [7] rho
//    [8..10] v [11] c  [12] P/rho^2*omega^-1  [13] m/rho  [14] P  [15] safeInv(omega)
//  general record: 24 doubles
//    [0..2] x  [3] det H  [4..9] H (xx,xy,xz,yy,yz,zz)  [10] m  [11] caller index
//    [12] rho  [13..15] v  [16] c  [17] Prho  [18] m/rho  [19] P  [20] safeInv(omega)
//    [21] Cl multiplier  [22] Cq multiplier  [23] shear correction
// Unused vector components (nu < 3) are zero.
#pragma once

namespace sphgpu {

template <bool ISO> struct RL;
template <> struct RL<true> {
  static constexpr int STRIDE = 16;
  static constexpr int X = 0, HINV = 3, M = 4, HDET = 5, ORIG = 6, RHO = 7, V = 8, CS = 11, PRHO = 12, MOR = 13,
                       P = 14, SOM = 15;
};
template <> struct RL<false> {
  static constexpr int STRIDE = 24;
  static constexpr int X = 0, HDET = 3, H = 4, M = 10, ORIG = 11, RHO = 12, V = 13, CS = 16, PRHO = 17, MOR = 18,
                       P = 19, SOM = 20, FCL = 21, FCQ = 22, FSH = 23;
};

// The FP64 density sum's neighbour record (isotropic H), 48 B: x, 1/h, m and the
// caller index; three 16-B loads per neighbour, and the gathered array (115 MB
// at 2.4 M slots) fits the Infinity Cache.  det H = (1/h)^D is recomputed.
struct DensRec {
  static constexpr int STRIDE = 6;
  static constexpr int X = 0, HINV = 3, M = 4, ORIG = 5;
};

}  // namespace sphgpu
