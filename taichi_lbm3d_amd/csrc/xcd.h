// Block order across the XCDs of a gfx950 GPU.
//
// The hardware deals the workgroups of a launch to its 8 XCDs round robin:
// hardware block h runs on XCD h % 8, and each XCD has its own L2.  Workgroups
// that share cache lines (neighbouring cells of a lattice, neighbouring bricks
// of a convolution) therefore mostly miss in each other's L2 when consecutive
// logical blocks are consecutive hardware blocks.  xcd_block hands each XCD
// runs of C consecutive logical blocks instead: within every group of 8C, XCD
// k takes blocks kC .. kC+C-1, and all XCDs still sweep the range together.
// The last, partial group is split into 8 equal runs, so no XCD gets a longer
// tail.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

// The j-th of XCD k's logical blocks among nb (nb past its last).
__host__ __device__ __forceinline__ uint32_t xcd_block(uint32_t C, uint32_t k, uint32_t j, uint32_t nb) {
  const uint32_t full = nb / (8u * C);  // whole groups
  if (j < full * C) return (j / C) * (8u * C) + k * C + (j % C);
  const uint32_t jt = j - full * C, ct = (nb - full * 8u * C + 7u) / 8u;
  const uint32_t b = full * 8u * C + k * ct + jt;
  return jt < ct && b < nb ? b : nb;
}

// Hardware blocks a launch of nb logical blocks needs under runs of C (C = 0:
// plain order, nb rounded up to a multiple of 8): every XCD gets as many
// hardware blocks as its longest list.
inline int64_t xcd_grid(int64_t nb, int64_t C) {
  if (C <= 0) return (nb + 7) & ~int64_t(7);
  const int64_t per = 8 * C, full = nb / per, tail = nb - full * per;
  return full * per + 8 * ((tail + 7) / 8);
}
