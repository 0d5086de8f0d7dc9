// Shapes of the full DP by the width of B, and the forward kernel variants
// of the DP's rounds.  Written by mauvehip/dp_autotune.py from measurements
// on one MI355X: every candidate's DP milliseconds per workload (forward +
// trace-back); chosen: shapes=rows256_s8_s16 wide=0 narrow=1 (a
// change must gain more than 1% over the best before it)
//   shapes=rows256_s8_s16 wide=0 narrow=1: whole 2.424 ms (forward 1.818), batch 2188.458 ms (forward 2105.835)
//   shapes=rows128_s4_s8_s16 wide=0 narrow=1: whole 2.989 ms (forward 2.386), batch 2181.672 ms (forward 2099.090)
//   shapes=rows128_s8_s16 wide=0 narrow=1: whole 3.362 ms (forward 2.738), batch 2184.885 ms (forward 2102.285)
//   shapes=rows256_s8to1024_s16 wide=0 narrow=1: whole 2.425 ms (forward 1.819), batch 2186.431 ms (forward 2103.787)
//   shapes=rows256_s16 wide=0 narrow=1: whole 2.429 ms (forward 1.822), batch 2186.768 ms (forward 2104.195)
//   shapes=rows256_s8to1024_s16 wide=2 narrow=1: whole 2.474 ms (forward 1.864), batch 2247.095 ms (forward 2162.635)
//   shapes=rows256_s8to1024_s16 wide=3 narrow=1: whole 2.521 ms (forward 1.926), batch 2112.089 ms (forward 2029.465)
//   shapes=rows256_s8to1024_s16 wide=4 narrow=1: whole 2.560 ms (forward 1.950), batch 2247.936 ms (forward 2165.282)
//   shapes=rows256_s8to1024_s16 wide=0 narrow=5: whole 2.418 ms (forward 1.819), batch 2185.095 ms (forward 2102.480)
//   shapes=rows256_s8to1024_s16 wide=0 narrow=0: whole 2.430 ms (forward 1.819), batch 2185.338 ms (forward 2102.703)
#pragma once
#include <cstdint>

namespace mh {
// upper bounds (inclusive) of the classes of lb; the last class is open
constexpr uint32_t DP_TUNE_LB_CLASS[] = {32, 64, 96, 128, 160, 192, 224, 256, 384, 512, 1024};
constexpr int DP_TUNE_NCLASS = (int)(sizeof(DP_TUNE_LB_CLASS) / sizeof(DP_TUNE_LB_CLASS[0])) + 1;
// per class: 0 = four jobs a wavefront on 16-lane rows (the fewest columns
// that hold the job; only for lb <= 256 and a + b < 1024, else stripes of 8),
// 4, 8 or 16 = one job a wavefront in stripes of 64 lanes x that many columns
constexpr uint8_t DP_TUNE_SHAPE[DP_TUNE_NCLASS] = {0, 0, 0, 0, 0, 0, 0, 0, 8, 8, 16, 16};
// forward kernel variants (indices into the variant table of dp.hip) of
// rounds with a shape wider than 8 columns a lane, and of the other rounds
constexpr int DP_TUNE_WIDE_VARIANT = 0;
constexpr int DP_TUNE_NARROW_VARIANT = 1;
}  // namespace mh
