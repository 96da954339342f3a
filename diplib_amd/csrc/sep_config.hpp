// Tile constants shared by the launch tables.
#pragma once

namespace dipamd {
constexpr int kLinePF = 8;  // outputs per lane, line kernel (float)
constexpr int kLinePD = 4;  // outputs per lane, line kernel (double)
constexpr int kSweepP = 8;  // outputs per thread and step, sweep kernel

}  // namespace dipamd
