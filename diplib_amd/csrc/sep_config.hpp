// Tile constants shared by the launch tables.
#pragma once

namespace dipamd {
constexpr int kLinePF = 8;  // outputs per lane, line kernel (float)
constexpr int kLinePD = 4;  // outputs per lane, line kernel (double)
constexpr int kSweepP = 8;  // outputs per thread and step, sweep kernel
constexpr int kColPF = 8;   // outputs per thread and step, column-tile sweep (float)
constexpr int kColPD = 4;   // outputs per thread and step, column-tile sweep (double)
constexpr int kColsweepMinBorder = 16;  // shorter filters on strided axes use the global sweep
constexpr int kZSweepMaxR = 12;         // the register-window sweep takes symmetric radii 1..12

}  // namespace dipamd
