// runs.h - the maximal runs of plasma cells in one row of the medium, for
// the tiler (which tiles hold plasma, which are plasma through and through).
// Free of HIP so that the host tests compile it alone.
#ifndef FDTD2D_RUNS_H
#define FDTD2D_RUNS_H

#include <algorithm>

namespace fdtd2d {

// Calls emit(first, end) for each maximal run [first, end) of cells i < n
// with positive(i) != 0, in order. Blocks of 64 cells with no change of
// state are passed over after one counting pass (the plasma of a row is one
// or a few long runs).
template <class Positive, class Emit>
inline void scan_runs(int n, Positive positive, Emit emit) {
  constexpr int kBlock = 64;
  int run = -1;  // first column of the open run
  for (int i = 0; i < n; i += kBlock) {
    const int m = std::min(kBlock, n - i);
    int count = 0;
    for (int k = 0; k < m; ++k) count += positive(i + k) != 0;
    if (count == (run >= 0 ? m : 0)) continue;
    for (int k = 0; k < m; ++k) {
      const bool pos = positive(i + k) != 0;
      if (pos && run < 0) {
        run = i + k;
      } else if (!pos && run >= 0) {
        emit(run, i + k);
        run = -1;
      }
    }
  }
  if (run >= 0) emit(run, n);
}

}  // namespace fdtd2d

#endif
