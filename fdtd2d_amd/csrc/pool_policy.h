// pool_policy.h - host-side scheduling arithmetic of pools (time-slicing
// turns, choice of the list copy a rebuild writes), free of HIP so that the
// host tests compile it alone.
#ifndef FDTD2D_POOL_POLICY_H
#define FDTD2D_POOL_POLICY_H

namespace fdtd2d {
namespace policy {

// First launch at or after g in which a shot that runs every k-th launch
// (launches 0, k, 2k, ...) takes its turn.
inline int first_turn(int g, int k) { return (g + k - 1) / k * k; }

// Launches from g to the next turn of period k (g >= 0; k when g is a turn).
inline int to_next_turn(int g, int k) { return k - g % k; }

// The list copy a rebuild writes, among n copies (valid: holds a shot set;
// last_g: last launch that read it): an unused copy if there is one, else
// the least recently read one, never the current copy while n > 1.
inline int pick_copy(const bool *valid, const int *last_g, int n, int cur) {
  int nb = -1;
  for (int i = 0; i < n; ++i) {
    if (i == cur && n > 1) continue;
    if (nb < 0 || (!valid[i] && valid[nb]) || (valid[i] == valid[nb] && last_g[i] < last_g[nb])) nb = i;
  }
  return nb;
}

}  // namespace policy
}  // namespace fdtd2d

#endif
