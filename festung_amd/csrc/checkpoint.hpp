// Binomial checkpointing of time steps for the discrete adjoint: the
// schedule that reverses n time steps with c state snapshots (the initial
// state included) in the least number of forward steps (Griewank's binomial
// bound; the "revolve" schedule of Griewank & Walther, ACM TOMS 26 (2000)).
#pragma once
#include <cstdint>
#include <vector>

namespace festung {

// One action of a schedule.  Slot ids 0 .. c-1 are snapshot buffers (slot 0
// holds the initial state); kRot is a pair of scratch state buffers.
struct CkAction {
  enum Kind { ADVANCE, REVERSE } kind;
  int32_t from;   // step index of the source state
  int32_t steps;  // ADVANCE: steps to take (>= 1)
  int32_t src;    // buffer holding state `from`
  int32_t dst;    // ADVANCE: buffer receiving state from + steps
};
constexpr int32_t kRot = -1;

// Forward steps (not counting the one recomputation with stored stage states
// that precedes each step's reversal) needed to reverse n steps with s free
// snapshots besides the one holding the start state:
//   s = 0: n (n - 1) / 2;  s >= 1: r n - C(s + 1 + r, s + 2), r the least
//   integer with C(s + 1 + r, s + 1) >= n.
int64_t binomial_advances(int64_t n, int64_t s);

// Actions reversing steps 0 .. n-1 (REVERSE actions for j = n-1 down to 0,
// each preceded by the advances that bring state j into a buffer), with c
// snapshot slots.  The schedule is simulated on the host before it is
// returned: every action reads the state it claims.
std::vector<CkAction> binomial_schedule(int32_t n, int32_t c);

}  // namespace festung
