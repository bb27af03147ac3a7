// Grid-completion ticket of the iteration kernels.
//
// Every workgroup of a launch counts in once; the last one to count in
// closes the iteration (it reads the iteration's counters, decides what
// comes next and advances the control block).  Workgroup b takes one
// returning atomic on the counter of group b % kTicketGroups; the last of a
// group takes one on the top counter; the last of those is the last
// workgroup of the grid (a grid of at most kTicketGroups workgroups counts
// on the top counter alone).  Counters reset themselves: the last of a group
// zeroes its group counter and the last workgroup the top counter, so the
// board is all zero between launches.
//
// Atomics that hit one cache line serialize on MI355X (about 12 ns each),
// so every counter has a 128-byte line of its own.  Measured on a 5376-
// workgroup grid (21 per compute unit): 28 us per launch with 64 adjacent
// counters, 4.5 us with padded ones; 3.1 us for a launch without a ticket.
//
// The board lives inside the kernel's control block (ticket_board member),
// not in an allocation of its own: with the board in a separate 8 KB
// hipMalloc allocation the 64-source K-wide kernel ran 1.6x slower on
// every iteration (same code, same counters; measured, cause not
// identified), inside the control block's allocation it runs at full speed.
#ifndef KITES_DETAIL_TICKET_HPP_
#define KITES_DETAIL_TICKET_HPP_

#include <hip/hip_runtime.h>

namespace kites {
namespace detail {

constexpr unsigned kTicketGroups = 64;
constexpr unsigned kTicketStride = 32;  // unsigned words per counter: one 128-byte line
constexpr unsigned kTicketWords = (kTicketGroups + 1) * kTicketStride;

// The board as a member of a control block; the control block's header
// (the fields before the board) is what the host copies back.
struct alignas(128) ticket_board {
  unsigned w[kTicketWords];
};

// Counts the calling workgroup in (call from one thread); true in the last
// workgroup of the grid.  The caller orders its own counter updates before
// this call (their results have returned).
__device__ __forceinline__ bool take_ticket(unsigned* board) {
  unsigned* const top = board + kTicketGroups * kTicketStride;
  if (gridDim.x <= kTicketGroups) {
    // every group would hold one workgroup: the top counter alone (one
    // round trip instead of two; at most kTicketGroups atomics on its line)
    if (atomicAdd(top, 1u) != gridDim.x - 1) return false;
    *top = 0;
    return true;
  }
  unsigned const ng = kTicketGroups;
  unsigned const g = blockIdx.x % ng;
  unsigned const in_group = (gridDim.x - g + ng - 1) / ng;
  unsigned* const cg = board + g * kTicketStride;
  if (atomicAdd(cg, 1u) != in_group - 1) return false;
  *cg = 0;
  if (atomicAdd(top, 1u) != ng - 1) return false;
  *top = 0;
  return true;
}

// Zeroes the board's counters (grid-stride over `tid`, `nt` threads).
__device__ __forceinline__ void clear_ticket_board(unsigned* board, unsigned tid, unsigned nt) {
  for (unsigned i = tid; i <= kTicketGroups; i += nt) board[i * kTicketStride] = 0;
}

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_TICKET_HPP_
