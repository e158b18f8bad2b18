// Fused HODLR matvec (three launches, no host synchronisation).
#pragma once
#include <cstdint>

#include "common.hpp"

namespace hmat {
namespace kern {

struct MvLeaf {
  int64_t start;
  int32_t len, pad_;
};
struct MvLevel {  // one tree depth c >= 1 (index c - 1)
  const double* V;
  const double* US;
  int64_t ld;
};
struct MvEntry {  // (leaf, depth): the depth-c segment s holding the leaf
  int32_t vcols;  // columns of V at s = rank of block (sib(s), s)
  int32_t ucols;  // columns of (U S) at s = rank of block (s, sib(s))
  int32_t tin;    // global index of sib(s): its t vector multiplies (U S)_s
  int32_t pad_;
};
struct MvSeg {  // one segment (global index = position in this list)
  int32_t level, leaf0, leaf1, vcols;
};

// Device-resident plan; every pointer is device memory.
struct MvPlan {
  const MvLeaf* leaves = nullptr;
  const MvLevel* lvls = nullptr;
  const MvEntry* ent = nullptr;  // nleaf x L
  const MvSeg* segs = nullptr;
  int nleaf = 0, L = 0, nseg = 0, KM = 0;
};

// part: nleaf * L * KM * nrhs doubles, t: nseg * KM * nrhs doubles.
void hodlr_matvec_fused(const MvPlan& p, const double* Dm, int64_t lm, const double* X, int64_t ldx,
                        int nrhs, double* part, double* t, double* Y, int64_t ldy, hipStream_t s);
size_t mv_lds_bytes(const MvPlan& p, int64_t lm, int nrhs);

}  // namespace kern
}  // namespace hmat
