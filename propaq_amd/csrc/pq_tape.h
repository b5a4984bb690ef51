// pq_tape.h -- device evaluation of a compiled surrogate tape.
//
// A compiled surrogate (crates/surrogate/src/symcoeff.rs:748-850) is a
// straight-line program: op i is a scalar, a sum of two earlier ops, or an
// earlier op times a constant, cos(theta_p) or sin(theta_p).  On the device
// the ops are ordered by level (an op's level is one more than its deepest
// operand), so every op of a level can run at once, and each op is evaluated
// for a batch of B parameter sets: value[op][s], s fastest, so one wavefront
// reads and writes 64 consecutive doubles of one op.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pq {

enum TapeKind : uint32_t { kTapeScalar = 0, kTapeAdd = 1, kTapeScale = 2, kTapeCos = 3, kTapeSin = 4 };

// 24 bytes: val is the scalar or the Scale factor; a, b operand op indices;
// p the parameter index of Cos / Sin.
struct TapeOp {
  double val;
  uint32_t a, b, p, kind;
};

// The device form of the tape: only ops whose value is read more than once
// (or is a term's root) are stored; a chain of up to kChain single-operand
// ops (Scale, Cos, Sin) read only by one op is folded into that op's operand
// and applied in registers.  The arithmetic and its order are the tape's, so
// values are unchanged; the traffic drops to one value written and one per
// stored operand read.
constexpr int kChain = 4;

struct FusedOp {
  double val;                  // Scalar value
  double f[2 * kChain];        // Scale factor of each step (operand a's steps, then b's)
  uint32_t a, b;               // stored operands, device order
  uint32_t step[2 * kChain];   // (param << 3) | kind (kTapeScale / kTapeCos / kTapeSin), innermost first
  uint32_t kind, na, nb, pad;  // kind: kTapeScalar, kTapeAdd (a + b) or kTapeScale (operand a alone);
                               // dwords, so a uniform op is read with scalar loads
};

constexpr uint32_t kTapeEmpty = 0xFFFFFFFFu;  // a term whose coefficient pruned to nothing
constexpr int kTermChunk = 32;                // terms summed per thread, pairwise in registers
constexpr int kBlockLevelThreads = 1024;

// Ops [lo, lo + n) of one level, for B parameter sets.  lut holds cos/sin of
// every parameter per set: lut[(2p) B + s] = cos, lut[(2p+1) B + s] = sin.
hipError_t launch_tape_level(const FusedOp* ops, uint32_t lo, uint32_t n, const double* lut, double* val,
                             uint32_t B, hipStream_t st);

// Levels [l0, l1) in one workgroup with a barrier between levels: for runs of
// levels too small to fill the chip, where a launch per level would cost more
// than the work.  level_off is the device copy of the level offsets.
hipError_t launch_tape_levels_block(const FusedOp* ops, const uint32_t* level_off, uint32_t l0, uint32_t l1,
                                    const double* lut, double* val, uint32_t B, hipStream_t st);

// One pass of the pairwise term sum: out[c][s] = the pairwise sum of the
// kTermChunk values of chunk c (zero padded).  First pass (roots non-null):
// value t = overlap[t] * val[roots[t]][s] (0 for an empty root); later
// passes: value t = in[t][s].
hipError_t launch_term_sum(const uint32_t* roots, const double* overlap, const double* in, uint64_t n, uint32_t B,
                           double* out, hipStream_t st);

}  // namespace pq
