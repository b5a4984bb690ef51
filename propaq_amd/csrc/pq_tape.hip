// pq_tape.hip -- gfx950 kernels that evaluate a compiled surrogate tape for a
// batch of parameter sets (SurrogateModel::evaluate / evaluate_batch,
// crates/surrogate/src/model.rs:63-103, CompiledCoeff::evaluate_into,
// crates/surrogate/src/symcoeff.rs:836-850).
//
// The reference walks the tape once per parameter set on a CPU thread.  Here
// a level of independent ops runs across the chip for every set at once:
// thread (op, s) reads its operands' values for set s and writes its own.
// The work is a few FP64 operations per 16-24 bytes, so the kernels are
// HBM-bound; value[op][s] keeps a wavefront's accesses on 64 consecutive
// doubles of one op and the op record is a wavefront-uniform load.
#include "pq_tape.h"

namespace pq {
namespace {

// The steps [off, off + n) applied to x, innermost first: f * x for Scale,
// cos or sin of the step's parameter times x (symcoeff.rs:843-847).  The
// loop has a fixed trip count so no array is indexed at run time, and the op
// is read in place (a uniform load when every lane has the same op).
__device__ __forceinline__ double apply_chain(const FusedOp* __restrict__ op, int off, int n, double x,
                                              const double* __restrict__ lut, uint32_t B, uint32_t s) {
#pragma unroll
  for (int j = 0; j < kChain; ++j) {
    if (j < n) {
      const uint32_t st = op->step[off + j];
      const uint32_t k = st & 7u, p = st >> 3;
      if (k == kTapeScale)
        x = op->f[off + j] * x;
      else
        x = lut[(uint64_t)(2 * p + (k == kTapeSin ? 1 : 0)) * B + s] * x;
    }
  }
  return x;
}

__device__ __forceinline__ double eval_fused(const FusedOp* __restrict__ op, uint32_t s,
                                             const double* __restrict__ lut, const double* __restrict__ val,
                                             uint32_t B) {
  const uint32_t kind = op->kind;
  if (kind == kTapeScalar) return op->val;
  const double x = apply_chain(op, 0, op->na, val[(uint64_t)op->a * B + s], lut, B, s);
  if (kind != kTapeAdd) return x;
  return x + apply_chain(op, kChain, op->nb, val[(uint64_t)op->b * B + s], lut, B, s);
}

__device__ __forceinline__ void eval_op(const FusedOp* __restrict__ ops, uint32_t o, uint32_t s,
                                        const double* __restrict__ lut, double* __restrict__ val, uint32_t B) {
  val[(uint64_t)o * B + s] = eval_fused(ops + o, s, lut, val, B);
}

__global__ void __launch_bounds__(256) k_tape_level(const FusedOp* __restrict__ ops, uint32_t lo, uint64_t work,
                                                    const double* __restrict__ lut, double* __restrict__ val,
                                                    uint32_t B) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < work; t += stride) {
    const uint32_t o = lo + (uint32_t)(t / B);
    eval_op(ops, o, (uint32_t)(t % B), lut, val, B);
  }
}

// Wide batches: a workgroup takes one op for kWideSets consecutive sets, so
// the op record is a uniform (scalar) load and no thread divides by B; each
// thread keeps kWidePer independent loads in flight.
constexpr uint32_t kWideThreads = 256;
constexpr uint32_t kWidePer = 4;
constexpr uint32_t kWideSets = kWideThreads * kWidePer;

__global__ void __launch_bounds__(kWideThreads) k_tape_level_wide(const FusedOp* __restrict__ ops, uint32_t lo,
                                                                  uint64_t n_blocks, uint32_t tiles,
                                                                  const double* __restrict__ lut,
                                                                  double* __restrict__ val, uint32_t B) {
  for (uint64_t blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    const uint32_t o = lo + (uint32_t)(blk / tiles);
    const uint32_t s0 = (uint32_t)(blk % tiles) * kWideSets + threadIdx.x;
    const FusedOp* op = ops + o;
    double x[kWidePer];
#pragma unroll
    for (uint32_t j = 0; j < kWidePer; ++j) {
      const uint32_t s = s0 + j * kWideThreads;
      if (s < B) x[j] = eval_fused(op, s, lut, val, B);
    }
    double* po = val + (uint64_t)o * B;
#pragma unroll
    for (uint32_t j = 0; j < kWidePer; ++j) {
      const uint32_t s = s0 + j * kWideThreads;
      if (s < B) po[s] = x[j];
    }
  }
}

// A workgroup's stores are visible to its other waves after the barrier:
// they share one CU and its L1.
__global__ void __launch_bounds__(kBlockLevelThreads)
    k_tape_levels_block(const FusedOp* __restrict__ ops, const uint32_t* __restrict__ level_off, uint32_t l0,
                        uint32_t l1, const double* __restrict__ lut, double* __restrict__ val, uint32_t B) {
  for (uint32_t l = l0; l < l1; ++l) {
    const uint32_t lo = level_off[l];
    const uint64_t work = (uint64_t)(level_off[l + 1] - lo) * B;
    for (uint64_t t = threadIdx.x; t < work; t += kBlockLevelThreads) {
      const uint32_t o = lo + (uint32_t)(t / B);
      eval_op(ops, o, (uint32_t)(t % B), lut, val, B);
    }
    __syncthreads();
  }
}

// Pairwise sum of kTermChunk values per thread, in registers: level by
// level, v[i] = v[2i] + v[2i+1] (adjacent pairs).  Chunks are a power of two
// long, so chunked passes give the same tree as one pass over the whole
// zero-padded sequence.
__global__ void __launch_bounds__(256) k_term_sum(const uint32_t* __restrict__ roots,
                                                  const double* __restrict__ overlap,
                                                  const double* __restrict__ in, uint64_t n, uint32_t B,
                                                  uint64_t n_chunks, double* __restrict__ out) {
  const uint64_t work = n_chunks * B;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < work; t += stride) {
    const uint64_t c = t / B;
    const uint32_t s = (uint32_t)(t % B);
    double v[kTermChunk];
#pragma unroll
    for (int i = 0; i < kTermChunk; ++i) {
      const uint64_t idx = c * kTermChunk + i;
      double x = 0.0;
      if (idx < n) {
        if (roots) {
          const uint32_t r = roots[idx];
          x = r == kTapeEmpty ? 0.0 : overlap[idx] * in[(uint64_t)r * B + s];
        } else {
          x = in[idx * B + s];
        }
      }
      v[i] = x;
    }
#pragma unroll
    for (int w = kTermChunk / 2; w >= 1; w /= 2) {
#pragma unroll
      for (int i = 0; i < w; ++i) v[i] = v[2 * i] + v[2 * i + 1];
    }
    out[c * B + s] = v[0];
  }
}

unsigned grid_for(uint64_t work) {
  const uint64_t blocks = (work + 255) / 256;
  return (unsigned)(blocks < 262144 ? (blocks ? blocks : 1) : 262144);
}

}  // namespace

hipError_t launch_tape_level(const FusedOp* ops, uint32_t lo, uint32_t n, const double* lut, double* val,
                             uint32_t B, hipStream_t st) {
  const uint64_t work = (uint64_t)n * B;
  if (work == 0) return hipSuccess;
  if (B >= kWideThreads / 2) {
    const uint32_t tiles = (B + kWideSets - 1) / kWideSets;
    const uint64_t n_blocks = (uint64_t)n * tiles;
    const unsigned grid = (unsigned)(n_blocks < 1048576 ? n_blocks : 1048576);
    hipLaunchKernelGGL(k_tape_level_wide, dim3(grid), dim3(kWideThreads), 0, st, ops, lo, n_blocks, tiles, lut, val,
                       B);
  } else {
    hipLaunchKernelGGL(k_tape_level, dim3(grid_for(work)), dim3(256), 0, st, ops, lo, work, lut, val, B);
  }
  return hipGetLastError();
}

hipError_t launch_tape_levels_block(const FusedOp* ops, const uint32_t* level_off, uint32_t l0, uint32_t l1,
                                    const double* lut, double* val, uint32_t B, hipStream_t st) {
  if (l1 <= l0) return hipSuccess;
  hipLaunchKernelGGL(k_tape_levels_block, dim3(1), dim3(kBlockLevelThreads), 0, st, ops, level_off, l0, l1, lut,
                     val, B);
  return hipGetLastError();
}

hipError_t launch_term_sum(const uint32_t* roots, const double* overlap, const double* in, uint64_t n, uint32_t B,
                           double* out, hipStream_t st) {
  const uint64_t n_chunks = (n + kTermChunk - 1) / kTermChunk;
  if (n_chunks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_term_sum, dim3(grid_for(n_chunks * B)), dim3(256), 0, st, roots, overlap, in, n, B, n_chunks,
                     out);
  return hipGetLastError();
}

}  // namespace pq
