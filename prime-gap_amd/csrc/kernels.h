// kernels.h -- kernel entry points and the registry the engine dispatches on.
//
// A kernel variant is (radix, digits, lanes per number, test kind).  Each
// instantiation unit (prp_r28_*.hip, ...) registers the variants it compiles;
// the engine picks the smallest variant whose digits cover a batch.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pgap {

constexpr int kMaxBases = 8;

struct Bases {
    uint32_t n;
    uint32_t b[kMaxBases];
};

enum Kind { KIND_FERMAT2 = 0, KIND_SPRP = 1 };

// Batch launch: numbers are rows of `limbs` words; verdicts out.
//   count    grid size (upper bound of the work)
//   d_count  if non-null, the actual count is read from device memory (so a
//            round can size its work on the device without a host round trip)
//   d_idx    if non-null, work item i tests row d_idx[i] and writes out[d_idx[i]]
typedef hipError_t (*batch_launch_fn)(const uint32_t *d_n, uint32_t limbs, uint32_t count, const uint32_t *d_count,
                                      const uint32_t *d_idx, Bases bases, uint8_t *d_out, int waves_per_simd,
                                      hipStream_t stream);

struct Variant {
    int radix;       // 28 or 32
    int digits;      // digits (radix-sized) per number
    int lanes;       // lanes cooperating on one number
    int kind;        // Kind
    int max_bits;    // largest N the variant accepts (lazy reduction needs N < R/16)
    batch_launch_fn launch;
    int (*blocks_per_cu)();  // resident 256-thread workgroups per CU (occupancy query)
};

// Registry: filled by every instantiation unit at library load.
void register_variant(const Variant &v);
const Variant *find_variant(int kind, int bits, int radix, int lanes);
size_t variant_count();
const Variant *variant_at(size_t i);

}  // namespace pgap
