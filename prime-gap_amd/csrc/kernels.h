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

// FERMAT2 and SPRP have one-lane and lane-group variants; LUCAS (the strong
// Lucas test, lucas_lg.h) has lane-group variants only.
enum Kind { KIND_FERMAT2 = 0, KIND_SPRP = 1, KIND_LUCAS = 2 };

// Batch launch: numbers are rows of `limbs` words; verdicts out.  A strong-PRP
// launch tests one base, bases.b[0] (bases.n must be 1; callers run further
// bases on the survivors through d_idx): a loop over bases inside the kernel
// keeps the whole test live across iterations and costs the lane-group
// kernels a wavefront per SIMD.
//   count    grid size (upper bound of the work)
//   d_count  if non-null, the actual count is read from device memory (so a
//            round can size its work on the device without a host round trip)
//   d_idx    if non-null, work item i tests row d_idx[i] and writes out[d_idx[i]]
//   blocks_cap  if > 0, at most this many 256-thread workgroups per CU (an
//            occupancy cap the autotuner can choose; enforced by reserving
//            LDS the kernel does not use)
typedef hipError_t (*batch_launch_fn)(const uint32_t *d_n, uint32_t limbs, uint32_t count, const uint32_t *d_count,
                                      const uint32_t *d_idx, Bases bases, uint8_t *d_out, int blocks_cap,
                                      hipStream_t stream);

// Dynamic LDS bytes that leave room for at most `cap` workgroups per CU
// (160 KiB of LDS per CU on gfx950); 0 when cap <= 0.  Sets the kernel's
// dynamic-LDS limit when the reservation exceeds the 64 KiB default.
template <class K>
inline size_t occupancy_cap_bytes(K kernel, int cap) {
    if (cap <= 0) return 0;
    const size_t lds_per_cu = 160 * 1024;
    const size_t bytes = (lds_per_cu / (size_t)cap) & ~(size_t)255;  // cap + 1 groups would not fit
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return bytes;
}

// Streaming-scan launch: one grid runs a round's confirmation tests and its
// Fermat tests together, so a round costs one test's latency.
//   rows [0, nc)                       strong PRP to base d_base[i], nc = min(d_ctr[3], conf_max)
//   rows [conf_rows, conf_rows + nf)   Fermat base 2, nf = min(d_ctr[0], max_items)
// Confirmation items come first in the grid, padded to whole wavefronts, so
// every wavefront runs one of the two tests.
typedef hipError_t (*scan_launch_fn)(const uint32_t *d_n, uint32_t limbs, uint32_t conf_rows, uint32_t conf_max,
                                     uint32_t max_items, const uint32_t *d_ctr, const uint8_t *d_base,
                                     uint8_t *d_out, int blocks_cap, hipStream_t stream);

struct Variant {
    int radix;       // digit radix in bits (28)
    int digits;      // digits (radix-sized) per number
    int lanes;       // lanes cooperating on one number
    int kind;        // Kind
    int max_bits;    // largest N the variant accepts: Fermat N < R/16; strong PRP N < R/2^8,
                     // which covers bases up to 8 (see select_bits); strong Lucas N < R/2^54
    batch_launch_fn launch;
    int (*blocks_per_cu)();  // resident 256-thread workgroups per CU (occupancy query)
    scan_launch_fn scan = nullptr;  // set on strong-PRP variants (max_bits is the SPRP limit)
    int (*scan_blocks_per_cu)() = nullptr;  // occupancy of the scan kernel
};

// Width to select a variant with.  Strong PRP multiplies by the base as a
// scalar: x < 2N becomes x < 2bN, and the next lazy squaring returns < 2N
// while 4b^2 N < R.  Variants register max_bits for bases up to 8 (N < R/2^8);
// a base in [9, 63] needs N < R/2^14, six bits more headroom.
inline uint32_t select_bits(int kind, uint32_t bits, const Bases &b) {
    if (kind != KIND_SPRP) return bits;
    for (uint32_t i = 0; i < b.n; ++i)
        if (b.b[i] > 8) return bits + 6;
    return bits;
}

// Registry: filled by every instantiation unit at library load.
void register_variant(const Variant &v);
const Variant *find_variant(int kind, int bits, int radix, int lanes);
size_t variant_count();
const Variant *variant_at(size_t i);

}  // namespace pgap
