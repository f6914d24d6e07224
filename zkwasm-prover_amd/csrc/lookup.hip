// Lookup multiplicities m(X) for gfx950.
//
// Drop-in for `prepare_table_lookup` and `calc_m` (cuda/bn254.cu:1637-1912,
// caller src/lib.rs:404-504).  Semantics kept: for every usable input row
// (row < unusable_rows_start) the table row it matches is the LARGEST usable
// row holding an equal value, and m[row] += (number of such inputs); m
// accumulates over calls (one per input of a lookup), so the caller zeroes it
// first.
//
// Method: instead of the reference's radix sort of (low k+2 bits of the
// Montgomery limb 0, row) pairs, a scan backwards through equal-key runs, a
// sort of the matched rows and a log-depth merge, the table's usable rows go
// into an open-addressing hash table (linear probing, 8n slots in the
// caller's start_offset buffer, so the load factor is at most 1/8): a row is
// inserted by compare-and-swap into an empty slot, or folded by atomic max
// into the slot that already holds an equal value.  calc_m then probes once
// per input row, counts matches with atomic increments on u32 counters (the
// caller's matched_index buffer) and adds the counts to m.  Two passes over
// the table and one over the input, no sort.
//
// An input that matches no usable table row cannot be counted (the reference
// traps); here its row is counted nowhere and the word sorted_matched_index[0]
// is set to 1 (it is cleared to 0 by every calc_m call).
#include "zkcore.h"

using namespace zk;

namespace {

constexpr uint32_t kEmpty = 0xffffffffu;
constexpr uint32_t kT = 256;

uint32_t blocks_for(uint32_t n) {
    uint32_t b = (n + kT - 1) / kT;
    return b < 65536 ? (b ? b : 1) : 65536;
}

__device__ __forceinline__ uint32_t fr_hash(const Fr &v) {
    uint32_t h = v.v[0] * 0x9e3779b1u;
#pragma unroll
    for (int i = 1; i < 8; i++) h = (h ^ v.v[i]) * 0x85ebca6bu;
    return h ^ (h >> 15);
}

__device__ __forceinline__ bool fr_eq(const Fr &a, const Fr &b) {
    uint32_t d = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d |= a.v[i] ^ b.v[i];
    return d == 0;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// insert the usable table rows; slots hold row indices
__global__ void k_table_insert(const Fr *__restrict__ table, uint32_t rows, uint32_t *slots, uint32_t mask) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < rows; j += gridDim.x * blockDim.x) {
        const Fr v = fe_load<FrParams>(&table[j]);
        uint32_t h = fr_hash(v) & mask;
        for (uint32_t probe = 0; probe <= mask; probe++, h = (h + 1) & mask) {
            uint32_t occ = atomicCAS(&slots[h], kEmpty, j);
            if (occ == kEmpty) break;  // inserted
            if (fr_eq(fe_load<FrParams>(&table[occ]), v)) {
                atomicMax(&slots[h], j);  // equal value: keep the largest row
                break;
            }
        }
    }
}

// count each usable input row against its table row
__global__ void k_input_match(const Fr *__restrict__ table, const Fr *__restrict__ input, uint32_t rows,
                              const uint32_t *__restrict__ slots, uint32_t mask, uint32_t *counts,
                              uint32_t *missing) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += gridDim.x * blockDim.x) {
        const Fr v = fe_load<FrParams>(&input[i]);
        uint32_t h = fr_hash(v) & mask;
        bool found = false;
        for (uint32_t probe = 0; probe <= mask; probe++, h = (h + 1) & mask) {
            const uint32_t occ = slots[h];
            if (occ == kEmpty) break;
            if (fr_eq(fe_load<FrParams>(&table[occ]), v)) {
                atomicAdd(&counts[occ], 1u);
                found = true;
                break;
            }
        }
        if (!found) *missing = 1u;
    }
}

// m[j] += counts[j] for the usable rows
__global__ void k_add_counts(Fr *m, const uint32_t *__restrict__ counts, uint32_t rows) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < rows; j += gridDim.x * blockDim.x) {
        const uint32_t c = counts[j];
        if (c) fe_store<FrParams>(&m[j], fe_load<FrParams>(&m[j]) + fe_from_u32<FrParams>(c));
    }
}

}  // namespace

extern "C" {

// cuda/bn254.cu:1867-1912 (caller src/lib.rs:431-449).  Every buffer is the
// reference's n = 2^k Fr (32n bytes); the hash table uses start_offset (8n
// u32 slots); table_lowest_u32, sorted_table_lowest_u32, index, sorted_index
// and the sort temp storage are not needed.
hipError_t prepare_table_lookup(void *table, void *table_lowest_u32, void *sorted_table_lowest_u32, void *index,
                                void *sorted_index, void *start_offset, void *candidate_sort_temp_storage,
                                unsigned candidate_sort_temp_storage_bytes, unsigned k, unsigned unusable_rows_start,
                                hipStream_t stream) {
    (void)table_lowest_u32;
    (void)sorted_table_lowest_u32;
    (void)index;
    (void)sorted_index;
    (void)candidate_sort_temp_storage;
    (void)candidate_sort_temp_storage_bytes;
    if (!table || !start_offset || k < 1 || k > 27 || unusable_rows_start > (1u << k)) return hipErrorInvalidValue;
    const uint32_t nslots = 8u << k;
    k_fill_u32<<<blocks_for(nslots), kT, 0, stream>>>((uint32_t *)start_offset, nslots, kEmpty);
    ZK_CHECK(hipGetLastError());
    if (!unusable_rows_start) return hipSuccess;
    k_table_insert<<<blocks_for(unusable_rows_start), kT, 0, stream>>>((const Fr *)table, unusable_rows_start,
                                                                       (uint32_t *)start_offset, nslots - 1);
    return hipGetLastError();
}

// cuda/bn254.cu:1794-1865 (caller src/lib.rs:466-492).  m (n Fr) accumulates
// the counts of this input; matched_index (n u32 of its 32n bytes) holds the
// per-row counters; sorted_matched_index[0] is the "input not in table" flag.
hipError_t calc_m(void *m, void *table, void *input, const void *sorted_index, const void *start_offset,
                  void *matched_index, void *sorted_matched_index, void *candidate_sort_temp_storage,
                  unsigned candidate_sort_temp_storage_bytes, unsigned k, unsigned unusable_rows_start,
                  hipStream_t stream) {
    (void)sorted_index;
    (void)candidate_sort_temp_storage;
    (void)candidate_sort_temp_storage_bytes;
    if (!m || !table || !input || !start_offset || !matched_index || !sorted_matched_index || k < 1 || k > 27 ||
        unusable_rows_start > (1u << k))
        return hipErrorInvalidValue;
    const uint32_t n = 1u << k, rows = unusable_rows_start;
    uint32_t *counts = (uint32_t *)matched_index;
    uint32_t *missing = (uint32_t *)sorted_matched_index;
    k_fill_u32<<<blocks_for(n), kT, 0, stream>>>(counts, n, 0u);
    ZK_CHECK(hipGetLastError());
    k_fill_u32<<<1, 1, 0, stream>>>(missing, 1, 0u);
    ZK_CHECK(hipGetLastError());
    if (!rows) return hipSuccess;
    k_input_match<<<blocks_for(rows), kT, 0, stream>>>((const Fr *)table, (const Fr *)input, rows,
                                                       (const uint32_t *)start_offset, (8u << k) - 1, counts, missing);
    ZK_CHECK(hipGetLastError());
    k_add_counts<<<blocks_for(rows), kT, 0, stream>>>((Fr *)m, counts, rows);
    return hipGetLastError();
}

}  // extern "C"
