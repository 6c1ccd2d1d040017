// Internal host-side helpers shared by the ntt_amd library sources.
#pragma once

#include <cstddef>
#include <cstdint>

#include "../../include/ntt_amd.h"

// Largest transform size accepted (the two-pass engine; every size up to it
// is checked against the oracle on the GPU).
#define NTT_MAX_LOG_N 20
#define NTT_MAX_STAGE_MERGING 5
// Sizes the fused limb kernel handles (one workgroup holds 2^16 words).
#define FUSED_MIN_LOG_N 10
#define FUSED_MAX_LOG_N 16
// gfx950: 160 KiB of LDS per CU, all of it addressable by one workgroup.
#define NTT_LDS_BYTES_GFX950 (160u * 1024u)

namespace ntt {

uint32_t powmod(uint32_t b, uint64_t e, uint32_t m);
uint32_t mont_q_inv_host(uint32_t q);
uint32_t to_mont_host(uint32_t x, uint32_t q);
uint32_t r2_host(uint32_t q);
uint32_t from_mont_host(uint32_t x, uint32_t q, uint32_t q_inv);
bool is_prime_u32(uint32_t n);
uint32_t smallest_primitive_root(uint32_t q);
int check_modulus(int logN, uint32_t q);
int make_runtime(int logN, uint32_t q, uint32_t g, ntt_modulus_runtime* out);
void twiddles_bitrev(int logN, uint32_t q, uint32_t root, bool mont, uint32_t* out);
void reference_config(int logN, ntt_transform_config* c);
void default_config(int logN, ntt_transform_config* c);
int validate_config(int logN, const ntt_transform_config* c, size_t max_lds_bytes);

// LDS words of a tile of `words` entries after the one-word-per-32 padding
// applied by the kernels (lds_pad), in bytes.
inline size_t lds_bytes_for(size_t words) { return (words + (words >> 5) + 1) * sizeof(uint32_t); }

}  // namespace ntt
