// Exposes the host Fr arithmetic of the prover core (host_fr.h) to the CPU
// tests, which compare it with the oracle.
#include "host_fr.h"

extern "C" {
void hf_add(zkh::Fr *r, const zkh::Fr *a, const zkh::Fr *b) { *r = zkh::fr_add(*a, *b); }
void hf_sub(zkh::Fr *r, const zkh::Fr *a, const zkh::Fr *b) { *r = zkh::fr_sub(*a, *b); }
void hf_mul(zkh::Fr *r, const zkh::Fr *a, const zkh::Fr *b) { *r = zkh::fr_mul(*a, *b); }
void hf_inv(zkh::Fr *r, const zkh::Fr *a) { *r = zkh::fr_inv(*a); }
void hf_from_u64(zkh::Fr *r, unsigned long long v) { *r = zkh::fr_from_u64(v); }
}
