// CPU harness for the optimized binary GCD (csrc/inv_bingcd.h) on the BN254
// scalar and base field moduli.
#include <stdint.h>

#include "inv_bingcd.h"

namespace {
const uint32_t kR[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                        0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
const uint32_t kP[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                        0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
}  // namespace

extern "C" {
// field 0: r (minv 0xefffffff), 1: p (minv 0xe4866389); returns 1 when b ended at 1
int ib_inv(uint32_t *v, const uint32_t *y, int field) {
    return zkinv::bingcd(v, y, field ? kP : kR, field ? 0xe4866389u : 0xefffffffu) ? 1 : 0;
}
int ib_bits() { return zkinv::kBits; }
int ib_steps() { return zkinv::kSteps; }
}
