// Generated by gen_pi_bits.py: exact bits of 2/pi and pi/2, most significant word first.
#pragma once
#include <stdint.h>
// floor(2/pi * 2^256)
static const uint64_t GR_TWO_OVER_PI_256[4] = {0xa2f9836e4e441529ull, 0xfc2757d1f534ddc0ull, 0xdb6295993c439041ull, 0xfe5163abdebbc561ull};
// floor(pi/2 * 2^255)
static const uint64_t GR_PIO2_256[4] = {0xc90fdaa22168c234ull, 0xc4c6628b80dc1cd1ull, 0x29024e088a67cc74ull, 0x020bbea63b139b22ull};
