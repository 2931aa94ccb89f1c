"""Writes pi_bits.h: the two tables libm's large-argument reduction (rem_pio2_large of the
Rust libm crate 0.2.16, musl's __rem_pio2_large, fdlibm's __kernel_rem_pio2) multiplies by,
computed exactly with integer arithmetic (Machin's formula pi = 16 atan(1/5) - 4 atan(1/239)
to 1700 bits plus guard bits, then truncated):

- GR_TWO_OVER_PI_24[k], k < 66: bits 24k+1 .. 24k+24 of 2/pi after the binary point;
- GR_PIO2_24[k], k < 8: pi/2 cut into consecutive 24-bit pieces, piece k scaled by
  2^-(23 + 24k), so that the eight sum to pi/2 truncated after 215 bits.

The golden spiral's cos/sin reduce arguments past 2^20 pi/2 with them (libm_trig.h).

    python graph_render_amd/csrc/gen_pi_bits.py > graph_render_amd/csrc/pi_bits.h
"""


def atan_inv(x: int, bits: int) -> int:
    """atan(1/x) * 2^bits, truncated, with guard bits absorbed by the caller."""
    one = 1 << bits
    total, term, k, sign = 0, one // x, 1, 1
    while term:
        total += sign * (term // k)
        term //= x * x
        k += 2
        sign = -sign
    return total


def main() -> None:
    guard = 64
    bits = 1700
    scale = bits + guard
    pi = 16 * atan_inv(5, scale) - 4 * atan_inv(239, scale)  # pi * 2^scale
    two_over_pi = (2 << (scale + 66 * 24)) // pi  # floor(2/pi * 2^(66*24))
    chunks = [(two_over_pi >> (24 * (65 - k))) & 0xFFFFFF for k in range(66)]
    assert chunks[0] == 0xA2F983  # 2/pi = 0.A2F9836E4E44...
    pieces = []
    for k in range(8):
        sh = 23 + 24 * k
        piece = (pi >> (scale + 1 - sh)) & 0xFFFFFF  # floor(pi/2 * 2^sh) mod 2^24
        pieces.append(float.hex(piece / 2 ** sh))
    print("// Generated by gen_pi_bits.py: 2/pi in 24-bit chunks and pi/2 in 24-bit pieces, exact.")
    print("#pragma once")
    print("#include <stdint.h>")
    print("// GR_TWO_OVER_PI_24[k] = floor(2/pi * 2^(24(k+1))) mod 2^24")
    print("static const int32_t GR_TWO_OVER_PI_24[66] = {")
    for r in range(0, 66, 6):
        print("    " + ", ".join(f"0x{v:06X}" for v in chunks[r:r + 6]) + ",")
    print("};")
    print("// GR_PIO2_24[k] = (floor(pi/2 * 2^(23+24k)) mod 2^24) * 2^-(23+24k)")
    print("static const double GR_PIO2_24[8] = {" + ", ".join(pieces) + "};")


if __name__ == "__main__":
    main()
