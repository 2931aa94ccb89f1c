"""Writes pi_bits.h: 256 bits of 2/pi and of pi/2, computed exactly with integer arithmetic
(Machin's formula pi = 16 atan(1/5) - 4 atan(1/239) to 600 bits, then truncated). The
golden spiral's libm cos/sin reduce arguments past 2^20 pi/2 with these (libm_trig.h).

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
    bits = 600
    pi = 16 * atan_inv(5, bits + guard) - 4 * atan_inv(239, bits + guard)  # pi * 2^(600+64)
    scale = bits + guard
    # 2/pi * 2^256, truncated: bits 1..256 after the binary point (2/pi < 1)
    two_over_pi = (2 << (scale + 256)) // pi
    # pi/2 * 2^255, truncated (pi/2 < 2, so 256 bits)
    pio2 = pi >> (scale + 1 - 255)
    assert two_over_pi >> 256 == 0 and pio2 >> 256 == 0 and pio2 >> 255 == 1

    def words(v: int) -> str:
        w = [(v >> (64 * (3 - i))) & (2**64 - 1) for i in range(4)]
        return ", ".join(f"0x{x:016x}ull" for x in w)

    print("// Generated by gen_pi_bits.py: exact bits of 2/pi and pi/2, most significant word first.")
    print("#pragma once")
    print("#include <stdint.h>")
    print("// floor(2/pi * 2^256)")
    print(f"static const uint64_t GR_TWO_OVER_PI_256[4] = {{{words(two_over_pi)}}};")
    print("// floor(pi/2 * 2^255)")
    print(f"static const uint64_t GR_PIO2_256[4] = {{{words(pio2)}}};")


if __name__ == "__main__":
    main()
