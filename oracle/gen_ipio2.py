"""Test infrastructure (oracle): writes ipio2.h, the two tables of the large-argument
reduction the reference's libm cos/sin use (the Rust libm crate 0.2.16's rem_pio2_large,
musl's __rem_pio2_large, Sun fdlibm's __kernel_rem_pio2):

- IPIO2[k], k < 66: the 24-bit chunks of 2/pi, IPIO2[k] = floor(2/pi * 2^(24(k+1))) mod 2^24;
- PIO2_24[k], k < 8: pi/2 cut into 24-bit pieces, PIO2_24[k] = (floor(pi/2 * 2^(23+24k)) mod
  2^24) * 2^-(23+24k), each exact in a double.

Both come from pi computed exactly with integer arithmetic (Machin's formula
pi = 16 atan(1/5) - 4 atan(1/239), 2000 bits plus guard bits, truncated), not typed in.

    python oracle/gen_ipio2.py > oracle/ipio2.h
"""


def atan_inv(x: int, bits: int) -> int:
    """atan(1/x) * 2^bits, truncated (the caller's guard bits absorb the error)."""
    one = 1 << bits
    total, term, k, sign = 0, one // x, 1, 1
    while term:
        total += sign * (term // k)
        term //= x * x
        k += 2
        sign = -sign
    return total


def tables():
    bits, guard = 2000, 64
    scale = bits + guard
    pi = 16 * atan_inv(5, scale) - 4 * atan_inv(239, scale)  # pi * 2^scale
    ipio2 = []
    for k in range(66):
        sh = 24 * (k + 1)
        ipio2.append(((2 << (scale + sh)) // pi) & 0xFFFFFF)
    pio2 = []
    for k in range(8):
        sh = 23 + 24 * k
        chunk = (pi >> (scale + 1 - sh)) & 0xFFFFFF  # floor(pi/2 * 2^sh) mod 2^24
        pio2.append((chunk, sh))
    return ipio2, pio2


def main() -> None:
    ipio2, pio2 = tables()
    print("/* Generated by gen_ipio2.py: 2/pi in 24-bit chunks and pi/2 in 24-bit pieces,")
    print(" * from pi computed exactly (test infrastructure, the oracle's libm restatement). */")
    print("#pragma once")
    print("#include <stdint.h>")
    print("static const int32_t IPIO2[66] = {")
    for r in range(0, 66, 6):
        print("    " + ", ".join(f"0x{v:06X}" for v in ipio2[r:r + 6]) + ",")
    print("};")
    print("/* PIO2_24[k] = chunk * 2^-shift, written as exact hexadecimal doubles */")
    print("static const double PIO2_24[8] = {")
    for chunk, sh in pio2:
        print(f"    {float.hex(chunk / 2 ** sh)},")
    print("};")


if __name__ == "__main__":
    main()
