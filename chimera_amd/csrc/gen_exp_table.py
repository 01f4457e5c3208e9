"""Write exp_table.hpp: the 2^(k/128) table of the table-driven exp in libm_exp.hpp.

For k = 0..127, with v = 2^(k/128) evaluated to 80 significant digits:
  H_k = v rounded to the nearest double,
  T_k = (v - H_k) / H_k rounded to the nearest double,
so that 2^(k/128) = H_k (1 + T_k) to about 2^-106.  Entry 2k holds the bits of T_k,
entry 2k+1 the bits of H_k less k << 45 (k/128 of an exponent unit), so that adding
the integer part of the scaled argument, shifted into the exponent field, gives the
bits of 2^(ki/128) for any integer ki.

    python3 chimera_amd/csrc/gen_exp_table.py      # rewrites exp_table.hpp beside it
"""
from __future__ import annotations

import decimal
import struct
from pathlib import Path

N = 128


def _bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def table() -> list[int]:
    ctx = decimal.Context(prec=80)
    out = []
    for k in range(N):
        v = ctx.power(decimal.Decimal(2), ctx.divide(decimal.Decimal(k), decimal.Decimal(N)))
        h = float(v)                                   # nearest double (via the decimal string)
        t = float(ctx.divide(ctx.subtract(v, decimal.Decimal(h)), decimal.Decimal(h)))
        out += [_bits(t), (_bits(h) - ((k << 52) // N)) & 0xFFFFFFFFFFFFFFFF]
    return out


def render() -> str:
    rows = table()
    lines = ["// Generated by gen_exp_table.py: 2^(k/128) = H_k (1 + T_k); entry 2k is the",
             "// bits of T_k, entry 2k+1 the bits of H_k minus (k << 45).",
             "#pragma once", "", "#include <cstdint>", "", "namespace chimera {", "",
             "static constexpr uint64_t EXP_TABLE_2K128[256] = {"]
    for k in range(N):
        lines.append(f"    0x{rows[2 * k]:016x}ull, 0x{rows[2 * k + 1]:016x}ull,")
    lines += ["};", "", "}  // namespace chimera", ""]
    return "\n".join(lines)


if __name__ == "__main__":
    Path(__file__).with_name("exp_table.hpp").write_text(render())
