"""Generate hfv_pow_tables.hpp: the tables of hfv_pow.hpp, computed with
60-digit decimal arithmetic.

  log:  48 intervals of width 1/64 covering m in [0.75, 1.5); per interval
        c = RN(1/m_mid) and -log(c) as a double-double (hi, lo).
  exp:  2^(j/64), j = 0..63, as double-doubles.
  ln 2: split for exact products: LN2_HI has its low 11 bits zero (k * LN2_HI
        is exact for |k| < 2^11); LN2_64_HI = ln2/64 with its low 17 bits zero
        (n * LN2_64_HI exact for |n| < 2^17).

Usage: python3 gen_pow_tables.py > hfv_pow_tables.hpp
"""
from decimal import Decimal, getcontext

getcontext().prec = 60


def rn(d: Decimal) -> float:
    return float(d)  # correctly rounded conversion


def dd(d: Decimal):
    hi = rn(d)
    lo = rn(d - Decimal(hi))
    return hi, lo


def keep_bits(x: float, bits: int) -> float:
    import math
    m, e = math.frexp(x)
    return math.ldexp(round(math.ldexp(m, bits)), e - bits)


def main():
    ln2 = Decimal(2).ln()
    out = ["// Generated by gen_pow_tables.py; do not edit.", "#pragma once", "",
           "namespace hfv {", ""]
    hi = keep_bits(rn(ln2), 42)
    out.append(f"constexpr double POW_LN2_HI = {hi.hex()};")
    out.append(f"constexpr double POW_LN2_LO = {rn(ln2 - Decimal(hi)).hex()};")
    l64 = ln2 / 64
    hi64 = keep_bits(rn(l64), 36)
    out.append(f"constexpr double POW_LN2_64_HI = {hi64.hex()};")
    out.append(f"constexpr double POW_LN2_64_LO = {rn(l64 - Decimal(hi64)).hex()};")
    out.append(f"constexpr double POW_64_OVER_LN2 = {rn(64 / ln2).hex()};")
    out.append("")
    out.append("// log: {c, -log(c) hi, -log(c) lo} for m in [0.75 + i/64, 0.75 + (i+1)/64)")
    out.append("__device__ __constant__ const double POW_LOG_TAB_D[48 * 3] = {")
    rows = []
    for i in range(48):
        lo_m = Decimal(3) / 4 + Decimal(i) / 64
        mid = lo_m + Decimal(1) / 128
        c = rn(1 / mid)
        nl = -(Decimal(c).ln())
        h, l = dd(nl)
        rows.append(f"    {c.hex()}, {h.hex()}, {l.hex()},")
    out += rows
    out.append("};")
    out.append("static const double POW_LOG_TAB_H[48 * 3] = {")
    out += rows
    out.append("};")
    out.append("")
    out.append("// exp: 2^(j/64) = hi + lo")
    out.append("__device__ __constant__ const double POW_EXP_TAB_D[64 * 2] = {")
    rows = []
    for j in range(64):
        h, l = dd(Decimal(2) ** (Decimal(j) / 64))
        rows.append(f"    {h.hex()}, {l.hex()},")
    out += rows
    out.append("};")
    out.append("static const double POW_EXP_TAB_H[64 * 2] = {")
    out += rows
    out.append("};")
    out.append("")
    out.append("}  // namespace hfv")
    print("\n".join(out))


if __name__ == "__main__":
    main()
