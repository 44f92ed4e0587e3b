"""Write tests/golden/hp_ops.json: known answers for the truncating HP arithmetic.

The reference's HP arithmetic is GMP mpf (complex.h:4), whose published contract is
"results truncated toward zero" (and mpf_get_d truncates).  Both HP
implementations here keep P = 64 * ((bits + 127) // 64 + 1) bits and truncate the
exact result; mpmath computes the same thing independently with rounding mode
'd' (toward zero) at precision P.  Operands are first parsed from decimal at P
bits with the same rounding.

Each record: op, a, b, iv, dv, bits -> sign, exp, mantissa as 32-bit words (most
significant first) of value = sign * 0.m * 2^exp.
Usage: python tests/golden/make_hp_goldens.py
"""
import json
import os
import random

import mpmath
from mpmath.libmp import from_float, from_int, from_str, mpf_add, mpf_div, mpf_mul, mpf_sub

HERE = os.path.dirname(os.path.abspath(__file__))


def prec_bits(bits):
    return 64 * ((bits + 127) // 64 + 1)


def encode(x, P):
    sign, man, exp, bc = x
    if man == 0:
        return 0, 0, [0] * (P // 32)
    s = -1 if sign else 1
    # value = man * 2^exp = 0.M * 2^(exp + bc), M = man << (P - bc)
    M = man << (P - bc)
    words = [(M >> (32 * (P // 32 - 1 - i))) & 0xFFFFFFFF for i in range(P // 32)]
    return s, exp + bc, words


def rand_dec(rng):
    m = rng.randint(1, 10 ** rng.randint(1, 70))
    e = rng.randint(-120, 40)
    s = "-" if rng.random() < 0.5 else ""
    if rng.random() < 0.3:
        return "%s%s.%se%d" % (s, str(m)[:1], str(m)[1:] or "0", e)
    return "%s%de%d" % (s, m, e)


def main():
    rng = random.Random(20261202)
    recs = []
    specials = ["0", "1", "-1", "0.25", "-0.75", "1e-30", "-0.5", "3", "1024", "0.1"]
    for bits in (64, 109, 200, 341, 672):
        P = prec_bits(bits)
        for t in range(24):
            a = rand_dec(rng) if t >= 10 else specials[t]
            b = rand_dec(rng) if t % 7 else specials[t % len(specials)]
            A = from_str(a, P, "d")
            B = from_str(b, P, "d")
            for op in range(7):
                iv = rng.randint(1, 8000) if op == 4 else rng.randint(-8000, 8000)
                dv = rng.uniform(-3, 3)
                if op == 0:
                    r = mpf_add(A, B, P, "d")
                elif op == 1:
                    r = mpf_sub(A, B, P, "d")
                elif op == 2:
                    r = mpf_mul(A, B, P, "d")
                elif op == 3:
                    r = mpf_mul(A, from_int(iv), P, "d")
                elif op == 4:
                    r = mpf_div(A, from_int(iv), P, "d")
                elif op == 5:
                    r = A
                else:
                    r = mpf_mul(A, from_float(dv), P, "d")
                s, e, w = encode(r, P)
                recs.append({"op": op, "a": a, "b": b, "iv": iv, "dv": dv, "bits": bits, "sign": s, "exp": e,
                             "words": w})
    with open(os.path.join(HERE, "hp_ops.json"), "w") as f:
        json.dump({"_about": "mpmath %s round-down known answers; see make_hp_goldens.py" % mpmath.__version__,
                   "records": recs}, f)
    print(len(recs), "records")


if __name__ == "__main__":
    main()
