#!/usr/bin/env python3
"""Generate bls12_381_consts.hpp: every arithmetic constant the HIP library
uses, derived from the BLS parameter x = -0xd201000000010000.

Fp elements are emitted as 14 little-endian limbs of 28 bits in Montgomery
form with R = 2^392; Fr elements as 8 little-endian 32-bit limbs in Montgomery
form with R = 2^256.  Exponents stay 12 x 32-bit words.  The 11-isogeny
coefficients come from h2c_iso_consts.h (derived by the oracle-side script and
pinned there against the CFRG generator vectors); this script converts them.

Run: python3 cubbs-cred_amd/csrc/gen_consts.py  (rewrites the header in place)
"""
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
X = -0xD201000000010000
P = ((X - 1) ** 2 * (X ** 4 - X ** 2 + 1)) // 3 + X
R = X ** 4 - X ** 2 + 1
NL = 14          # Fp limbs
LB = 28          # bits per Fp limb
RP = 1 << (NL * LB)   # Montgomery radix of Fp: 2^392
RR = 1 << 256


def limbs(v, n):
    return [(v >> (32 * i)) & 0xFFFFFFFF for i in range(n)]


def limbs28(v):
    assert 0 <= v < RP
    return [(v >> (LB * i)) & ((1 << LB) - 1) for i in range(NL)]


def mont(v):
    return v * RP % P


def montr(v):
    return v * RR % R


# ----- Fp2 helpers (u^2 = -1) as tuples
def f2mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2pow(a, e):
    r = (1, 0)
    while e:
        if e & 1:
            r = f2mul(r, a)
        a = f2mul(a, a)
        e >>= 1
    return r


def ec_add(p1, p2):
    if p1 is None:
        return p2
    if p2 is None:
        return p1
    (x1, y1), (x2, y2) = p1, p2
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, P - 2, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, P - 2, P) % P
    x3 = (lam * lam - x1 - x2) % P
    return x3, (lam * (x1 - x3) - y1) % P


def ec_mul(pt, k):
    acc = None
    while k:
        if k & 1:
            acc = ec_add(acc, pt)
        pt = ec_add(pt, pt)
        k >>= 1
    return acc


def decompress_g1(hexs):
    b = bytearray(bytes.fromhex(hexs))
    sign = (b[0] >> 5) & 1
    b[0] &= 0x1F
    x = int.from_bytes(b, "big")
    y = pow((x ** 3 + 4) % P, (P + 1) // 4, P)
    assert y * y % P == (x ** 3 + 4) % P
    if (y > (P - 1) // 2) != bool(sign):
        y = P - y
    return x, y


def sliding_window(e, w):
    """Left-to-right sliding-window schedule of e: first digit, then (squarings,
    odd digit) steps, then trailing squarings."""
    bits = bin(e)[2:]
    steps, i, pending = [], 0, 0
    while i < len(bits):
        if bits[i] == "0":
            pending += 1
            i += 1
            continue
        j = min(i + w, len(bits))
        while bits[j - 1] == "0":
            j -= 1
        steps.append((pending + (j - i) if steps else 0, int(bits[i:j], 2)))
        pending, i = 0, j
    v = 0
    for sq, d in steps:
        v = (v << sq) + d
    assert v << pending == e
    return steps, pending


def main():
    out = []
    emit = out.append

    def arr(name, vals, n, typ="uint32_t"):
        flat = []
        for v in vals:
            flat += limbs28(v) if n == "fp" else limbs(v, n)
        body = ", ".join("0x%08xu" % w for w in flat)
        emit("static constexpr %s %s[%d] = {%s};" % (typ, name, len(flat), body))

    pinv = (-pow(P, -1, 1 << 32)) % (1 << 32)
    pinv28 = (-pow(P, -1, 1 << LB)) % (1 << LB)
    rinv = (-pow(R, -1, 1 << 32)) % (1 << 32)
    emit("// Generated by gen_consts.py from x = -0xd201000000010000. Do not edit.")
    emit("#pragma once")
    emit("#include <stdint.h>")
    emit("namespace bls {")
    emit("// Fp elements: %d limbs of %d bits, Montgomery radix R = 2^%d" % (NL, LB, NL * LB))
    emit("static constexpr int FP_LIMBS = %d;" % NL)
    emit("static constexpr int FP_LIMB_BITS = %d;" % LB)
    arr("P28", [P], "fp")
    arr("P2_28", [2 * P], "fp")
    arr("P4_28", [4 * P], "fp")
    emit("static constexpr uint32_t P28_INV = 0x%08xu;" % pinv28)
    arr("P_R2", [RP * RP % P], "fp")
    arr("P_ONE", [mont(1)], "fp")
    # 2^384 R^2 mod p: mont_mul(hi, P_WIDE_HI) = Montgomery form of hi * 2^384
    arr("P_WIDE_HI", [(1 << 384) * RP * RP % P], "fp")
    # safegcd inversion (fp_inv): p in 13 signed 30-bit limbs, -p^-1 mod 2^30,
    # and R^3 mod p (fp_mul(x^-1, R^3) = Montgomery form of the inverse)
    emit("static constexpr int32_t P_S30[13] = {%s};" % ", ".join(
        "0x%08x" % ((P >> (30 * i)) & ((1 << 30) - 1)) for i in range(13)))
    emit("static constexpr uint32_t P_INV30 = 0x%08xu;" % ((-pow(P, -1, 1 << 30)) % (1 << 30)))
    arr("P_R3", [RP ** 3 % P], "fp")
    arr("P_MINUS_2", [P - 2], 12)
    arr("P_SQRT_EXP", [(P + 1) // 4], 12)
    # (p+1)/4 as a 4-bit sliding-window schedule for fp_sqrt
    steps, tail = sliding_window((P + 1) // 4, 4)
    emit("static constexpr int SQRT_W4_STEPS = %d;" % len(steps))
    emit("static constexpr int SQRT_W4_TAIL = %d;" % tail)
    emit("static constexpr uint8_t SQRT_W4_SQ[%d] = {%s};" % (len(steps), ", ".join(str(a) for a, _ in steps)))
    emit("static constexpr uint8_t SQRT_W4_DIG[%d] = {%s};" % (len(steps), ", ".join(str(b) for _, b in steps)))
    arr("P_LEGENDRE_EXP", [(P - 1) // 2], 12)
    arr("P_MINUS3_DIV4", [(P - 3) // 4], 12)
    arr("P_HALF", [(P - 1) // 2], "fp")  # canonical (not Montgomery), lexicographic sign
    arr("R_MOD", [R], 8)
    emit("static constexpr uint32_t R_INV = 0x%08xu;" % rinv)
    arr("R_R2", [RR * RR % R], 8)
    arr("R_ONE", [montr(1)], 8)
    arr("R_MINUS_2", [R - 2], 8)
    arr("FP_B", [mont(4)], "fp")
    # Frobenius: gamma1[k] = xi^(k(p-1)/6) in Fp2 ; gamma2[k] = xi^(k(p^2-1)/6) in Fp
    xi = (1, 1)
    g = f2pow(xi, (P - 1) // 6)
    gam = [(1, 0)]
    for k in range(1, 6):
        gam.append(f2mul(gam[-1], g))
    emit("// gamma1[k] = xi^(k (p-1)/6) as (c0, c1), Montgomery")
    arr("FROB1", [mont(c) for gk in gam for c in gk], "fp")
    g2 = f2pow(xi, (P * P - 1) // 6)
    gam2 = [(1, 0)]
    for k in range(1, 6):
        gam2.append(f2mul(gam2[-1], g2))
    assert all(c[1] == 0 for c in gam2)
    emit("// gamma2[k] = xi^(k (p^2-1)/6), an element of Fp")
    arr("FROB2", [mont(c[0]) for c in gam2], "fp")
    # G2 generator (affine, Montgomery): x = x0 + x1 u
    g2x0 = 0x024AA2B2F08F0A91260805272DC51051C6E47AD4FA403B02B4510B647AE3D1770BAC0326A805BBEFD48056C8C121BDB8
    g2x1 = 0x13E02B6052719F607DACD3A088274F65596BD0D09920B61AB5DA61BBDC7F5049334CF11213945D57E5AC7D055D042B7E
    g2y0 = 0x0CE5D527727D6E118CC9CDC6DA2E351AADFD9BAA8CBDD3A76D429A695160D12C923AC9CC3BACA289E193548608B82801
    g2y1 = 0x0606C4A02EA734CC32ACD2B02BC28B99CB3E287E85A763AF267492AB572E99AB3F370D275CEC1DA1AAA9075FF05F79BE
    # y^2 = x^3 + 4(1+u)
    x = (g2x0, g2x1)
    y = (g2y0, g2y1)
    lhs = f2mul(y, y)
    rhs = f2mul(x, f2mul(x, x))
    rhs = ((rhs[0] + 4) % P, (rhs[1] + 4) % P)
    assert lhs == rhs
    arr("G2_GEN", [mont(g2x0), mont(g2x1), mont(g2y0), mont(g2y1)], "fp")
    arr("G2_GEN_NEG", [mont(g2x0), mont(g2x1), mont((-g2y0) % P), mont((-g2y1) % P)], "fp")
    # G1 generator and BBS P1
    g1 = decompress_g1("97f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb")
    assert ec_mul(g1, R) is None
    arr("G1_GEN", [mont(g1[0]), mont(g1[1])], "fp")
    p1 = decompress_g1("a8ce256102840821a3e94ea9025e4662b205762f9776b3a766c872b948f1fd225e7c59698588e70d11406d161b4e28c9")
    arr("BBS_P1", [mont(p1[0]), mont(p1[1])], "fp")
    # GLV: beta (cube root of unity in Fp) with (beta x, y) = lambda P, lambda = x^2 - 1
    lam = (X * X - 1) % R
    assert (lam * lam + lam + 1) % R == 0
    beta = None
    for gg in range(2, 50):
        z = pow(gg, (P - 1) // 3, P)
        if z != 1:
            for cand in (z, z * z % P):
                if ec_mul(g1, lam) == (cand * g1[0] % P, g1[1]):
                    beta = cand
            break
    assert beta is not None
    arr("GLV_BETA", [mont(beta)], "fp")
    arr("GLV_LAMBDA", [lam], 8)
    # Barrett constant for k div lambda, k < 2^256: mu = floor(2^256 / lambda)
    mu = (1 << 256) // lam
    arr("GLV_MU", [mu], 8)
    emit("static constexpr int GLV_MU_BITS = %d;" % mu.bit_length())
    # hash_to_curve: SSWU Z, A', B' and isogeny coefficients
    iso = open(os.path.join(HERE, "h2c_iso_consts.h")).read()
    A_iso = int(re.search(r'H2C_ISO_A "([0-9a-f]+)"', iso).group(1), 16)
    B_iso = int(re.search(r'H2C_ISO_B "([0-9a-f]+)"', iso).group(1), 16)
    arr("SSWU_A", [mont(A_iso)], "fp")
    arr("SSWU_B", [mont(B_iso)], "fp")
    arr("SSWU_Z", [mont(11)], "fp")
    # -B/A and B/(Z A)
    arr("SSWU_MINUS_B_OVER_A", [mont((-B_iso) * pow(A_iso, P - 2, P) % P)], "fp")
    arr("SSWU_B_OVER_ZA", [mont(B_iso * pow(11 * A_iso, P - 2, P) % P)], "fp")
    for name in ("XNUM", "XDEN", "YNUM", "YDEN"):
        block = re.search(r"H2C_ISO_%s\[(\d+)\] = \{([^}]*)\}" % name, iso, re.S)
        vals = [int(v, 16) for v in re.findall(r'"([0-9a-f]+)"', block.group(2))]
        arr("ISO_" + name, [mont(v) for v in vals], "fp")
        emit("static constexpr int ISO_%s_LEN = %d;" % (name, len(vals)))
    emit("}  // namespace bls")
    path = os.path.join(HERE, "bls12_381_consts.hpp")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
