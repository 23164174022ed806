#!/usr/bin/env python3
"""Derive the BLS12-381 G1 hash-to-curve constants from the curve parameter.

Test infrastructure (oracle side).  Nothing here is shipped in the product.

The BBS generators (`crates/cubbs-bbs-protocol/src/bbsplus/generators.rs:119-148`)
are RFC 9380 `hash_to_curve` outputs (suite BLS12381G1_XMD:SHA-256_SSWU_RO_).
That suite maps a field element to the 11-isogenous curve
    E': y^2 = x^3 + A' x + B'
with simplified SWU (Z = 11) and then applies an 11-isogeny E' -> E.

Instead of transcribing the 50-odd isogeny coefficients, this script:
  1. derives p and r from the BLS parameter x = -0xd201000000010000;
  2. checks that E' (A', B' as published in RFC 9380 section 8.8.1) has the
     same group order as E: y^2 = x^3 + 4 (isogenous curves over Fp do);
  3. factors the 11-division polynomial of E', builds every rational
     11-isogeny with Kohel/Velu formulas, keeps those landing on a j = 0
     curve, and composes them with the isomorphisms onto E;
  4. selects the unique candidate whose full hash_to_curve reproduces the
     CFRG fixture generators Q1, H_1..H_10 (tests/golden/cfrg/generators.json);
  5. writes the coefficients to a C header (as hex strings) for the oracle
     and for the HIP library.

Run:  python3 oracle/derive_h2c_iso.py [--out-oracle F] [--out-product F]
"""
import argparse
import hashlib
import json
import os
import random
import sys

X_BLS = -0xD201000000010000
P = ((X_BLS - 1) ** 2 * (X_BLS ** 4 - X_BLS ** 2 + 1)) // 3 + X_BLS
R = X_BLS ** 4 - X_BLS ** 2 + 1
# RFC 9380 section 8.8.1 (BLS12-381 G1 suite): E' coefficients and Z.
A_ISO = 0x144698A3B8E9433D693A02C96D4982B0EA985383EE66A8D8E8981AEFD881AC98936F8DA0E0F97F5CF428082D584C1D
B_ISO = 0x12E2908D11688030018B12E8753EEE3B2016C1F0F24F4070A0B9C14FCEF35EF55A23215A316CEAA5D1CC48E98E172BE0
Z_SSWU = 11
B_E = 4

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def is_probable_prime(n, rounds=32):
    if n < 2:
        return False
    for q in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29):
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    rng = random.Random(1)
    for _ in range(rounds):
        a = rng.randrange(2, n - 1)
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


# ---------------------------------------------------------------- Fp[x] utils
def pnorm(a):
    while a and a[-1] == 0:
        a.pop()
    return a


def padd(a, b):
    n = max(len(a), len(b))
    return pnorm([((a[i] if i < len(a) else 0) + (b[i] if i < len(b) else 0)) % P for i in range(n)])


def psub(a, b):
    n = max(len(a), len(b))
    return pnorm([((a[i] if i < len(a) else 0) - (b[i] if i < len(b) else 0)) % P for i in range(n)])


def pscale(a, c):
    return pnorm([x * c % P for x in a])


def pmul(a, b):
    if not a or not b:
        return []
    out = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                out[i + j] += x * y
    return pnorm([v % P for v in out])


def pdivmod(a, b):
    a = list(a)
    b = pnorm(list(b))
    inv = pow(b[-1], P - 2, P)
    q = [0] * max(0, len(a) - len(b) + 1)
    while len(pnorm(a)) >= len(b):
        c = a[-1] * inv % P
        k = len(a) - len(b)
        q[k] = c
        for i, y in enumerate(b):
            a[k + i] = (a[k + i] - c * y) % P
        pnorm(a)
    return pnorm(q), pnorm(a)


def pmod(a, b):
    return pdivmod(a, b)[1]


def pgcd(a, b):
    a, b = pnorm(list(a)), pnorm(list(b))
    while b:
        a, b = b, pmod(a, b)
    if a:
        a = pscale(a, pow(a[-1], P - 2, P))
    return a


def pderiv(a):
    return pnorm([(i * a[i]) % P for i in range(1, len(a))])


def ppowmod(base, e, m):
    result = [1]
    base = pmod(base, m)
    while e:
        if e & 1:
            result = pmod(pmul(result, base), m)
        base = pmod(pmul(base, base), m)
        e >>= 1
    return result


def pcompose_mod(g, h, m):
    """g(h(x)) mod m by Horner."""
    out = []
    for c in reversed(g):
        out = padd(pmod(pmul(out, h), m), [c])
    return out


def division_poly_reduced(n, a, b):
    """f_n with psi_n = f_n (n odd), psi_n = 2y f_n (n even); y^2 = F."""
    F = [b, a, 0, 1]
    F2x16 = pscale(pmul(F, F), 16)
    f = {0: [], 1: [1], 2: [1]}
    f[3] = pnorm([(-a * a) % P, 12 * b % P, 6 * a % P, 0, 3])
    f[4] = pscale(pnorm([(-8 * b * b - a ** 3) % P, (-4 * a * b) % P, (-5 * a * a) % P, 20 * b % P, 5 * a % P, 0, 1]), 2)
    for k in range(5, n + 1):
        m = k // 2
        if k % 2 == 1:
            t1 = pmul(f[m + 2], pmul(f[m], pmul(f[m], f[m])))
            t2 = pmul(f[m - 1], pmul(f[m + 1], pmul(f[m + 1], f[m + 1])))
            if m % 2 == 0:
                f[k] = psub(pmul(F2x16, t1), t2)
            else:
                f[k] = psub(t1, pmul(F2x16, t2))
        else:
            t1 = pmul(f[m + 2], pmul(f[m - 1], f[m - 1]))
            t2 = pmul(f[m - 2], pmul(f[m + 1], f[m + 1]))
            f[k] = pmul(f[m], psub(t1, t2))
    return f[n]


def equal_degree_split(g, d, rng):
    """Split a product of distinct degree-d irreducibles into its factors."""
    if len(g) - 1 == d:
        return [g]
    xp = ppowmod([0, 1], P, g)
    while True:
        a = [rng.randrange(P) for _ in range(len(g) - 1)]
        a = pnorm(a)
        if not a:
            continue
        # trace map T(a) = a + a^p + ... + a^(p^(d-1)) lands in Fp (mod each factor)
        t = list(a)
        cur = list(a)
        for _ in range(d - 1):
            cur = pcompose_mod(cur, xp, g)
            t = padd(t, cur)
        s = ppowmod(t, (P - 1) // 2, g)
        s = psub(s, [1])
        h = pgcd(g, s)
        if 0 < len(h) - 1 < len(g) - 1:
            q, rem = pdivmod(g, h)
            assert not rem
            return equal_degree_split(h, d, rng) + equal_degree_split(pscale(q, pow(q[-1], P - 2, P)), d, rng)


def kohel_isogeny(h, a, b):
    """Velu isogeny with kernel polynomial h (degree 5) from y^2=x^3+ax+b.

    Returns (A, B, xnum, xden, ynum, yden) with X = xnum/xden, Y = y*ynum/yden.
    """
    hp = pderiv(h)
    vpoly = [2 * a % P, 0, 6]                      # 6x^2 + 2a
    upoly = pscale([b, a, 0, 1], 4)                # 4(x^3 + ax + b)
    R1 = pmod(pmul(vpoly, hp), h)
    R2 = pmod(pmul(upoly, hp), h)
    # X = [x h^2 + R1 h + R2 h' - R2' h] / h^2
    h2 = pmul(h, h)
    num = padd(padd(pmul([0, 1], h2), pmul(R1, h)), psub(pmul(R2, hp), pmul(pderiv(R2), h)))
    # Y = y * (N' h - 2 N h') / h^3
    ynum = psub(pmul(pderiv(num), h), pscale(pmul(num, hp), 2))
    yden = pmul(h2, h)
    # codomain: power sums of the roots via Newton identities
    n = len(h) - 1
    e = [1] + [((-1) ** k) * h[n - k] % P for k in range(1, n + 1)]  # elementary symmetric
    p1 = e[1]
    p2 = (e[1] * p1 - 2 * e[2]) % P
    p3 = (e[1] * p2 - e[2] * p1 + 3 * e[3]) % P
    v = (6 * p2 + 2 * a * n) % P
    w = (10 * p3 + 6 * a * p1 + 4 * b * n) % P
    A = (a - 5 * v) % P
    B = (b - 7 * w) % P
    return A, B, num, h2, ynum, yden


def fp_sqrt(a):
    s = pow(a, (P + 1) // 4, P)
    return s if s * s % P == a % P else None


def cube_roots_of_unity():
    for g in range(2, 100):
        z = pow(g, (P - 1) // 3, P)
        if z != 1:
            return [1, z, z * z % P]
    raise RuntimeError("no cube root of unity")


def sixth_roots(c):
    """All u in Fp with u^6 = c."""
    # find one root by exponentiating in the subgroup structure (brute: use Tonelli-like via pow)
    out = []
    zs = cube_roots_of_unity()
    s = fp_sqrt(c)
    if s is None:
        return out
    for sq in (s, (-s) % P):
        # cube roots of sq
        # p = 1 mod 3: cube root exists iff sq^((p-1)/3) == 1; find via exponent search
        if pow(sq, (P - 1) // 3, P) != 1:
            continue
        # Adleman-Manders-Miller is overkill; p-1 = 3^k * m with small k
        m, k = P - 1, 0
        while m % 3 == 0:
            m //= 3
            k += 1
        # find non-cubic residue
        for g in range(2, 1000):
            if pow(g, (P - 1) // 3, P) != 1:
                break
        # brute-force over the 3-Sylow part
        inv3 = pow(3, -1, m) if m % 3 else None
        cand = pow(sq, inv3, P) if inv3 is not None else None
        # cand^3 = sq * (unit in 3-Sylow); fix up with 3-Sylow search
        gs = pow(g, m, P)
        sylow = [pow(gs, i, P) for i in range(3 ** k)]
        for t in sylow:
            u = cand * t % P
            if pow(u, 3, P) == sq:
                for zz in zs:
                    out.append(u * zz % P)
                break
    return sorted(set(out))


# ---------------------------------------------------------------- hash_to_curve (Python, for selection)
def expand_message_xmd(msg, dst, n):
    ell = (n + 31) // 32
    dst_prime = dst + bytes([len(dst)])
    b0 = hashlib.sha256(b"\x00" * 64 + msg + n.to_bytes(2, "big") + b"\x00" + dst_prime).digest()
    bs = [hashlib.sha256(b0 + b"\x01" + dst_prime).digest()]
    for i in range(2, ell + 1):
        prev = bytes(x ^ y for x, y in zip(b0, bs[-1]))
        bs.append(hashlib.sha256(prev + bytes([i]) + dst_prime).digest())
    return b"".join(bs)[:n]


def sswu(u):
    A, B, Z = A_ISO, B_ISO, Z_SSWU
    tv1 = (Z * Z * pow(u, 4, P) + Z * u * u) % P
    tv1 = pow(tv1, P - 2, P) if tv1 else 0
    if tv1 == 0:
        x1 = B * pow(Z * A % P, P - 2, P) % P
    else:
        x1 = (-B) * pow(A, P - 2, P) * (1 + tv1) % P
    gx1 = (x1 ** 3 + A * x1 + B) % P
    y1 = fp_sqrt(gx1)
    if y1 is not None:
        x, y = x1, y1
    else:
        x = Z * u * u * x1 % P
        y = fp_sqrt((x ** 3 + A * x + B) % P)
    if (u % 2) != (y % 2):
        y = (-y) % P
    return x, y


def peval(c, x):
    acc = 0
    for v in reversed(c):
        acc = (acc * x + v) % P
    return acc


def ec_add(p1, p2, a=0):
    if p1 is None:
        return p2
    if p2 is None:
        return p1
    x1, y1 = p1
    x2, y2 = p2
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = (3 * x1 * x1 + a) * pow(2 * y1, P - 2, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, P - 2, P) % P
    x3 = (lam * lam - x1 - x2) % P
    return x3, (lam * (x1 - x3) - y1) % P


def ec_mul(pt, k, a=0):
    acc = None
    if k < 0:
        k = -k
        pt = (pt[0], (-pt[1]) % P)
    while k:
        if k & 1:
            acc = ec_add(acc, pt, a)
        pt = ec_add(pt, pt, a)
        k >>= 1
    return acc


def compress_g1(pt):
    if pt is None:
        return bytes([0xC0]) + bytes(47)
    x, y = pt
    b = bytearray(x.to_bytes(48, "big"))
    b[0] |= 0x80
    if y > (P - 1) // 2:
        b[0] |= 0x20
    return bytes(b)


def hash_to_curve(msg, dst, iso):
    xn, xd, yn, yd = iso
    ub = expand_message_xmd(msg, dst, 128)
    pts = []
    for i in range(2):
        u = int.from_bytes(ub[64 * i:64 * i + 64], "big") % P
        x, y = sswu(u)
        X = peval(xn, x) * pow(peval(xd, x), P - 2, P) % P
        Y = y * peval(yn, x) % P * pow(peval(yd, x), P - 2, P) % P
        assert (Y * Y - X ** 3 - B_E) % P == 0
        pts.append((X, Y))
    return ec_mul(ec_add(pts[0], pts[1]), 1 - X_BLS)


def bbs_generators(count, api_id):
    seed_dst = api_id + b"SIG_GENERATOR_SEED_"
    gen_dst = api_id + b"SIG_GENERATOR_DST_"
    v = expand_message_xmd(api_id + b"MESSAGE_GENERATOR_SEED", seed_dst, 48)
    out = []
    for i in range(1, count + 1):
        v = expand_message_xmd(v + i.to_bytes(8, "big"), seed_dst, 48)
        out.append(v)
    return out, gen_dst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out-oracle", default=os.path.join(HERE, "h2c_iso_consts.h"))
    ap.add_argument("--out-product", default=os.path.join(REPO, "cubbs-cred_amd", "csrc", "h2c_iso_consts.h"))
    ap.add_argument("--fixtures", default=os.path.join(REPO, "tests", "golden", "cfrg", "generators.json"))
    args = ap.parse_args()

    assert P.bit_length() == 381 and R.bit_length() == 255
    assert is_probable_prime(P) and is_probable_prime(R)
    order = P + 1 - (X_BLS + 1)
    assert order % R == 0
    rng = random.Random(2023)
    # E' is isogenous to E: same point count.
    for _ in range(4):
        while True:
            x = rng.randrange(P)
            y = fp_sqrt((x ** 3 + A_ISO * x + B_ISO) % P)
            if y is not None:
                break
        assert ec_mul((x, y), order, A_ISO) is None, "E' order mismatch"
    print("E' order check passed", file=sys.stderr)

    f11 = division_poly_reduced(11, A_ISO, B_ISO)
    assert len(f11) - 1 == 60
    f11 = pscale(f11, pow(f11[-1], P - 2, P))
    xp = ppowmod([0, 1], P, f11)
    # gcd with x^(p^5) - x  -> product of degree 1 and degree 5 irreducibles
    cur = xp
    for _ in range(4):
        cur = pcompose_mod(cur, xp, f11)
    g15 = pgcd(f11, psub(cur, [0, 1]))
    g1 = pgcd(f11, psub(xp, [0, 1]))
    g5 = pdivmod(g15, g1)[0] if g1 and len(g1) > 1 else g15
    print("deg linear part", len(g1) - 1, "deg quintic part", len(g5) - 1, file=sys.stderr)
    kernels = equal_degree_split(g5, 5, rng) if len(g5) > 1 else []
    if len(g1) - 1 == 5:
        # all five kernel x-coordinates are rational: Frobenius acts as +-1
        kernels.append(g1)
    fixtures = json.load(open(args.fixtures))
    api_id = b"BBS_BLS12381G1_XMD:SHA-256_SSWU_RO_H2G_HM2S_"
    seeds, gen_dst = bbs_generators(len(fixtures["MsgGenerators"]) + 1, api_id)
    expect = [fixtures["Q1"]] + fixtures["MsgGenerators"]

    found = None
    for h in kernels:
        A, B, xn, xd, yn, yd = kohel_isogeny(h, A_ISO, B_ISO)
        if A != 0:
            continue
        for u in sixth_roots(B_E * pow(B, P - 2, P) % P):
            u2, u3 = u * u % P, pow(u, 3, P)
            iso = (pscale(xn, u2), xd, pscale(yn, u3), yd)
            pt = hash_to_curve(seeds[0], gen_dst, iso)
            if compress_g1(pt).hex() == expect[0]:
                found = iso
                break
        if found:
            break
    assert found, "no isogeny candidate reproduces Q1"
    for i, s in enumerate(seeds):
        assert compress_g1(hash_to_curve(s, gen_dst, found)).hex() == expect[i], i
    print("isogeny reproduces Q1 and H_1..H_%d" % (len(seeds) - 1), file=sys.stderr)

    xn, xd, yn, yd = found
    # normalise denominators to be monic (RFC 9380 convention) by scaling num+den together
    cx = pow(xd[-1], P - 2, P)
    xn, xd = pscale(xn, cx), pscale(xd, cx)
    cy = pow(yd[-1], P - 2, P)
    yn, yd = pscale(yn, cy), pscale(yd, cy)
    assert xd[-1] == 1 and yd[-1] == 1

    def arr(name, c):
        body = ",\n".join('  "%096x"' % v for v in c)
        return "static const char *const %s[%d] = {\n%s\n};\n" % (name, len(c), body)

    text = ("/* Generated by oracle/derive_h2c_iso.py from x = -0xd201000000010000 and the\n"
            " * published E' coefficients of the BLS12381G1_XMD:SHA-256_SSWU_RO_ suite.\n"
            " * 11-isogeny E' -> E, coefficients low degree first, big-endian hex. */\n"
            "#pragma once\n"
            '#define H2C_ISO_A "%096x"\n#define H2C_ISO_B "%096x"\n#define H2C_SSWU_Z %d\n'
            % (A_ISO, B_ISO, Z_SSWU)
            + arr("H2C_ISO_XNUM", xn) + arr("H2C_ISO_XDEN", xd)
            + arr("H2C_ISO_YNUM", yn) + arr("H2C_ISO_YDEN", yd))
    for path in (args.out_oracle, args.out_product):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as fh:
            fh.write(text)
        print("wrote", path, file=sys.stderr)


if __name__ == "__main__":
    main()
