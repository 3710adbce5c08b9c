"""Generates aes_bs_sbox.h: the AES S-box without its affine constant,
x -> A * x^-1 in GF(2^8), as a straight-line Boolean circuit on bit-slice
words, for the bitsliced AES-128 (aes_bs.hip).

Derivation (FIPS-197 S-box, tower-field inversion):
  * GF(2^8) = GF(2)[a]/(a^8 + a^4 + a^3 + a + 1) (the AES field);
  * GF(2^4) = GF(2)[z]/(z^4 + z + 1), GF((2^4)^2) = GF(2^4)[y]/(y^2 + y + lam)
    with y^2 + y + lam irreducible;
  * an isomorphism a -> beta (beta a root of the AES polynomial in the tower
    field), i.e. an 8x8 bit matrix M into the tower basis (bits 0-3: the y^0
    coordinate l, bits 4-7: the y coordinate h);
  * (h y + l)^-1 = (h N^-1) y + ((h + l) N^-1), N = lam h^2 + h l + l^2
    (the norm: (h y + l) times its conjugate h (y + 1) + l);
  * GF(2^4) products as AND-XOR sums (c_k = XOR_i a_i & (b z^i)_k), the
    GF(2^4) inverse from its algebraic normal form, the linear maps (M, the
    squarings, lam*, A M^-1) as XOR sums with greedy sharing of common pairs.
The S-box constant 0x63 is left out: the kernel adds it through the round
keys (MixColumns maps a column of equal bytes to itself, so the constant
passes unchanged into every round key after the first).

The circuit is checked here on all 256 inputs (bit-parallel over the truth
table) before the header is written; tests/test_aes_bs_host.py checks the
written header the same way after compiling it with g++.

    python3 cuot_amd/csrc/gen_aes_bs.py        # rewrites aes_bs_sbox.h
"""
import itertools
import os
import sys

AES_POLY = 0x11B


def gf8_mul(a, b):
    r = 0
    while b:
        if b & 1:
            r ^= a
        a <<= 1
        if a & 0x100:
            a ^= AES_POLY
        b >>= 1
    return r


def gf8_inv(a):
    if a == 0:
        return 0
    r = 1
    for _ in range(254):
        r = gf8_mul(r, a)
    return r


def affine(u):
    out = 0
    for i in range(8):
        bit = 0
        for s in (0, 4, 5, 6, 7):
            bit ^= (u >> ((i + s) % 8)) & 1
        out |= bit << i
    return out


SBOX = [affine(gf8_inv(x)) ^ 0x63 for x in range(256)]
SBOX_NC = [affine(gf8_inv(x)) for x in range(256)]  # what the circuit computes


def gf4_mul(a, b):
    r = 0
    for i in range(4):
        if (b >> i) & 1:
            r ^= a << i
    for d in (6, 5, 4):
        if (r >> d) & 1:
            r ^= 0b10011 << (d - 4)
    return r


def gf4_inv(a):
    for b in range(16):
        if gf4_mul(a, b) == 1:
            return b
    return 0


def find_lam():
    for lam in range(1, 16):
        if all(gf4_mul(y, y) ^ y ^ lam for y in range(16)):
            return lam
    raise RuntimeError("no irreducible y^2 + y + lam")


LAM = find_lam()


def t_mul(p, q):  # tower elements as (h << 4) | l
    ph, pl, qh, ql = p >> 4, p & 15, q >> 4, q & 15
    hh = gf4_mul(ph, qh)
    h = hh ^ gf4_mul(ph, ql) ^ gf4_mul(pl, qh)
    l = gf4_mul(hh, LAM) ^ gf4_mul(pl, ql)
    return (h << 4) | l


def t_pow(p, e):
    r = 1
    for _ in range(e):
        r = t_mul(r, p)
    return r


def find_beta():
    # roots of a^8 + a^4 + a^3 + a + 1 in the tower field
    roots = []
    for b in range(2, 256):
        v = t_pow(b, 8) ^ t_pow(b, 4) ^ t_pow(b, 3) ^ b ^ 1
        if v == 0:
            roots.append(b)
    return roots


def matrix_of(beta):
    """M[j] = mask of AES bits feeding tower bit j (tower = M * aes)."""
    cols = [t_pow(beta, i) for i in range(8)]
    return [sum(((cols[i] >> j) & 1) << i for i in range(8)) for j in range(8)]


def mat_apply(rows, x):
    return sum((bin(rows[j] & x).count("1") & 1) << j for j in range(len(rows)))


def mat_inv(rows):
    n = len(rows)
    a = [(rows[j], 1 << j) for j in range(n)]
    for c in range(n):
        p = next(r for r in range(c, n) if (a[r][0] >> c) & 1)
        a[c], a[p] = a[p], a[c]
        for r in range(n):
            if r != c and (a[r][0] >> c) & 1:
                a[r] = (a[r][0] ^ a[c][0], a[r][1] ^ a[c][1])
    return [a[j][1] for j in range(n)]


# ---- circuit builder ------------------------------------------------------------
class Circuit:
    def __init__(self, ninputs):
        self.ops = []  # (name, kind, args)
        self.n = 0
        self.inputs = [f"x{i}" for i in range(ninputs)]

    def new(self, kind, *args):
        name = f"t{self.n}"
        self.n += 1
        self.ops.append((name, kind, args))
        return name

    def xor(self, terms):
        """XOR of a list of signals with 3-input ops; returns a signal."""
        terms = list(terms)
        assert terms
        while len(terms) > 1:
            if len(terms) >= 3:
                terms = [self.new("x3", terms[0], terms[1], terms[2])] + terms[3:]
            else:
                terms = [self.new("x2", terms[0], terms[1])]
        return terms[0]

    def linear(self, rows, sig):
        """Outputs XOR_{i in rows[j]} sig[i]; shares common pairs greedily (Paar)."""
        rows = [set(i for i in range(len(sig)) if (r >> i) & 1) for r in rows]
        sig = list(sig)
        while True:
            best, cnt = None, 1
            for a, b in itertools.combinations(range(len(sig)), 2):
                c = sum(1 for r in rows if a in r and b in r)
                if c > cnt:
                    best, cnt = (a, b), c
            if best is None:
                break
            a, b = best
            sig.append(self.new("x2", sig[a], sig[b]))
            k = len(sig) - 1
            for r in rows:
                if a in r and b in r:
                    r -= {a, b}
                    r.add(k)
        out = []
        for r in rows:
            if not r:
                out.append("0")
            elif len(r) == 1:
                out.append(sig[next(iter(r))])
            else:
                out.append(self.xor([sig[i] for i in sorted(r)]))
        return out

    def andxor(self, pairs, acc=None):
        """acc ^ XOR_k (p_k & q_k), one AND-XOR op per pair."""
        for a, b in pairs:
            acc = self.new("and", a, b) if acc is None else self.new("ax", a, b, acc)
        return acc

    def gf4_mul(self, a, b_rows):
        """a * b in GF(2^4); b_rows[i] = the 4 signals of b * z^i."""
        return [self.andxor([(a[i], b_rows[i][k]) for i in range(4)]) for k in range(4)]

    def gf4_shifts(self, b):
        rows = [list(b)]
        for _ in range(3):
            p = rows[-1]
            # p * z with z^4 = z + 1: [p3, p0 ^ p3, p1, p2]
            rows.append([p[3], self.new("x2", p[0], p[3]), p[1], p[2]])
        return rows


def gf4_inv_anf():
    """ANF of the GF(2^4) inverse: per output bit, the set of monomials
    (frozensets of input indices; the empty set is the constant 1)."""
    out = []
    for k in range(4):
        coef = {}
        for mono in itertools.chain.from_iterable(itertools.combinations(range(4), r) for r in range(5)):
            # Moebius transform
            v = 0
            for sub in itertools.chain.from_iterable(itertools.combinations(mono, r) for r in range(len(mono) + 1)):
                x = sum(1 << i for i in sub)
                v ^= (gf4_inv(x) >> k) & 1
            if v:
                coef[frozenset(mono)] = 1
        out.append(set(coef))
    return out


def build(beta):
    M = matrix_of(beta)
    Minv = mat_inv(M)
    Arows = [sum(1 << ((i + s) % 8) for s in (0, 4, 5, 6, 7)) for i in range(8)]
    # output map: A * Minv (aes bit i of out from tower bits)
    Aout = []
    for i in range(8):
        r = 0
        for j in range(8):
            if (Arows[i] >> j) & 1:
                r ^= Minv[j]
        Aout.append(r)
    c = Circuit(8)
    t = c.linear(M, c.inputs)
    l, h = t[0:4], t[4:8]
    # N = lam h^2 + l^2 + h l: linear part from (l, h) bits, then the product on top
    lin_rows = []
    for k in range(4):
        r = 0
        for j in range(4):  # contribution of l_j and h_j to bit k
            if (gf4_mul(1 << j, 1 << j) >> k) & 1:
                r |= 1 << j
            if (gf4_mul(gf4_mul(1 << j, 1 << j), LAM) >> k) & 1:
                r |= 1 << (4 + j)
        lin_rows.append(r)
    lin = c.linear(lin_rows, l + h)
    lrows = c.gf4_shifts(l)
    N = [c.andxor([(h[i], lrows[i][k]) for i in range(4)], None if lin[k] == "0" else lin[k]) for k in range(4)]
    # N^-1 from the ANF: pair products shared
    anf = gf4_inv_anf()
    pairs = {}
    for s in anf:
        for mono in s:
            if len(mono) >= 2:
                a, b = sorted(mono)[:2]
                pairs.setdefault((a, b), None)
    for (a, b) in sorted(pairs):
        pairs[(a, b)] = c.new("and", N[a], N[b])
    Ni = []
    for k in range(4):
        lin_terms, prod_terms = [], []
        const = False
        for mono in sorted(anf[k], key=lambda m: (len(m), sorted(m))):
            ms = sorted(mono)
            if len(ms) == 0:
                const = True
            elif len(ms) == 1:
                lin_terms.append(N[ms[0]])
            elif len(ms) == 2:
                lin_terms.append(pairs[(ms[0], ms[1])])
            else:
                prod_terms.append((pairs[(ms[0], ms[1])], ms[2:]))
        acc = c.xor(lin_terms) if lin_terms else None
        for base, rest in prod_terms:
            if len(rest) == 1:
                acc = c.new("and", base, N[rest[0]]) if acc is None else c.new("ax", base, N[rest[0]], acc)
            else:  # degree 4: (ab)(cd)
                cd = pairs.get((rest[0], rest[1])) or c.new("and", N[rest[0]], N[rest[1]])
                acc = c.new("and", base, cd) if acc is None else c.new("ax", base, cd, acc)
        assert not const
        Ni.append(acc)
    nrows = c.gf4_shifts(Ni)
    hn = c.gf4_mul(h, nrows)
    hl = [c.new("x2", h[i], l[i]) for i in range(4)]
    ln = c.gf4_mul(hl, nrows)
    out = c.linear(Aout, ln + hn)
    return c, out


def evaluate(c, out):
    # bit-parallel over the 256 inputs: signal = 256-bit integer
    env = {}
    for i in range(8):
        env[f"x{i}"] = sum(((x >> i) & 1) << x for x in range(256))
    env["0"] = 0
    full = (1 << 256) - 1
    for name, kind, args in c.ops:
        v = [env[a] for a in args]
        if kind == "x2":
            r = v[0] ^ v[1]
        elif kind == "x3":
            r = v[0] ^ v[1] ^ v[2]
        elif kind == "and":
            r = v[0] & v[1]
        elif kind == "ax":
            r = (v[0] & v[1]) ^ v[2]
        else:
            raise ValueError(kind)
        env[name] = r & full
    res = [0] * 256
    for i, o in enumerate(out):
        w = env[o]
        for x in range(256):
            res[x] |= ((w >> x) & 1) << i
    return res


def emit(c, out, path):
    counts = {}
    for _, kind, _ in c.ops:
        counts[kind] = counts.get(kind, 0) + 1
    lines = [
        "// aes_bs_sbox.h -- generated by gen_aes_bs.py (do not edit): the AES S-box",
        "// without its constant 0x63, x -> A * x^-1 over GF(2^8), on bit-slice words",
        "// (bit i of a byte across 32 blocks; x0 = least significant bit).  Tower-field",
        f"// inversion over GF(2^4)[y]/(y^2 + y + {LAM}), GF(2^4) = GF(2)[z]/(z^4 + z + 1);",
        f"// {len(c.ops)} operations: " + ", ".join(f"{v} {k}" for k, v in sorted(counts.items())) + ".",
        "// Checked on all 256 inputs by the generator and by tests/test_aes_bs_host.py.",
        "#pragma once",
        "",
        "// x2: a ^ b; x3: a ^ b ^ c; and: a & b; ax: (a & b) ^ c",
        "#define AES_BS_SBOX(x0, x1, x2, x3, x4, x5, x6, x7, y0, y1, y2, y3, y4, y5, y6, y7) \\",
        "  do { \\",
    ]
    for name, kind, args in c.ops:
        if kind == "x2":
            e = f"{args[0]} ^ {args[1]}"
        elif kind == "x3":
            e = f"BS_X3({args[0]}, {args[1]}, {args[2]})"
        elif kind == "and":
            e = f"{args[0]} & {args[1]}"
        else:
            e = f"({args[0]} & {args[1]}) ^ {args[2]}"
        lines.append(f"    const uint32_t {name} = {e}; \\")
    for i, o in enumerate(out):
        lines.append(f"    y{i} = {o}; \\")
    lines.append("  } while (0)")
    lines.append("")
    with open(path, "w") as f:
        f.write("\n".join(lines))
    return counts


def main():
    best = None
    for beta in find_beta():
        c, out = build(beta)
        if best is None or len(c.ops) < len(best[0].ops):
            best = (c, out, beta)
    c, out, beta = best
    got = evaluate(c, out)
    if got != SBOX_NC:
        print("circuit does not match the S-box", file=sys.stderr)
        return 1
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "aes_bs_sbox.h")
    counts = emit(c, out, path)
    print(f"beta={beta:#04x} lam={LAM} ops={len(c.ops)} {counts} -> {path}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
