#!/usr/bin/env python3
"""Generate ``c2r_gen.hpp``: straight-line, fully unrolled complex-to-real
inverse DFTs (unnormalised, numpy/FFTW DC and Nyquist convention: the
imaginary parts of X_0 and X_H are ignored) for every profile length
NB = 2M with M <= 64 whose prime factors are all <= 7.

Algorithm (all sizes):
  1. pack the H+1 = M+1 half-spectrum bins into M complex values
       Z_k = E_k + i O_k,  E_k = X_k + conj(X_{M-k}),
       O_k = (X_k - conj(X_{M-k})) exp(+2 pi i k / NB)
     (X_0 and X_M enter through their real parts only);
  2. z = sum_k Z_k exp(+2 pi i k n / M), an M-point mixed-radix
     decimation-in-time transform, radices 4, 2, 3, 5, 7;
  3. out[2n] = Re z_n, out[2n+1] = Im z_n.
Twiddles are float literals rounded once from exact double values.  The emitted
functions compile as host and device code (CS_HD), so the same source is
checked on the CPU against numpy.fft.irfft.

Usage: python gen_c2r.py > c2r_gen.hpp
"""
import math
import sys


def smooth7(n):
    for f in (2, 3, 5, 7):
        while n % f == 0:
            n //= f
    return n == 1


SIZES = [2 * m for m in range(1, 65) if smooth7(m)]


def flit(v):
    if v == 0.0:
        return "0.0f"
    return repr(float(v)) + "f" if "e" in repr(float(v)) or "." in repr(float(v)) else repr(float(v)) + ".0f"


class Gen:
    def __init__(self):
        self.lines = []
        self.n = 0

    def tmp(self, expr):
        name = f"t{self.n}"
        self.n += 1
        self.lines.append(f"  const float {name} = {expr};")
        return name

    def add(self, a, b):
        return (self.tmp(f"{a[0]} + {b[0]}"), self.tmp(f"{a[1]} + {b[1]}"))

    def sub(self, a, b):
        return (self.tmp(f"{a[0]} - {b[0]}"), self.tmp(f"{a[1]} - {b[1]}"))

    def mulc(self, a, c, s):
        """a * (c + i s) with exact simplification of 0, +-1 components."""
        if abs(s) < 1e-300 and abs(c - 1.0) < 1e-300:
            return a
        if abs(s) < 1e-300 and abs(c + 1.0) < 1e-300:
            return (self.tmp(f"-{a[0]}"), self.tmp(f"-{a[1]}"))
        if abs(c) < 1e-300 and abs(s - 1.0) < 1e-300:
            return (self.tmp(f"-{a[1]}"), a[0])
        if abs(c) < 1e-300 and abs(s + 1.0) < 1e-300:
            return (a[1], self.tmp(f"-{a[0]}"))
        cf, sf = flit(c), flit(s)
        re = self.tmp(f"{a[0]} * {cf} - {a[1]} * {sf}")
        im = self.tmp(f"{a[0]} * {sf} + {a[1]} * {cf}")
        return (re, im)


def tw(e, n):
    """exp(+2 pi i e / n) from exact rational reduction."""
    e %= n
    if e == 0:
        return 1.0, 0.0
    if 4 * e == n:
        return 0.0, 1.0
    if 2 * e == n:
        return -1.0, 0.0
    if 4 * e == 3 * n:
        return 0.0, -1.0
    a = 2.0 * math.pi * e / n
    return math.cos(a), math.sin(a)


def dft_small(g, xs):
    """Direct r-point inverse DFT of complex values xs (r = 2, 3, 4, 5, 7)."""
    r = len(xs)
    if r == 1:
        return xs
    if r == 2:
        return [g.add(xs[0], xs[1]), g.sub(xs[0], xs[1])]
    if r == 4:
        a = g.add(xs[0], xs[2])
        b = g.sub(xs[0], xs[2])
        c = g.add(xs[1], xs[3])
        d = g.sub(xs[1], xs[3])
        di = g.mulc(d, 0.0, 1.0)          # +i (inverse transform)
        return [g.add(a, c), g.add(b, di), g.sub(a, c), g.sub(b, di)]
    out = []
    for u in range(r):
        acc = xs[0]
        for q in range(1, r):
            c, s = tw(q * u, r)
            acc = g.add(acc, g.mulc(xs[q], c, s))
        out.append(acc)
    return out


def fft_rec(g, xs):
    n = len(xs)
    if n == 1:
        return xs
    r = None
    for f in (4, 2, 3, 5, 7):
        if n % f == 0:
            r = f
            break
    if r is None:
        r = n
    m = n // r
    subs = [fft_rec(g, xs[q::r]) for q in range(r)]
    out = [None] * n
    for k in range(m):
        t = []
        for q in range(r):
            c, s = tw(q * k, n)
            t.append(g.mulc(subs[q][k], c, s))
        y = dft_small(g, t)
        for u in range(r):
            out[k + m * u] = y[u]
    return out


def gen_size(NB):
    M = NB // 2
    g = Gen()
    Z = []
    for k in range(M):
        if k == 0:
            a = ("xr[0]", "0.0f")
            b = (f"xr[{M}]", "0.0f")
        else:
            a = (f"xr[{k}]", f"xi[{k}]")
            b = (f"xr[{M - k}]", f"(-xi[{M - k}])")
        E = g.add(a, b)
        D = g.sub(a, b)
        c, s = tw(k, NB)
        O = g.mulc(D, c, s)
        Z.append((g.tmp(f"{E[0]} - {O[1]}"), g.tmp(f"{E[1]} + {O[0]}")))
    z = fft_rec(g, Z)
    body = list(g.lines)
    for nn in range(M):
        body.append(f"  out[{2 * nn}] = {z[nn][0]};")
        body.append(f"  out[{2 * nn + 1}] = {z[nn][1]};")
    return body


def main():
    o = []
    o.append("// GENERATED by gen_c2r.py -- do not edit.  Unnormalised c2r transforms")
    o.append("// (numpy DC/Nyquist convention) for profile lengths NB:")
    o.append("// " + " ".join(str(s) for s in SIZES))
    o.append("#pragma once")
    o.append("#ifndef CS_HD")
    o.append("#define CS_HD")
    o.append("#endif")
    o.append("namespace cs_c2r {")
    o.append("template <int NB> struct Fixed { static constexpr bool available = false; };")
    for NB in SIZES:
        body = gen_size(NB)
        o.append(f"template <> struct Fixed<{NB}> {{")
        o.append("  static constexpr bool available = true;")
        o.append("  static CS_HD inline void run(const float* __restrict__ xr, "
                 "const float* __restrict__ xi, float* __restrict__ out) {")
        o.extend(["  " + ln for ln in body])
        o.append("  }")
        o.append("};")
    o.append("#define CS_C2R_SIZES(X) " + " ".join(f"X({s})" for s in SIZES))
    o.append("}  // namespace cs_c2r")
    sys.stdout.write("\n".join(o) + "\n")


if __name__ == "__main__":
    main()
