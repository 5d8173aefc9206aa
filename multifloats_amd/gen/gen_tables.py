#!/usr/bin/env python3
"""Generate the constant tables used by the DD kernels (multifloats_amd/csrc/dd_tables.h).

Every table here is a closed form evaluated with mpmath at 400 bits and rounded
limb by limb (hi = RN(v), lo = RN(v - hi), third limb likewise). Nothing is
taken from another project's tables.

Tables that the reference's exp/log/sin algorithms need and whose values are
fixed by mathematics (so a bit-exact restatement must use the same values):
  - exp: ln2 Cody-Waite split with a 33-bit leading piece (reference
    src/float64x2/exp_log.inc:247-268), exp2 Taylor coefficients ln2^j/j!
    and the 257-entry table 2^(k/256) (exp_log.inc:26-72);
  - log: c_k = 2/((2k+1) ln2) and the 32 centres 1+(2i+1)/64 with log2 values
    (exp_log.inc:280-308);
  - sin: pi/2 three-piece split, Taylor coefficients of sin(x)/x and cos(x),
    pi/4 and 1/sqrt(2) (src/float64x2/trig.inc:114-187);
  - exp2/log10/expm1/log1p/sinh: log10(2), the exp2 clamps, and the Taylor
    coefficients 1/(k+1)!, 2/(2k+1) and 1/(2k+1)! (exp_log.inc:119-128,
    343-353, 407-469; hyp.inc:2-18);
  - asinh/atanh: the Taylor coefficients (-1)^k (2k)!/(4^k k!^2 (2k+1)) and
    1/(2k+1) of asinh(x)/x and atanh(x)/x in x^2 (inv_trig.inc:226-280).

The erf, erfc and erfcx tables are this project's own design (the reference fits erf with
rational approximations; see csrc/dd_math.h for the piecewise Taylor scheme):
  - row 0:   erf(x) = x + x*Q(x^2) for |x| < 1/4, Q(z) = (2/sqrt(pi) - 1)
             + sum_{n>=1} (2/sqrt(pi)) (-1)^n z^n / (n! (2n+1));
  - row i>0: Taylor expansion of erf about the centre c_i = 1/4 + (i - 1/2)/16,
             erf(c_i + t) = a0 + t*(a1 + a2 t + ...), |t| <= 1/32, with a0
             stored as three limbs;
  - erfc on [1/4, 4): Taylor rows of erfc about the same centres, all
             coefficients double-double (relative accuracy for small erfc);
  - erfcx = exp(x^2) erfc(x): Taylor rows about (i + 1/2)/16 on [0, 4),
             32 centres per binade on [4, 32) and -(i + 1/2)/16 on [-2, 0),
             coefficients from the ODE
             y' = 2xy - 2/sqrt(pi) at 450 bits; 1/sqrt(pi) for the
             asymptotic series beyond 32;
  - Ozaki CRT form: 40 pairwise-coprime moduli <= 256, 1/m in 180-bit
             fixed point, and the CRT inverses and products for every
             modulus count (exact integers).

Run:  python3 multifloats_amd/gen/gen_tables.py   (rewrites csrc/dd_tables.h)
"""
import os
import sys

import mpmath
from mpmath import mp, mpf

mp.prec = 400

ERF_ND = 10       # Horner coefficients kept as double-double
ERF_NT = 18       # total Horner coefficients (the rest are plain doubles)
ERF_X0 = mpf(1) / 4
ERF_H = mpf(1) / 16
ERF_XMAX = 9      # |x| >= 9 rounds to +-1 (erfc(9) < 2^-121)
ERF_ROWS = 1 + int((ERF_XMAX - ERF_X0) / ERF_H)

ERFC_NT = 22      # DD Taylor coefficients per erfc row on [1/4, 4)
ERFC_ROWS = 60    # centres 1/4 + (i - 1/2)/16, i = 1..60
ERFCX_NT = 20     # DD Taylor coefficients per erfcx row
ERFCX_ROWS = 192  # 64 rows of width 1/16 on [0, 4), 32 per binade on [4, 32), 32 on [-2, 0)
ERFCX_ASYM = 16   # terms of the asymptotic series for x >= 32


def limbs(v, n=2):
    out = []
    r = mpf(v)
    for _ in range(n):
        d = float(r)
        out.append(d)
        r = r - mpf(d)
    return out


def hx(d):
    if d == 0.0:
        return "0.0" if str(d)[0] != "-" else "-0.0"
    return float.hex(d)


def arr(name, vals, per_line=3):
    lines = []
    for i in range(0, len(vals), per_line):
        lines.append("    " + ", ".join(hx(v) for v in vals[i:i + per_line]) + ",")
    return "MF_TABLE double %s[%d] = {\n%s\n};\n" % (name, len(vals), "\n".join(lines))


def scalar(name, v):
    return "MF_SCALAR double %s = %s;\n" % (name, hx(v))


def dd_pair(name, v):
    h, l = limbs(v)
    return scalar(name + "_hi", h) + scalar(name + "_lo", l)


def main():
    ln2 = mpmath.log(2)
    pi = mpmath.pi
    out = []
    out.append("/* Generated by multifloats_amd/gen/gen_tables.py -- do not edit. */\n")
    out.append("#pragma once\n")
    out.append("#if defined(__HIPCC__)\n#define MF_TABLE static __constant__ const\n"
               "#define MF_SCALAR static constexpr\n#elif defined(__cplusplus)\n"
               "#define MF_TABLE static const\n#define MF_SCALAR static constexpr\n#else\n"
               "#define MF_TABLE static const\n#define MF_SCALAR static const\n#endif\n\n")

    out.append("/* log2(e), ln(2), pi/4, 1/sqrt(2) as double-doubles */\n")
    out.append(dd_pair("mf_log2e", 1 / ln2))
    out.append(dd_pair("mf_ln2", ln2))
    out.append(dd_pair("mf_pi_quarter", pi / 4))
    out.append(dd_pair("mf_inv_sqrt2", 1 / mpmath.sqrt(2)))

    # ln2 split: leading piece keeps the top 33 significand bits of RN(ln2).
    import struct
    ln2_d = float(ln2)
    bits = struct.unpack("<Q", struct.pack("<d", ln2_d))[0]
    cw1 = struct.unpack("<d", struct.pack("<Q", bits & ~((1 << 20) - 1)))[0]
    cw2 = float(ln2 - mpf(cw1))
    cw3 = float(ln2 - mpf(cw1) - mpf(cw2))
    out.append("\n/* ln2 = cw1 + cw2 + cw3; cw1 has a 33-bit significand so n*cw1 is exact */\n")
    out.append(scalar("mf_ln2_cw1", cw1) + scalar("mf_ln2_cw2", cw2) + scalar("mf_ln2_cw3", cw3))
    # exp clamps: the largest x below the overflow edge is RN(1024*ln2) rounded
    # down to the double 1024*RN(ln2); below -750 every result underflows to 0.
    out.append(scalar("mf_exp_max", 1024.0 * ln2_d))
    out.append(scalar("mf_exp_min", -750.0))

    ph = limbs(pi / 2, 3)
    out.append("\n/* pi/2 = cw1 + cw2 + cw3 (three round-to-nearest limbs) */\n")
    out.append(scalar("mf_pio2_cw1", ph[0]) + scalar("mf_pio2_cw2", ph[1]) + scalar("mf_pio2_cw3", ph[2]))

    c = [ln2 ** j / mpmath.factorial(j) for j in range(10)]
    out.append("\n/* 2^s - 1 = s*(c1 + c2 s + ... + c9 s^8), c_j = ln2^j / j! */\n")
    out.append(arr("mf_exp2c_hi", [limbs(v)[0] for v in c]))
    out.append(arr("mf_exp2c_lo", [limbs(v)[1] for v in c]))
    t = [mpf(2) ** (mpf(k - 128) / 256) for k in range(257)]
    out.append("\n/* T[k] = 2^((k-128)/256), k = 0..256 */\n")
    out.append(arr("mf_exp2t_hi", [limbs(v)[0] for v in t]))
    out.append(arr("mf_exp2t_lo", [limbs(v)[1] for v in t]))

    lc = [2 / ((2 * k + 1) * ln2) for k in range(11)]
    out.append("\n/* log2 series in t^2: c_k = 2/((2k+1) ln2) */\n")
    out.append(arr("mf_log2c_hi", [limbs(v)[0] for v in lc]))
    out.append(arr("mf_log2c_lo", [limbs(v)[1] for v in lc]))
    centres = [1.0 + (2 * i + 1) / 64.0 for i in range(32)]
    vals = [mpmath.log(mpf(x)) / ln2 for x in centres]
    out.append("\n/* centres 1 + (2i+1)/64 and log2(centre) */\n")
    out.append(arr("mf_log2_centre", centres))
    out.append(arr("mf_log2v_hi", [limbs(v)[0] for v in vals]))
    out.append(arr("mf_log2v_lo", [limbs(v)[1] for v in vals]))
    out.append("/* third limb of log2(centre), used by pow */\n")
    out.append(arr("mf_log2v_x", [limbs(v, 3)[2] for v in vals]))

    st = [(-1) ** k / mpmath.factorial(2 * k + 1) for k in range(13)]
    ct = [(-1) ** k / mpmath.factorial(2 * k) for k in range(13)]
    out.append("\n/* sin(x)/x and cos(x) Taylor coefficients in x^2 */\n")
    out.append(arr("mf_sinc_hi", [limbs(v)[0] for v in st]))
    out.append(arr("mf_sinc_lo", [limbs(v)[1] for v in st]))
    out.append(arr("mf_cosc_hi", [limbs(v)[0] for v in ct]))
    out.append(arr("mf_cosc_lo", [limbs(v)[1] for v in ct]))

    # ---- exp2 / log10 / expm1 / log1p / sinh ----
    out.append("\n/* log10(2); exp2 clamps of the reference (exp_log.inc:119-128): below\n"
               "   -1080 the result is 0, above 1024 - 2^-42 it is +inf */\n")
    out.append(dd_pair("mf_log10_2", mpmath.log10(2)))
    out.append(scalar("mf_exp2_min", -1080.0) + scalar("mf_exp2_max", 1024.0 - 2.0 ** -42))
    l10 = limbs(mpmath.log(10) / ln2, 3)
    out.append("\n/* log2(10) in three round-to-nearest limbs; ln(10) and 1/ln(10) */\n")
    out.append(scalar("mf_log2_10_hi", l10[0]) + scalar("mf_log2_10_lo", l10[1]) + scalar("mf_log2_10_lo2", l10[2]))
    out.append(dd_pair("mf_ln10", mpmath.log(10)))
    out.append("/* pi as a double-double (sinpi, cospi, tanpi) */\n")
    out.append(dd_pair("mf_pi", pi))
    out.append(dd_pair("mf_inv_pi", 1 / pi))
    ac = [mpf(-1) ** k / (2 * k + 1) for k in range(1, 18)]
    out.append("\n/* atan(u) = u + u^3 Q(u^2): Q's coefficients (-1)^k / (2k+1), k = 1..17 */\n")
    out.append(arr("mf_atanc_hi", [limbs(v)[0] for v in ac]))
    out.append(arr("mf_atanc_lo", [limbs(v)[1] for v in ac]))
    at = [limbs(mpmath.atan(mpf(k) / 64), 3) for k in range(8, 513)]
    out.append("\n/* atan(k/64), k = 8..512, in three round-to-nearest limbs */\n")
    out.append(arr("mf_atant", [v for t in at for v in t]))
    out.append(dd_pair("mf_inv_ln10", 1 / mpmath.log(10)))
    em = [1 / mpmath.factorial(k + 1) for k in range(25)]
    out.append("\n/* expm1(x)/x Taylor coefficients: 1/(k+1)!, k = 0..24 */\n")
    out.append(arr("mf_expm1c_hi", [limbs(v)[0] for v in em]))
    out.append(arr("mf_expm1c_lo", [limbs(v)[1] for v in em]))
    lp = [mpf(2) / (2 * k + 1) for k in range(18)]
    out.append("\n/* log1p(x) = t*P(t^2), t = x/(2+x): 2/(2k+1), k = 0..17 */\n")
    out.append(arr("mf_log1pc_hi", [limbs(v)[0] for v in lp]))
    out.append(arr("mf_log1pc_lo", [limbs(v)[1] for v in lp]))
    sh = [1 / mpmath.factorial(2 * k + 1) for k in range(13)]
    out.append("\n/* sinh(x)/x Taylor coefficients in x^2: 1/(2k+1)!, k = 0..12 */\n")
    out.append(arr("mf_sinhc_hi", [limbs(v)[0] for v in sh]))
    out.append(arr("mf_sinhc_lo", [limbs(v)[1] for v in sh]))
    ash = [(-1) ** k * mpmath.factorial(2 * k) / (4 ** k * mpmath.factorial(k) ** 2 * (2 * k + 1)) for k in range(15)]
    out.append("\n/* asinh(x)/x Taylor coefficients in x^2: (-1)^k (2k)! / (4^k k!^2 (2k+1)), k = 0..14 */\n")
    out.append(arr("mf_asinhc_hi", [limbs(v)[0] for v in ash]))
    out.append(arr("mf_asinhc_lo", [limbs(v)[1] for v in ash]))
    ath = [mpf(1) / (2 * k + 1) for k in range(15)]
    out.append("\n/* atanh(x)/x Taylor coefficients in x^2: 1/(2k+1), k = 0..14 */\n")
    out.append(arr("mf_atanhc_hi", [limbs(v)[0] for v in ath]))
    out.append(arr("mf_atanhc_lo", [limbs(v)[1] for v in ath]))

    # ---- erf ----
    two_rtpi = 2 / mpmath.sqrt(pi)
    rows_base = []      # three limbs of a0 (row 0: unused)
    rows_dd_hi, rows_dd_lo, rows_d = [], [], []
    centres_e = []
    for r in range(ERF_ROWS):
        if r == 0:
            p = [two_rtpi - 1] + [two_rtpi * (-1) ** n / (mpmath.factorial(n) * (2 * n + 1))
                                  for n in range(1, ERF_NT)]
            base = [0.0, 0.0, 0.0]
            centres_e.append(0.0)
        else:
            x0 = ERF_X0 + (r - mpf(1) / 2) * ERF_H
            assert float(x0) == x0
            g = two_rtpi * mpmath.exp(-x0 * x0)
            a = [g * (-1) ** (n - 1) * mpmath.hermite(n - 1, x0) / mpmath.factorial(n)
                 for n in range(1, ERF_NT + 1)]
            p = a
            base = limbs(mpmath.erf(x0), 3)
            centres_e.append(float(x0))
        rows_base.extend(base)
        for j in range(ERF_ND):
            h, l = limbs(p[j])
            rows_dd_hi.append(h)
            rows_dd_lo.append(l)
        for j in range(ERF_ND, ERF_NT):
            rows_d.append(float(p[j]))
    out.append("\n/* erf: row 0 is x + x*Q(x^2) for |x| < 1/4; row i is the Taylor expansion\n"
               "   about mf_erf_centre[i] = 1/4 + (i - 1/2)/16 for 1/4 <= |x| < 9 */\n")
    out.append("#define MF_ERF_ND %d\n#define MF_ERF_NT %d\n#define MF_ERF_ROWS %d\n" % (ERF_ND, ERF_NT, ERF_ROWS))
    out.append(scalar("mf_erf_x0", float(ERF_X0)) + scalar("mf_erf_inv_h", float(1 / ERF_H))
               + scalar("mf_erf_xmax", float(ERF_XMAX)))
    out.append(arr("mf_erf_centre", centres_e))
    out.append(arr("mf_erf_base", rows_base))
    out.append(arr("mf_erf_ddhi", rows_dd_hi))
    out.append(arr("mf_erf_ddlo", rows_dd_lo))
    out.append(arr("mf_erf_d", rows_d))

    # ---- erfc on [1/4, 4): Taylor rows of erfc itself (relative accuracy) ----
    rtpi = 2 / mpmath.sqrt(pi)
    ec_centre, ec_base, ec_hi, ec_lo = [], [], [], []
    for i in range(1, ERFC_ROWS + 1):
        c0 = ERF_X0 + (i - mpf(1) / 2) * ERF_H
        g = -rtpi * mpmath.exp(-c0 * c0)
        ec_centre.append(float(c0))
        ec_base.extend(limbs(mpmath.erfc(c0), 3))
        for n in range(1, ERFC_NT + 1):
            v = g * (-1) ** (n - 1) * mpmath.hermite(n - 1, c0) / mpmath.factorial(n)
            h, l = limbs(v)
            ec_hi.append(h)
            ec_lo.append(l)
    out.append("\n/* erfc(c + t) = base + t*(a1 + a2 t + ... + a%d t^%d) about c = 1/4 + (i - 1/2)/16,\n"
               "   i = 1..%d, |t| <= 1/32; a_n = erfc^(n)(c)/n! */\n" % (ERFC_NT, ERFC_NT - 1, ERFC_ROWS))
    out.append("#define MF_ERFC_NT %d\n#define MF_ERFC_ROWS %d\n" % (ERFC_NT, ERFC_ROWS))
    out.append(arr("mf_erfc_centre", ec_centre))
    out.append(arr("mf_erfc_base", ec_base))
    out.append(arr("mf_erfc_hi", ec_hi))
    out.append(arr("mf_erfc_lo", ec_lo))

    # ---- erfcx = exp(x^2) erfc(x): Taylor rows from y' = 2xy - 2/sqrt(pi) ----
    def erfcx_coeffs(c0, nmax):
        y = [mpmath.exp(c0 * c0) * mpmath.erfc(c0)]
        y.append(2 * c0 * y[0] - rtpi)
        for n in range(1, nmax):
            y.append((2 * c0 * y[n] + 2 * y[n - 1]) / (n + 1))
        return y
    centres_x = [(mpf(i) + mpf(1) / 2) / 16 for i in range(64)]
    for e in (2, 3, 4):
        w = mpf(2) ** e / 32
        centres_x += [mpf(2) ** e + (j + mpf(1) / 2) * w for j in range(32)]
    centres_x += [-(mpf(i) + mpf(1) / 2) / 16 for i in range(32)]
    assert len(centres_x) == ERFCX_ROWS
    ex_base, ex_hi, ex_lo = [], [], []
    for c0 in centres_x:
        y = erfcx_coeffs(c0, ERFCX_NT)
        ex_base.extend(limbs(y[0], 3))
        for n in range(1, ERFCX_NT + 1):
            h, l = limbs(y[n])
            ex_hi.append(h)
            ex_lo.append(l)
    out.append("\n/* erfcx(c + t) = base + t*(b1 + b2 t + ... + b%d t^%d), b_n = erfcx^(n)(c)/n!, about\n"
               "   centres (i + 1/2)/16 (i < 64), 2^e + (j + 1/2) 2^e/32 (e = 2, 3, 4) and\n"
               "   -(i + 1/2)/16 (i < 32, rows 160..191) */\n"
               % (ERFCX_NT, ERFCX_NT - 1))
    out.append("#define MF_ERFCX_NT %d\n#define MF_ERFCX_ROWS %d\n#define MF_ERFCX_ASYM %d\n"
               % (ERFCX_NT, ERFCX_ROWS, ERFCX_ASYM))
    out.append(arr("mf_erfcx_centre", [float(c0) for c0 in centres_x]))
    out.append(arr("mf_erfcx_base", ex_base))
    out.append(arr("mf_erfcx_hi", ex_hi))
    out.append(arr("mf_erfcx_lo", ex_lo))
    out.append(dd_pair("mf_inv_sqrtpi", 1 / mpmath.sqrt(pi)))

    # ---- Ozaki CRT form: moduli and CRT constants of the INT8 GEMM ----
    mods = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179,
            173, 167, 163, 157, 151, 149, 139, 137, 131, 127, 113, 109, 107, 103, 101, 97, 89, 83, 79, 73, 71]
    for a in range(len(mods)):
        for b in range(a):
            assert mpmath.mpf(mods[a]) and __import__("math").gcd(mods[a], mods[b]) == 1
    rl = []        # 1/m as five 36-bit limbs: floor(2^180 / m) = sum L_j 2^(36 (4 - j))
    for mm in mods:
        big = (1 << 180) // mm
        rl += [float((big >> (36 * (4 - j))) & ((1 << 36) - 1)) for j in range(5)]
    out.append("\n/* Ozaki CRT form: pairwise-coprime moduli (<= 256, so balanced residues are int8),\n"
               "   and 1/m as five 36-bit limbs */\n")
    out.append("#define MF_OZ2_MODULI %d\n" % len(mods))
    out.append("#ifdef __cplusplus\nstatic constexpr int mf_oz2_m_host[%d] = {%s};\n#endif\n"
               % (len(mods), ", ".join(map(str, mods))))
    out.append("MF_TABLE int mf_oz2_m[%d] = {%s};\n" % (len(mods), ", ".join(map(str, mods))))
    out.append(arr("mf_oz2_rlimb", rl, per_line=4))
    # CRT constants for the first N moduli, N = 1..40 (read by the device-side
    # plan of a call): y[N-1][t] = (M_N / m_t)^-1 mod m_t and
    # M_N = prod_{t<N} m_t rounded to a double-double
    ys, mp = [], []
    for N in range(1, len(mods) + 1):
        MN = 1
        for mm in mods[:N]:
            MN *= mm
        ys += [pow(MN // mods[t] % mods[t], -1, mods[t]) for t in range(N)]
        ys += [0] * (len(mods) - N)
        h = float(MN)
        mp += [h, float(MN - int(h))]
    out.append("MF_TABLE int mf_oz2_y[%d] = {\n%s\n};\n"
               % (len(ys), ",\n".join("    " + ", ".join(map(str, ys[r:r + len(mods)]))
                                      for r in range(0, len(ys), len(mods)))))
    out.append(arr("mf_oz2_mp", mp, per_line=4))

    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "csrc", "dd_tables.h")
    with open(dst, "w") as f:
        f.write("".join(out))
    print("wrote", os.path.normpath(dst), file=sys.stderr)


if __name__ == "__main__":
    main()
