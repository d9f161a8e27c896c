"""Writes ff_log_table.h: the table and polynomial of the natural logarithm
used by the FF frequency-grid kernel (log_abs in ff_math.h).

Method (table-driven reduction, as in Tang's "Table-driven implementation of
the logarithm function", ACM TOMS 16(4), 1990): |x| = m 2^k with m in
[0.5, 1); the top 8 fraction bits of m select i, the interval
[0.5 + i/512, 0.5 + (i+1)/512) with centre c_i; with invc_i = RN(1/c_i),
    r = m invc_i - 1          (one fma, |r| < 2^-9),
    log|x| = k ln2 + logc_i + log1p(r),   logc_i = RN(-ln(invc_i)),
    log1p(r) = r + r^2 Q(r),
Q the degree-3 Chebyshev economization on [-2^-9, 2^-9] of the Taylor series
of (log1p(r) - r)/r^2 = sum_j (-1)^(j+1) r^j / (j+2); its error there is
below 3.2e-13 (times r^2 <= 2^-18: below 1.3e-18 of the logarithm).

Everything is computed in exact rational arithmetic (fractions) or with
50-digit decimals and rounded once to double.  Run from this directory:
    python3 ff_log_table.py > ff_log_table.h
"""
from decimal import Decimal, getcontext
from fractions import Fraction

BITS = 8                       # table index bits
N = 1 << BITS
H = Fraction(1, 1 << (BITS + 1))  # |r| bound
DEGREE = 3                     # degree of Q
TAYLOR = 16                    # Taylor terms before economization


def chebyshev(n):
    """Coefficients of T_n(t) in powers of t, lowest first."""
    a, b = [Fraction(1)], [Fraction(0), Fraction(1)]
    if n == 0:
        return a
    for _ in range(n - 1):
        c = [Fraction(0)] + [2 * v for v in b]
        for k, v in enumerate(a):
            c[k] -= v
        a, b = b, c
    return b


def economized_q():
    """Q's coefficients (lowest first) and the bound of the dropped terms."""
    coef = [Fraction((-1) ** (j + 1), j + 2) for j in range(TAYLOR + 1)]
    dropped = Fraction(1, 2 + TAYLOR + 1) * H ** (TAYLOR + 1) * 2  # tail
    for n in range(TAYLOR, DEGREE, -1):
        an = coef[n]
        tn = chebyshev(n)
        scale = Fraction(1, 2 ** (n - 1))
        # r^n = h^n 2^(1-n) T_n(r/h) - sum_{k<n} c_{n,k} 2^(1-n) h^(n-k) r^k
        for k in range(n):
            coef[k] -= an * tn[k] * scale * H ** (n - k)
        coef[n] = Fraction(0)
        dropped += abs(an) * scale * H ** n
    return coef[:DEGREE + 1], dropped


def table():
    getcontext().prec = 50
    rows = []
    for i in range(N):
        c = Fraction(1, 2) + Fraction(2 * i + 1, 4 * N)
        invc = float(1 / c)
        logc = float(-Decimal(invc).ln())
        rows.append((invc, logc))
    return rows


def header():
    q, dropped = economized_q()
    assert dropped < Fraction(32, 10 ** 14), float(dropped)
    out = [
        "// ff_log_table.h — generated by ff_log_table.py; do not edit.",
        "//",
        "// Table and polynomial of log_abs (ff_math.h): for interval i of the",
        "// reduced argument m in [0.5, 1), {invc_i, logc_i} with",
        "// invc_i = RN(1/c_i), c_i = 0.5 + (2i+1)/%d, logc_i = RN(-ln invc_i);" % (4 * N),
        "// Q(r) = q0 + q1 r + q2 r^2 + q3 r^3 ~ (log1p(r) - r)/r^2 on",
        "// |r| <= 2^-%d, within %.1e." % (BITS + 1, float(dropped)),
        "#pragma once",
        "",
        "#if defined(__HIPCC__)",
        "#define FF_LOG_TABLE_ATTR __device__",
        "#else",
        "#define FF_LOG_TABLE_ATTR",
        "#endif",
        "",
        "namespace gpp {",
        "namespace ff {",
        "",
        "constexpr int kLogTableBits = %d;" % BITS,
        "constexpr double kLogQ0 = %s;" % float(q[0]).hex(),
        "constexpr double kLogQ1 = %s;" % float(q[1]).hex(),
        "constexpr double kLogQ2 = %s;" % float(q[2]).hex(),
        "constexpr double kLogQ3 = %s;" % float(q[3]).hex(),
        "",
        "FF_LOG_TABLE_ATTR const double kLogTable[%d] = {" % (2 * N),
    ]
    for invc, logc in table():
        out.append("    %s, %s," % (invc.hex(), logc.hex()))
    out += ["};", "", "}  // namespace ff", "}  // namespace gpp", ""]
    return "\n".join(out)


if __name__ == "__main__":
    import sys
    sys.stdout.write(header())
