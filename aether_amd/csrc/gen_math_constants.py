"""Constants of the inverse-trigonometric, hyperbolic and power routines in dmath.hpp (and their oracle
twin), computed with mpmath at 120 digits and printed as C hex-float literals:

  atan(k/16) and pi/2 - atan(k/16), k = 0..16, as double-double (hi, lo) pairs;
  pi/2 and pi as double-double;
  the Taylor coefficients of tanh x = sum_j c_j x^(2j+1), j = 1..21 (from the Bernoulli numbers);
  2/3 as double-double.
The reciprocal-odd-integer coefficients (1/5, 1/7, ...) are written directly as divisions in the sources.
"""
import mpmath as mp

mp.mp.dps = 120


def dd(v):
    hi = float(v)
    lo = float(v - mp.mpf(hi))
    return hi, lo


def hexf(x):
    return float.hex(x)


def main():
    print("atan(k/16):")
    for k in range(17):
        h, l = dd(mp.atan(mp.mpf(k) / 16))
        print(f"  {{{hexf(h)}, {hexf(l)}}},")
    print("pi/2 - atan(k/16):")
    for k in range(17):
        h, l = dd(mp.pi / 2 - mp.atan(mp.mpf(k) / 16))
        print(f"  {{{hexf(h)}, {hexf(l)}}},")
    print("pi/2:", [hexf(v) for v in dd(mp.pi / 2)])
    print("pi:", [hexf(v) for v in dd(mp.pi)])
    print("2/3:", [hexf(v) for v in dd(mp.mpf(2) / 3)])
    print("tanh Taylor c_j, j = 1..21 (x^(2j+1)):")
    for j in range(1, 22):
        n = j + 1  # coefficient of x^(2n-1): 2^(2n) (2^(2n) - 1) B_(2n) / (2n)!
        c = mp.mpf(2) ** (2 * n) * (mp.mpf(2) ** (2 * n) - 1) * mp.bernoulli(2 * n) / mp.factorial(2 * n)
        print(f"  {hexf(float(c))},  // x^{2 * n - 1}")


if __name__ == "__main__":
    main()
