"""Regenerates the polynomial coefficients of mpgos_math.h.

Each kernel is a Chebyshev (near-minimax) fit, with mpmath at 50 digits, of the
function that remains after the leading Taylor terms are taken out:
  cos(r) = 1 - z/2 + z^2 C(z),      z = r^2, r in [0, pi/4]
  sin(r) = r + r z S(z)
  e^r    = 1 + r + r^2 E(r),         r in [-ln2/2, ln2/2]
  2 atanh(s) = 2 s + s z L(z),       z = s^2, s in [0, (sqrt2-1)/(sqrt2+1)]
Run:  python gen_math_coeffs.py   (prints hex literals, highest degree first)
"""
import mpmath as mp

mp.mp.dps = 50


def fit(g, lo, hi, n):
    poly, err = mp.chebyfit(g, [lo, hi], n, error=True)
    return poly, err


def main():
    zc = (mp.pi / 4) ** 2
    kernels = {
        "C (cos)": (lambda z: mp.mpf(1) / 24 if z == 0 else (mp.cos(mp.sqrt(z)) - 1 + z / 2) / z ** 2, 0, zc, 6),
        "S (sin)": (lambda z: -mp.mpf(1) / 6 if z == 0 else (mp.sin(mp.sqrt(z)) - mp.sqrt(z)) / (mp.sqrt(z) * z), 0, zc, 6),
        "E (exp)": (lambda r: mp.mpf(1) / 2 if r == 0 else (mp.exp(r) - 1 - r) / r ** 2, -mp.log(2) / 2, mp.log(2) / 2, 11),
        "L (log)": (lambda z: mp.mpf(2) / 3 if z == 0 else (2 * mp.atanh(mp.sqrt(z)) - 2 * mp.sqrt(z)) / (mp.sqrt(z) * z),
                    0, ((mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)) ** 2, 7),
    }
    for name, (g, lo, hi, n) in kernels.items():
        poly, err = fit(g, lo, hi, n)
        print(f"{name}: {n} coefficients, max fit error {mp.nstr(err, 3)}")
        for c in poly:
            print("   ", float(c).hex())


if __name__ == "__main__":
    main()
