"""Generate the polynomial coefficients used by vp_math.h.

Each table is a near-minimax fit in relative error (Lawson iteratively
reweighted least squares on Chebyshev nodes, targets from mpmath at 40
digits), rounded to float32 and printed as C hex-float literals:

* ERF_P   : erf(x)/x  as a polynomial in z = x^2 on [0, 1]           (degree 6)
* ERFCX_A : exp(x^2) erfc(x) on [1, 2.2] as a polynomial in the centred
            variable u = (1/x - m) * s mapping the interval onto [-1, 1]   (degree 8)
* ERFCX_B : same on [2.2, 4.5]                                        (degree 8)
* ERFCX_C : same on [4.5, 10.1]                                       (degree 8)
  (m, s) of each interval are printed beside the table.
* ERFINV_A: erfinv(x)/x as a polynomial in w - 2.5, w = -log(1-x^2) < 5    (degree 9)
* ERFINV_B: erfinv(x)/x as a polynomial in sqrt(w) - 3, w >= 5            (degree 9)
* CHORD_S : S(d) = sum_n d^n / (2n+1)!! = sqrt(pi/2) e^(d/2) erf(sqrt(d/2)) / sqrt(d)
            on [0, 9] (degree 10); the optical depth of a whole chord through a
            3-sigma-truncated Gaussian is 2 e^(-9/2) M dnf sqrt(d)/|w| S(d), with d
            the chord's squared half-length in the primitive's local frame

The erfinv variable substitution is the one of M. Giles, "Approximating the
erfinv function" (GPU Computing Gems, 2011); the coefficients are refitted
here.  Run:  python gen_math_coeffs.py  > coeffs.txt
"""
import numpy as np
import mpmath as mp

mp.mp.dps = 40


def cheb(a, b, n):
    k = np.arange(n)
    return 0.5 * (a + b) + 0.5 * (b - a) * np.cos(np.pi * (k + 0.5) / n)


def fit_rel(bx, ys, deg, iters=60):
    A = np.vstack([bx ** k for k in range(deg + 1)]).T
    w = np.ones_like(ys)
    for _ in range(iters):
        W = w / np.abs(ys)
        c, *_ = np.linalg.lstsq(A * W[:, None], ys * W, rcond=None)
        err = np.abs(A @ c - ys) / np.abs(ys)
        w = w * (err / err.max()) ** 0.5 + 1e-12
        w /= w.max()
    return c, err.max()


def emit(name, c, err):
    c32 = [float(np.float32(v)) for v in c]
    print(f"// {name}: degree {len(c)-1}, fit max rel err {err:.3e}")
    print(f"static constexpr float {name}[{len(c)}] = {{")
    for v in c32:
        print(f"    {float(v).hex().replace('0x', '0x')}f,  // {v:.9e}")
    print("};")


def main():
    xs = np.abs(cheb(0.0, 1.0, 3000)) + 1e-9
    ys = np.array([float(mp.erf(mp.mpf(x)) / mp.mpf(x)) for x in xs])
    emit("ERF_P", *fit_rel(xs * xs, ys, 6))
    for name, a, b in (("ERFCX_A", 1.0, 2.2), ("ERFCX_B", 2.2, 4.5), ("ERFCX_C", 4.5, 10.1)):
        xs = cheb(a, b, 3000)
        ys = np.array([float(mp.erfc(mp.mpf(x)) * mp.exp(mp.mpf(x) ** 2)) for x in xs])
        m = np.float32(0.5 * (1.0 / a + 1.0 / b))
        s = np.float32(2.0 / (1.0 / a - 1.0 / b))
        print(f"// {name} centring: m = {float(m).hex()}f ({m:.9e}), s = {float(s).hex()}f ({s:.9e})")
        emit(name, *fit_rel((1.0 / xs - float(m)) * float(s), ys, 8))
    ws = np.linspace(1e-7, 5.0, 5000)
    ys = np.array([float(mp.erfinv(mp.sqrt(1 - mp.exp(-mp.mpf(w)))) / mp.sqrt(1 - mp.exp(-mp.mpf(w)))) for w in ws])
    emit("ERFINV_A", *fit_rel(ws - 2.5, ys, 9))
    ws = np.linspace(5.0, 16.7, 5000)
    ys = np.array([float(mp.erfinv(mp.sqrt(1 - mp.exp(-mp.mpf(w)))) / mp.sqrt(1 - mp.exp(-mp.mpf(w)))) for w in ws])
    emit("ERFINV_B", *fit_rel(np.sqrt(ws) - 3.0, ys, 9))
    xs = cheb(0.0, 9.0, 4000)
    ys = np.array([1.0 if x == 0 else float(mp.sqrt(mp.pi / 2) * mp.exp(mp.mpf(x) / 2) * mp.erf(mp.sqrt(mp.mpf(x) / 2))
                                            / mp.sqrt(mp.mpf(x))) for x in xs])
    emit("CHORD_S", *fit_rel(xs, ys, 10, iters=80))


if __name__ == "__main__":
    main()
