"""The mixed path's FP32 sin/cos error bound (kSinCos32Err in
oscars_amd/csrc/lw_common.h) against the FP32 polynomials of sincos_coeffs.h,
evaluated on the CPU in the kernel's operation order (sincos_reduced32 in
om_device.h): each fmaf is an exact float64 product-sum rounded once to
float32, as the hardware's fused multiply-add rounds."""
import os
import re

import numpy as np

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oscars_amd", "csrc")


def _coeffs():
    txt = open(os.path.join(CSRC, "sincos_coeffs.h")).read()
    co = {m.group(1): float(m.group(2)) for m in re.finditer(r"#define (OM_\w+) \(([-0-9.e+]+)f?\)", txt)}
    return [np.float32(co["OM_S32_%d" % i]) for i in range(3)], [np.float32(co["OM_C32_%d" % i]) for i in range(3)]


def _bound():
    txt = open(os.path.join(CSRC, "lw_common.h")).read()
    return float(re.search(r"constexpr double kSinCos32Err = ([0-9.e+-]+);", txt).group(1))


def _fmaf(a, b, c):
    # float32 operands: the float64 product is exact; rounding the float64 sum
    # again to float32 can differ from the fused single rounding only in rare
    # near-halfway cases, by one float32 ulp, which the bound's margin covers
    return (a.astype(np.float64) * b + c).astype(np.float32)


def sincos32(r):
    S, Cc = _coeffs()
    f32 = np.float32
    rf = r.astype(f32)
    u = (rf * rf).astype(f32)
    ps = _fmaf(u, S[2], S[1])
    ps = _fmaf(u, ps, S[0])
    sr = _fmaf((rf * u).astype(f32), ps, rf)
    pc = _fmaf(u, Cc[2], Cc[1])
    pc = _fmaf(u, pc, Cc[0])
    cr = _fmaf((u * u).astype(f32), pc, _fmaf(np.full_like(u, -0.5), u, np.ones_like(u)))
    return sr, cr


def test_fp32_sincos_within_bound():
    rng = np.random.default_rng(7)
    e = _bound()
    worst = 0.0
    for it in range(4):
        r = rng.uniform(-np.pi / 4, np.pi / 4, 2_000_000)
        if it == 0:
            r[:6] = [np.pi / 4, -np.pi / 4, 0.0, 1e-300, np.nextafter(np.pi / 4, 0), -1e-8]
        s, c = sincos32(r)
        worst = max(worst, float(np.max(np.abs(s - np.sin(r)))), float(np.max(np.abs(c - np.cos(r)))))
    print("FP32 sin/cos worst error %.3e, bound %.3e" % (worst, e))
    assert worst < e
    # the bound keeps a margin but is not loose enough to hide a wrong polynomial
    assert worst > 0.5 * e


def test_bound_catches_a_perturbed_polynomial():
    # a plausible mistake (one coefficient off by 1e-3 relative) exceeds the bound
    S, Cc = _coeffs()
    r = np.linspace(-np.pi / 4, np.pi / 4, 200001)
    f32 = np.float32
    rf = r.astype(f32)
    u = (rf * rf).astype(f32)
    ps = _fmaf(u, S[2], S[1])
    ps = _fmaf(u, ps, f32(S[0] * 1.001))
    sr = _fmaf((rf * u).astype(f32), ps, rf)
    assert float(np.max(np.abs(sr - np.sin(r)))) > _bound()
