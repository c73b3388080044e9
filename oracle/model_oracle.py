"""CPU restatement of the attenuation pre-pass -- TEST INFRASTRUCTURE ONLY.

DISPERSION (dispersion.f:1-131; Aki & Richards eq. 5.88): with
f = ln(Re omega / (2 pi freqbase)) and q = 1/Q per node,

    c = c_base max(1 + f q / pi, 0.1),   c~ = c (1 - i q / 2),
    mu = vs~^2 rho,   lambda = vp~^2 rho - 2 mu,

skipped (lossless moduli kept) when Re omega <= 10 eps or freqbase <= 0.
Parity: pinned against the compiled reference in FP64 and FP32 and by
closed forms (tests/test_model_oracle.py).
"""
import numpy as np


def dispersion(vp, vs, qpi, qsi, rho, freqbase, omega_re, eps=10 * np.finfo(np.float64).eps):
    vp, vs, qpi, qsi, rho = (np.asarray(a, np.float64) for a in (vp, vs, qpi, qsi, rho))
    omegabase = freqbase * 2.0 * np.pi
    if not (omega_re > eps and omegabase > 0.0):
        mu = (vs * vs * rho).astype(np.complex128)
        return (vp * vp * rho - 2.0 * mu).astype(np.complex128), mu
    f = np.log(omega_re / omegabase)
    fs = np.maximum(1.0 + (qsi / np.pi) * f, 0.1)
    fp = np.maximum(1.0 + (qpi / np.pi) * f, 0.1)
    cs = vs * fs * (1.0 - 0.5j * qsi)
    cp = vp * fp * (1.0 - 0.5j * qpi)
    mu = cs * cs * rho
    return cp * cp * rho - 2.0 * mu, mu
