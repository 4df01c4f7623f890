"""The FP32C pair coefficients (A, Bx) of G_reg = A I + Bx x x^T as the
kernels evaluate them: a piecewise polynomial in s = r^2 (degree 7 per
interval, FP32 Horner), checked against the oracle's FP64 tensor
(ggem_stokes.C:304-357) over the whole range of pair distances and beyond
the cut-off, for several (alpha, xi)."""
import numpy as np
import pytest

import copss_ggem as G
import oracle as O
import workloads as W


def eval_table(tab, meta, s):
    """FP32 evaluation in the kernels' order of operations."""
    f32 = np.float32
    s = s.astype(f32)
    s1, inv1, inv2, d1, d2 = (f32(v) for v in meta[:5])
    n1, n2 = int(meta[5]), int(meta[6])
    r2 = s >= s1
    base = np.where(r2, s1, f32(0)).astype(f32)
    inv = np.where(r2, inv2, inv1).astype(f32)
    d = np.where(r2, d2, d1).astype(f32)
    j = np.floor((s - base) * inv).astype(np.int64)
    j = np.minimum(j, np.where(r2, n2, n1 - 1))
    i = np.where(r2, n1, 0) + j
    mid = ((j.astype(f32) + f32(0.5)) * d + base).astype(f32)
    y = ((s - mid) * (f32(2) * inv)).astype(f32)
    A = np.zeros_like(s)
    B = np.zeros_like(s)
    for k in range(7, -1, -1):
        A = (A * y + tab[i, k]).astype(f32)
        B = (B * y + tab[i, 8 + k]).astype(f32)
    return A.astype(np.float64), B.astype(np.float64)


def exact(cfg, s):
    """(A, Bx) from the oracle's tensor at x = (r, 0, 0)."""
    A = np.empty_like(s)
    B = np.empty_like(s)
    for n, sv in enumerate(s):
        r = np.sqrt(sv)
        Gm = O.green_local(cfg, np.array([r, 0.0, 0.0]))
        A[n] = Gm[1, 1]
        B[n] = (Gm[0, 0] - Gm[1, 1]) / sv
    return A, B


@pytest.mark.parametrize("alpha,br0", [(0.5, 1.0), (0.35, 1.0), (0.9, 1.0), (0.2, 1.0), (1.5, 0.5)])
def test_table_matches_fp64_tensor(alpha, br0):
    cfg = W.WORKLOADS["beads20k"]["cfg"](alpha=alpha)
    cfg["br0"] = br0
    c = O.consts(cfg)
    tab, meta = G.fp32c_table(c.one_eight_pi_mu, c.alpha, c.ksi)
    assert tab.shape[0] == int(meta[5]) + int(meta[6]) + 1
    assert np.all(tab[-1] == 0)
    # widths are powers of two and the intervals cover 20 decay lengths of
    # the slower scale, after which the table is zero
    k, m = max(c.alpha2, c.ksi2), min(c.alpha2, c.ksi2)
    assert np.log2(meta[3]) == np.round(np.log2(meta[3])) and meta[3] * k <= 1.4
    assert meta[0] + meta[6] * meta[4] >= 20 / m
    s_cut = float(meta[0] + meta[6] * meta[4])
    assert meta[7] == s_cut
    rng = np.random.default_rng(3)
    s = np.concatenate([np.linspace(1e-6, 1.0, 400), rng.uniform(0, s_cut, 1600),
                        np.linspace(0.9 * s_cut, 1.5 * s_cut, 200)])
    At, Bt = eval_table(tab, meta, s)
    Ae, Be = exact(cfg, s)
    a_scale, b_scale = np.max(np.abs(Ae)), np.max(np.abs(Be))
    ea = np.max(np.abs(At - Ae)) / a_scale
    eb = np.max(np.abs(Bt - Be)) / b_scale
    print("alpha %.2f xi %.3f intervals %d: max err A %.2e Bx %.2e" % (c.alpha, c.ksi, tab.shape[0], ea, eb))
    assert ea < 3e-7 and eb < 3e-7
    # s = 0 gives coeff_2, the reference's coincident-point value
    A0, _ = eval_table(tab, meta, np.zeros(1))
    assert abs(A0[0] - c.coeff_2) < 3e-7 * abs(c.coeff_2)


def test_table_series_join():
    """Near s = 0 the builder switches from the power series to the closed
    form; both sides agree with the oracle."""
    cfg = W.WORKLOADS["beads20k"]["cfg"]()
    c = O.consts(cfg)
    tab, meta = G.fp32c_table(c.one_eight_pi_mu, c.alpha, c.ksi)
    k = max(c.alpha2, c.ksi2)
    s = np.linspace(0.2 / k, 0.3 / k, 101)
    At, Bt = eval_table(tab, meta, s)
    Ae, Be = exact(cfg, s)
    assert np.max(np.abs(At - Ae)) < 3e-7 * np.max(np.abs(Ae))
    assert np.max(np.abs(Bt - Be)) < 3e-7 * np.max(np.abs(Be))
