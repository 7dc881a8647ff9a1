"""The scan kernels reject a combination before the logarithm-based FP32
screen when an upper bound on its MI is already below the threshold:

    MI = KL(P_XY || P_X P_Y) <= ln(1 + chi2/N) <= chi2/N,
    chi2/N = N/(n_c n_t) * sum_x (n_cx - n_x n_c/N)^2 / n_x,

computed in FP32 and accepted when bound * 1.0001 + 4e-6 >= tau - delta.
These tests emulate that FP32 evaluation (same operation order, rounding to
float32 after each step, the reciprocal correctly rounded -- the device's
v_rcp_f32 is within 1 ulp, far inside the margin) and check on random and
adversarial contingency tables that the raised bound never falls below the
exact MI (FP64), so the rejection can never drop a combination the screen
would have kept.
"""
import numpy as np
import pytest

F = np.float32


def exact_mi(c, t):
    """H(X) + H(Y) - H(X, Y) in FP64 from integer counts (nats)."""
    c = c.astype(np.float64)
    t = t.astype(np.float64)
    n = c.sum() + t.sum()

    def plogp(p):
        p = p[p > 0]
        return float((p * np.log(p)).sum())

    return -plogp((c + t) / n) - plogp(np.array([c.sum(), t.sum()]) / n) + plogp(c / n) + plogp(t / n)


def device_bound(c, t):
    """FP32 emulation of the kernels' bound (fmaf = one rounding)."""
    nc, nt = int(c.sum()), int(t.sum())
    n = nc + nt
    frac_c = F(nc / n)
    chi_scale = F(n / (nc * nt)) if nc and nt else F(0)
    s = F(0)
    for a0, a1 in zip(c.tolist(), t.tolist()):
        fx = F(a0 + a1)
        d = F(np.float64(a0) - np.float64(fx) * np.float64(frac_c))        # fmaf(-fx, frac_c, a0)
        r = F(1.0 / np.float64(max(fx, F(1))))
        s = F(np.float64(F(d * d)) * np.float64(r) + np.float64(s))          # fmaf(d*d, r, s)
    return F(s * chi_scale)


def raised(b):
    return float(F(np.float64(b) * np.float64(F(1.0001)) + np.float64(F(4e-6))))


def tables(rng, cells, count):
    for _ in range(count):
        kind = rng.integers(0, 5)
        nc = int(rng.choice([1, 7, 100, 1000, 8000, 50000]))
        nt = int(rng.choice([1, 7, 100, 1000, 8000, 50000]))
        if kind == 0:     # independent
            p = rng.dirichlet(np.ones(cells))
            c, t = rng.multinomial(nc, p), rng.multinomial(nt, p)
        elif kind == 1:   # strongly dependent
            c = rng.multinomial(nc, rng.dirichlet(np.full(cells, 0.3)))
            t = rng.multinomial(nt, rng.dirichlet(np.full(cells, 0.3)))
        elif kind == 2:   # disjoint supports (MI = H(Y))
            half = cells // 2
            pc = np.zeros(cells)
            pc[:half] = rng.dirichlet(np.ones(half))
            pt = np.zeros(cells)
            pt[half:] = rng.dirichlet(np.ones(cells - half))
            c, t = rng.multinomial(nc, pc), rng.multinomial(nt, pt)
        elif kind == 3:   # many empty cells, one planted cell
            p = np.zeros(cells)
            p[rng.integers(0, cells, 3)] = 1.0
            p /= p.sum()
            c, t = rng.multinomial(nc, p), rng.multinomial(nt, p)
            c[rng.integers(0, cells)] += nc // 10 + 1
        else:             # small perturbation of independence
            p = rng.dirichlet(np.ones(cells) * 5)
            q = np.clip(p + rng.normal(0, 0.002, cells), 1e-9, None)
            c, t = rng.multinomial(nc, p), rng.multinomial(nt, q / q.sum())
        yield c, t


@pytest.mark.parametrize("cells", [9, 27, 81])
def test_bound_never_below_mi(cells):
    rng = np.random.default_rng(cells)
    worst = np.inf
    for c, t in tables(rng, cells, 400):
        mi = exact_mi(c, t)
        b = device_bound(c, t)
        assert raised(b) >= mi, (c.tolist(), t.tolist(), mi, float(b))
        if mi > 1e-4:
            worst = min(worst, float(b) / mi)
    assert worst >= 1.0


def test_bound_tracks_two_mi_near_independence():
    """Near independence chi2/N ~ 2 MI: the bound is useful (rejects most
    combinations), not just valid."""
    rng = np.random.default_rng(5)
    p = rng.dirichlet(np.ones(27) * 5)
    ratios = []
    for _ in range(50):
        c, t = rng.multinomial(8000, p), rng.multinomial(8000, p)
        ratios.append(float(device_bound(c, t)) / exact_mi(c, t))
    assert 1.9 < float(np.median(ratios)) < 2.2


def test_bound_edge_cases():
    z = np.zeros(27, dtype=np.int64)
    one_class = z.copy()
    one_class[3] = 10
    assert device_bound(one_class, z) == 0.0 and exact_mi(one_class, z) == 0.0
    # identical distributions: MI 0, bound 0 up to rounding
    c = np.arange(27)
    assert abs(float(device_bound(c, c))) < 1e-6
    # one cell per class: MI = H(Y) = ln 2, chi2/N = 1
    a, b = z.copy(), z.copy()
    a[0], b[1] = 500, 500
    assert abs(exact_mi(a, b) - np.log(2)) < 1e-12
    assert abs(float(device_bound(a, b)) - 1.0) < 1e-6
