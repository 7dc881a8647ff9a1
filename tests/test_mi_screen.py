"""The FP32 screen that decides which combinations get the exact FP64 MI.

The scan kernels screen every combination with

    MI = H(Y) + (1/N) sum_x [n_cx ln(n_cx / n_x) + n_tx ln(n_tx / n_x)]

in FP32 (screen_sum_fn: per cell r = v_rcp_f32(max(n_x, 1)), then
s = fmaf(n_cx, v_log_f32(max(n_cx r, 7.9e-31)), s) and the same for the
controls, in cell order; ms = fmaf(s, float(1/N) * 0.69314718f, float(H(Y)))),
and compute the exact value only when ms >= tau - delta, delta = 2e-5 nats
(mi_constants).  The tables of orders 4 to 6 are split over L = 3^(k-3) lanes
(k_scan_km): each lane sums its 27 consecutive cells, then the L partial
sums are added in lane order.  A combination is dropped wrongly only if
exact - screen > delta, so |screen - exact| must stay below delta for every
table of every kernel shape.  The reference's MI is mi_compute.cpp:14-41 /
mi_compute_512.cpp:20-69; this screen only filters, the reported value is the
FP64 one.

Two checks:
* the worst-case bound from the operation order: every term is one-signed
  (n ln(n/n_x) <= 0), so a chain of m fused adds errs by at most
  (m - 1) u sum |t| (u = 2^-24), and sum |t| <= N log2(e) H(Y|X) <= N; each
  term's reciprocal, product and logarithm add at most 2 u / ln 2 + e_log
  per sample, e_log the logarithm's absolute error in log2 units (taken as
  2^-21, far above the v_log_f32 accuracy); the final fused multiply-add adds
  the rounding of float(1/N) * 0.69314718f and of H(Y).  This bound, in nats,
  is below delta / 2 for every kernel shape;
* an emulation of the device arithmetic (float32 after every operation,
  fmaf as one rounding) with the reciprocal and the logarithm pushed one ulp
  (plus the absolute floor) in the same direction for every cell -- the
  worst systematic bias -- on random and adversarial tables up to 2^23
  samples per class: |screen - exact| stays below that bound.
"""
import math

import numpy as np
import pytest

F = np.float32
U = 2.0 ** -24
E_LOG = 2.0 ** -21
DELTA = 2e-5
LN2_F = float(F(0.69314718))

# (cells per lane, lanes): the kernels' screen shapes
SHAPES = {
    "order 2 (scan2m, mfma2, direct)": (9, 1),
    "order 3 (scan3m, mfma3, direct)": (27, 1),
    "order 4 direct": (81, 1),
    "order 4 k_scan_km": (27, 3),
    "order 5 k_scan_km": (27, 9),
    "order 6 k_scan_km": (27, 27),
}


def worst_case_bound(cells, lanes):
    """Worst-case |screen - exact| in nats for a table of cells x lanes."""
    m_lane = 2 * cells                       # fused adds per lane
    chain = (m_lane - 1) + (lanes - 1)       # longest rounding chain of one term
    gamma = chain * U / (1 - chain * U)
    s_err = gamma + 2 * U / math.log(2) + E_LOG   # per sample, in log2 units (sum |t| <= N)
    # final fmaf: s * c + hy, c = float(float(1/N) * 0.69314718f); |s c| <= ln 2
    c_rel = 2 * U + abs(LN2_F - math.log(2)) / math.log(2)
    return s_err * math.log(2) * (1 + c_rel) + c_rel * math.log(2) + 2 * U * math.log(2) + U


def exact_mi(c, t):
    c, t = c.astype(np.float64), t.astype(np.float64)
    n = c.sum() + t.sum()

    def plogp(p):
        p = p[p > 0]
        return float((p * np.log(p)).sum())

    return -plogp((c + t) / n) - plogp(np.array([c.sum(), t.sum()]) / n) + plogp(c / n) + plogp(t / n)


def f32_step(x, direction):
    """float32 x moved one ulp in `direction` (+1, -1 or 0)."""
    if direction == 0:
        return F(x)
    return np.nextafter(F(x), F(np.inf if direction > 0 else -np.inf), dtype=F)


def device_screen(c, t, lanes, bias):
    """FP32 emulation of screen_sum_fn per lane + the lane sum + the final
    fmaf; `bias` +1 / -1 pushes every reciprocal and logarithm one ulp (and
    the logarithm by E_LOG more) up / down, 0 rounds correctly."""
    nc, nt = int(c.sum()), int(t.sum())
    n = nc + nt
    cells = len(c) // lanes
    parts = []
    for p in range(lanes):
        s = F(0)
        for x in range(p * cells, (p + 1) * cells):
            fc, ft = F(c[x]), F(t[x])
            r = f32_step(1.0 / np.float64(max(F(fc + ft), F(1))), bias)
            for f in (fc, ft):
                arg = max(F(f * r), F(7.9e-31))
                lg = f32_step(math.log2(float(arg)), bias)
                lg = F(np.float64(lg) + bias * E_LOG)
                s = F(np.float64(f) * np.float64(lg) + np.float64(s))   # fmaf
        parts.append(s)
    tot = parts[0]
    for v in parts[1:]:
        tot = F(tot + v)
    hy = 0.0
    for k in (nc, nt):
        if k:
            hy -= (k / n) * math.log(k / n)
    scale = F(F(1.0 / n) * F(0.69314718))
    return float(F(np.float64(tot) * np.float64(scale) + np.float64(F(hy))))


def tables(rng, cells, count):
    sizes = [1, 7, 100, 8000, 50000, 1 << 22, 1 << 23]
    for i in range(count):
        kind = i % 6
        nc, nt = int(rng.choice(sizes)), int(rng.choice(sizes))
        if kind == 0:     # independent, many cells: H(Y|X) near its maximum
            p = rng.dirichlet(np.ones(cells))
            c, t = rng.multinomial(nc, p), rng.multinomial(nt, p)
        elif kind == 1:   # uniform cells (every term the same size)
            p = np.full(cells, 1.0 / cells)
            c, t = rng.multinomial(nc, p), rng.multinomial(nt, p)
        elif kind == 2:   # strongly dependent
            c = rng.multinomial(nc, rng.dirichlet(np.full(cells, 0.3)))
            t = rng.multinomial(nt, rng.dirichlet(np.full(cells, 0.3)))
        elif kind == 3:   # disjoint supports (MI = H(Y))
            half = cells // 2
            pc, pt = np.zeros(cells), np.zeros(cells)
            pc[:half] = rng.dirichlet(np.ones(half))
            pt[half:] = rng.dirichlet(np.ones(cells - half))
            c, t = rng.multinomial(nc, pc), rng.multinomial(nt, pt)
        elif kind == 4:   # a few heavy cells, the rest single samples
            c, t = np.ones(cells, np.int64), np.ones(cells, np.int64)
            c[rng.integers(0, cells, 2)] += nc
            t[rng.integers(0, cells, 2)] += nt
        else:             # near independence, large classes (small MI, where the screen matters)
            p = rng.dirichlet(np.ones(cells) * 5)
            q = np.clip(p + rng.normal(0, 0.002, cells), 1e-9, None)
            c, t = rng.multinomial(nc, p), rng.multinomial(nt, q / q.sum())
        yield np.asarray(c, np.int64), np.asarray(t, np.int64)


def test_worst_case_bound_below_half_the_margin():
    for name, (cells, lanes) in SHAPES.items():
        b = worst_case_bound(cells, lanes)
        assert b < DELTA / 2, (name, b)
    # the longest chain (order 4 on one lane, 161 fused adds) stays the worst
    assert max(SHAPES.values(), key=lambda s: worst_case_bound(*s)) == (81, 1)


@pytest.mark.parametrize("name", list(SHAPES))
def test_emulated_screen_within_the_bound(name):
    cells, lanes = SHAPES[name]
    rng = np.random.default_rng(cells * 100 + lanes)
    bound = worst_case_bound(cells, lanes)
    count = 60 if cells * lanes >= 243 else 150
    worst = 0.0
    for c, t in tables(rng, cells * lanes, count):
        if c.sum() + t.sum() == 0 or c.sum() + t.sum() >= 1 << 24:
            continue
        ex = exact_mi(c, t)
        for bias in (-1, 0, 1):
            err = abs(device_screen(c, t, lanes, bias) - ex)
            assert err <= bound, (name, bias, err, bound)
            worst = max(worst, err)
    assert worst < DELTA / 2


def test_screen_edge_cases():
    z = np.zeros(27, np.int64)
    a, b = z.copy(), z.copy()
    a[0], b[1] = 500, 500          # MI = H(Y) = ln 2
    assert abs(device_screen(a, b, 1, 0) - math.log(2)) < 1e-6
    c = np.arange(27, dtype=np.int64)
    assert abs(device_screen(c, c, 1, 0)) < 1e-6    # identical distributions: MI 0
    one = z.copy()
    one[5] = 9
    assert abs(device_screen(one, z, 1, 0)) < 1e-6  # one class only
