"""GPU: the scan server under a live request stream.

Requests arrive (Poisson gaps, submitted from this thread) while earlier ones
are being stepped on the device; each result must equal the plain
single-search result on the same tables bit for bit.  The stream mixes:

* host-table requests small enough for the staged upload (pinned host
  staging, copies and transposes queued on the step stream);
* host-table requests larger than the server's pinned staging, which are
  uploaded synchronously on the ingest stream while staged uploads and steps
  are in flight (the device staging of a staged upload must not be handed to
  them before its transposes ran);
* requests on one resident dataset at rising orders (2, 3, 4, 5), whose
  binomial table grows while scans queued with the smaller table still read
  it.
"""
import numpy as np
import pytest

import fiuncho_amd as fa

pytestmark = pytest.mark.gpu


def synth(snps, cases, ctrls, seed):
    spec = fa.synth_spec(snps, cases, ctrls, seed=seed)
    words, _, cw, tw = fa.synth_generate(spec, words=True)
    return words, cw, tw


def same(a, b):
    return a[0].tolist() == b[0].tolist() and np.array_equal(a[1], b[1])


def plain(d, order, outputs):
    return fa.GpuSearch().run_arrays(d, order, fa.Distribution(d.snps, order - 1), outputs)


def test_stream_interleaved_with_running_steps():
    rng = np.random.default_rng(2024)
    # host-table requests: (snps, cases, ctrls, order, outputs)
    shapes = []
    for i in range(48):
        order = [2, 3, 2, 3, 4, 2, 3, 5][i % 8]
        snps = {2: int(rng.integers(150, 900)), 3: int(rng.integers(60, 220)), 4: int(rng.integers(24, 45)),
                5: int(rng.integers(14, 20))}[order]
        cases = int(rng.integers(300, 3000))
        ctrls = int(rng.integers(300, 3000))
        shapes.append((snps, cases, ctrls, order, int(rng.integers(5, 60))))
    # three requests above the pinned staging of a 64 MiB budget (16 MiB):
    # about 21 MB of tables each, uploaded on the synchronous path
    for k in range(3):
        shapes.insert(10 + 15 * k, (1000, 28000, 28000, 2, 30))
    reqs, expect = [], []
    print(f"case: serve stream, {len(shapes)} host-table requests, plain results", flush=True)
    for i, (snps, ca, ct, order, outputs) in enumerate(shapes):
        words, cw, tw = synth(snps, ca, ct, 7000 + i)
        reqs.append(dict(order=order, outputs=outputs, words=words, snps=snps, cases_words=cw, ctrls_words=tw,
                         cases=ca, ctrls=ct))
        d = fa.Dataset.from_words(words, snps, cw, tw, ca, ct)
        expect.append(plain(d, order, outputs))
        d.close()
    # the resident dataset's expected results, each from a fresh copy of it
    rwords, rcw, rtw = synth(40, 1800, 1500, 7999)
    resident_orders = [2, 3, 4, 5, 2, 4]
    resident_expect = {}
    for order in set(resident_orders):
        d = fa.Dataset.from_words(rwords, 40, rcw, rtw, 1800, 1500)
        resident_expect[order] = plain(d, order, 25)
        d.close()
    resident = fa.Dataset.from_words(rwords, 40, rcw, rtw, 1800, 1500)

    srv = fa.Server(fa.batch_config(memory_budget=64 << 20))
    gaps = rng.exponential(0.002, size=len(reqs))
    submitted = []            # (id, expected)
    r_iter = iter(resident_orders)
    t_stats = []
    print("case: serve stream, submitting", flush=True)
    import time
    for i, (r, g) in enumerate(zip(reqs, gaps)):
        time.sleep(float(g))
        submitted.append((srv.submit(**r), expect[i], i))
        if i % 8 == 3:
            order = next(r_iter, None)
            if order is not None:
                submitted.append((srv.submit(order=order, outputs=25, dataset=resident), resident_expect[order],
                                  f"resident order {order}"))
        if i % 5 == 4:
            # collect whatever is done already (poll), keep the rest running
            j = 0
            while j < len(submitted):
                rid, exp, tag = submitted[j]
                out = srv.wait(rid, timeout_ms=0.0)
                if out is None:
                    j += 1
                    continue
                assert same(out[:2], exp), tag
                t_stats.append(out[2])
                submitted.pop(j)
    for order in r_iter:
        submitted.append((srv.submit(order=order, outputs=25, dataset=resident), resident_expect[order],
                          f"resident order {order}"))
    print(f"case: serve stream, waiting for {len(submitted)} requests", flush=True)
    for rid, exp, tag in submitted:
        out = srv.wait(rid)
        assert same(out[:2], exp), tag
        t_stats.append(out[2])
    info = srv.info()
    srv.close()
    resident.close()
    assert info.completed == len(reqs) + len(resident_orders)
    assert info.memory_used == 0
    # requests really arrived while others were running on the device
    overlaps = sum(1 for a in t_stats for b in t_stats
                   if a is not b and b.start_ms < a.submit_ms < b.done_ms)
    assert overlaps > 0
    print(f"case: serve stream done, {overlaps} overlapping request pairs", flush=True)
