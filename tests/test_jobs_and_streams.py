"""Committed job lists and product-wide source rules (CPU only).

* The batched and served job lists under nixis_amd/jobs/ are the outputs of
  their generators (batch.batched_jobs, serve.served_jobs at one job per
  second), so the measured workloads are reproducible from either.
* No stream of raised or lowered priority is created anywhere in the
  product: work on streams of mixed priorities has hung these GPUs, and the
  served form's second stream is one default-priority stream made by
  _lib.side_stream.  torch's stream pool (torch.cuda.Stream()) creates
  raised-priority streams when first used, so it is not called either.
"""
import os
import re

import pytest

from nixis_amd import batch, serve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nixis_amd")


def test_batched_job_list_is_the_generators():
    jobs = batch.load_jobs(batch.BATCHED_64_PATH)
    assert jobs == batch.batched_jobs()
    assert len(jobs) == 64 and [j.seed for j in jobs] == list(range(64))
    assert all(64 <= j.d <= 1024 and j.d % 2 == 0 for j in jobs)
    assert all(4 <= j.n_octaves <= 16 and 0 <= j.erosion_iters <= 200 for j in jobs)


def test_served_trace_is_the_generators():
    jobs = batch.load_jobs(batch.SERVED_120_PATH)
    assert jobs == serve.served_jobs(120, rate=1.0)
    for rate in (3.0, 250.0):
        scaled = serve.load_trace(batch.SERVED_120_PATH, rate)
        want = serve.served_jobs(120, rate=rate)
        assert [j.d for j in scaled] == [j.d for j in want]
        assert all(abs(a.arrival - b.arrival) <= 1e-12 * max(1.0, b.arrival)
                   for a, b in zip(scaled, want))


def _product_sources():
    for base, _, files in os.walk(PKG):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                yield os.path.join(base, f)


PRIORITY_PATTERNS = [
    r"torch\.cuda\.Stream\s*\(",
    r"cuda\.streams\.Stream\s*\(",
    r"hipStreamCreateWithPriority",
    r"is_high_priority_stream",
    r"priority\s*=",
    r"get_stream_from_external\s*\(",
]


@pytest.mark.parametrize("pattern", PRIORITY_PATTERNS)
def test_no_prioritised_or_pooled_streams(pattern):
    hits = []
    for path in _product_sources():
        with open(path, encoding="utf-8", errors="replace") as f:
            for k, line in enumerate(f, 1):
                if re.search(pattern, line):
                    hits.append(f"{os.path.relpath(path, ROOT)}:{k}: {line.strip()}")
    assert not hits, "stream of another priority or from torch's pool:\n" + "\n".join(hits)


def test_side_stream_is_made_with_default_priority():
    with open(os.path.join(PKG, "_lib.py")) as f:
        src = f.read()
    assert "hipStreamCreateWithFlags" in src and "ExternalStream" in src


def test_product_never_touches_the_oracle():
    hits = []
    for path in _product_sources():
        with open(path, encoding="utf-8", errors="replace") as f:
            for k, line in enumerate(f, 1):
                if re.search(r"\boracle\b|nx_oracle|ora_", line) and not line.lstrip().startswith(("#", "//", "*", "/*")):
                    hits.append(f"{os.path.relpath(path, ROOT)}:{k}: {line.strip()}")
    assert not hits, "\n".join(hits)
