"""The batched workload itself (64 joins, N log-uniform 1e4..5e5, d in
{2,3,4,6,8}, uniform or Exp(40) data, 8..128 neighbours/pt, seed 5) in one
graph-recorded launch per pass, FP32+recheck and FP64, every instance
against the CPU oracle.

The comparison is exact and order-free per instance: |S|, the sum over
pairs of (row + 1) * neighbour ID and the sum of squared IDs (computed on
the device from the cell-order table, on the host from the oracle CSR in
original-ID order), plus a row-by-row check of a few rows per instance."""
import sys

import numpy as np
import pytest

import oracle as O
import sjamd
from sjamd import datasets as D

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fp_host(off, ids):
    rows = np.repeat(np.arange(off.size - 1, dtype=np.int64), np.diff(off.astype(np.int64)))
    v = ids.astype(np.int64)
    return int(off[-1]), int((v * (rows + 1)).sum()), int((v * v).sum())


def _fp_device(torch, o, i, r, n, tot):
    off = sjamd.device_copy(o, n + 1, torch.int64)
    ids = sjamd.device_copy(i, tot, torch.int32).to(torch.int64)
    rid = sjamd.device_copy(r, n, torch.int32).to(torch.int64)
    rr = torch.repeat_interleave(rid, off[1:] - off[:-1])
    return int(off[-1].item() - off[0].item()), int((ids * (rr + 1)).sum().item()), int((ids * ids).sum().item())


@pytest.fixture(scope="module")
def batch_and_refs():
    insts = D.multi_join_instances()
    batch = [(D.instance_points(i), i["d"], D.multi_eps(j)) for j, i in enumerate(insts)]
    refs = []
    for p, k, e in batch:
        off, ids = O.join(p, k, e, O.LID, sort_rows=False)
        refs.append((_fp_host(off, ids), off, ids))
    return batch, refs


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_many64_every_instance_matches_oracle(ctx, batch_and_refs, prec):
    import torch

    batch, refs = batch_and_refs
    _say(f"[many64] {prec}: {len(batch)} instances, {sum(p.shape[0] for p, _, _ in batch)} points")
    precision = sjamd.FP32_RECHECK if prec == "fp32" else sjamd.FP64
    m = ctx.many(batch, form="bunicomp", precision=precision, schedule=sjamd.WORKQUEUE)
    m.run()  # a replay after the one the plan ran
    torch.cuda.synchronize()
    rng = np.random.default_rng(5)
    for i, ((p, k, e), (fp, off, ids)) in enumerate(zip(batch, refs)):
        n = p.shape[0]
        o, d, r, tot = m.device_table(i)
        assert tot == fp[0], f"instance {i}: |S| {tot} vs oracle {fp[0]}"
        assert _fp_device(torch, o, d, r, n, tot) == fp, f"instance {i}"
        # a few rows, element by element
        goff = sjamd.download(o, n + 1, np.uint64).astype(np.int64)
        grow = sjamd.download(r, n, np.uint32).astype(np.int64)
        for rr in rng.integers(0, n, 3).tolist():
            pid = grow[rr]
            got = np.sort(sjamd.download(d + 4 * int(goff[rr]), int(goff[rr + 1] - goff[rr]), np.int32))
            want = np.sort(ids[int(off[pid]):int(off[pid + 1])])
            np.testing.assert_array_equal(got, want, err_msg=f"instance {i} point {pid}")
    assert m.stats()["total_pairs"] == sum(fp[0] for fp, _, _ in refs)
    m.free()
