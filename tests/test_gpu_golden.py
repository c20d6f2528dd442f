"""The HIP join reproduces every golden fixture exactly, for every
traversal form, both precisions and both schedules."""
import glob
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_hip_join_reproduces_fixture(ctx, path):
    import sjamd

    with np.load(path, allow_pickle=False) as z:
        pts, k, eps, off, ids = z["points"], int(z["k"]), float(z["eps"]), z["offsets"], z["ids"]
    print(f"[golden] {os.path.basename(path)} n={pts.shape[0]}", flush=True)
    sys.stdout.flush()
    for form in ("full", "unicomp", "bunicomp", "lid"):
        for prec in (sjamd.FP64, sjamd.FP32_RECHECK):
            for sched in (sjamd.STRIDED, sjamd.WORKQUEUE):
                r = ctx.self_join(pts, k, eps, form=form, precision=prec, schedule=sched,
                                  sort_rows=True)
                goff, gids = r.csr()
                np.testing.assert_array_equal(goff, off, err_msg=f"{form} {prec} {sched}")
                np.testing.assert_array_equal(gids, ids, err_msg=f"{form} {prec} {sched}")
                r.free()
