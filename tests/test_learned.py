"""What the trained warm-start network gains over plain GJK, on coherent hull pairs.

The network is trained on the GPU by the library's own kernels on plain GJK/EPA results, then
used as init_dir for contact (prediction 0) and distance on moving hull pairs, with GJK capped at
32 iterations and plain GJK on the cap.  Results must agree with plain GJK (the termination tests
are unchanged) and the GJK iteration count per distance query must drop."""
import pytest

import ncollide_amd as ncx
from ncollide_amd import learned

pytestmark = pytest.mark.gpu


def test_learned_warm_start_gains_over_plain(ctx):
    feat, y = learned.training_set(ctx, seed=51, nsamples=120_000, pool_hulls=8192)
    mlp, hist, _ = learned.train(ctx, feat, y, epochs=6)
    assert hist[-1] < 0.5 * hist[0]
    res = learned.evaluate_coherent(ctx, mlp, seed=52, npairs=8192, steps=4, placer_iters=16)
    p, l, o = res["plain"], res["learned"], res["oracle_dir"]
    print(f"distance GJK it/pair: plain {p['distance_gjk_iters_per_pair']:.2f}, learned "
          f"{l['distance_gjk_iters_per_pair']:.2f}, exact direction {o['distance_gjk_iters_per_pair']:.2f}; "
          f"contact GJK it/pair: plain {p['contact_gjk_iters_per_pair']:.2f}, learned "
          f"{l['contact_gjk_iters_per_pair']:.2f}; agreement {l['contact_agreement_with_plain']:.4f} / "
          f"{l['distance_agreement_with_plain']:.4f}")
    assert l["distance_gjk_iters_per_pair"] < 0.95 * p["distance_gjk_iters_per_pair"]
    assert l["contact_agreement_with_plain"] > 0.99
    assert l["distance_agreement_with_plain"] > 0.99
