"""The GEMM autotuner's host logic (CPU only): workload shapes, buckets and
the committed table's format."""


def test_autotune_shape_buckets():
    """The autotuner's workload shapes: ZC^T (N x N_mu x M, kind 1) and the FP64
    solve (N_mu x N x N_mu, kind 0) of every batch cell and configuration,
    grouped by (ceil log2 m, n, k, flags)."""
    from thcpp_amd import autotune, synthetic
    batch = synthetic.make_batch(count=2, seed=0, max_it=2)
    shapes = autotune.workload_shapes(batch)
    assert len(shapes) == 2 * len(batch) + 2 * len(autotune.CONFIG_WORKLOADS)
    s = batch[0]
    assert ("batch32", s.N, s.K, s.M, 1) in shapes and ("batch32", s.K, s.N, s.K, 0) in shapes
    assert ("diamond64_n64_k2048", 262144, 2048, 128, 1) in shapes
    assert autotune.bucket(262144, 2048, 128, 1) == (18, 11, 7, 5)
    assert autotune.bucket(2048, 262144, 2048, 0) == (11, 18, 11, 0)
    assert autotune.clog2(1) == 0 and autotune.clog2(1025) == 11


def test_committed_tuning_table_is_well_formed():
    import json
    import os
    from thcpp_amd.api import TUNING_FILE
    if not os.path.exists(TUNING_FILE):
        return
    with open(TUNING_FILE) as f:
        t = json.load(f)
    for e in t["entries"]:
        assert len(e) == 5 and 1 <= e[4] <= 5 and 0 <= e[3] < 16
        assert all(0 <= v <= 40 for v in e[:3])
