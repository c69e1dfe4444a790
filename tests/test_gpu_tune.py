"""Autotuner: launch knobs change speed, never clusters."""
import os

import numpy as np
import pytest

import gsdr
import oracle_ffi as O

pytestmark = pytest.mark.gpu


def _oracle(tmp_path, recs, t):
    fa = str(tmp_path / "in.fasta")
    recs.write_fasta(fa)
    ocl, _, _ = O.cluster_file(fa, t)
    r = gsdr.Records.read_fasta(fa, length_order=True)
    return r, ocl


@pytest.mark.parametrize("law,n,t", [(1, 3000, 0.95), (4, 4000, 0.80)])
def test_every_knob_setting_gives_the_oracle_clusters(engine, tmp_path, law, n, t):
    r, ocl = _oracle(tmp_path, gsdr.Records.generate(law, 1000 + law, n), t)
    store = engine.load(r, gsdr.auto_word_length(t))
    try:
        settings = [(2, 64, 8, 4, 1, 8), (0, 64, 8, 0, 0, 64), (8, 1, 1, 1, 1, 1), (3, 16, 16, 2, 0, 32),
                    (1, 7, 2, 8, 1, 2), (4, 64, -1, -1, -1, 64)]
        for swap, spec, fb, ab, fork, lf in settings:
            tu = gsdr.Tuning(swap_ratio=swap, spec_max=spec, filter_blocks=fb, align_blocks=ab, fork_shapes=fork,
                             list_factor=lf)
            engine.set_tuning(tu)
            got = engine.tuning()
            assert (got.swap_ratio, got.spec_max, got.filter_blocks, got.align_blocks, got.fork_shapes) == (
                swap, spec, fb, ab, fork)
            for mode in (gsdr.MODE_PLAIN, gsdr.MODE_BATCHED):
                cl, st = store.cluster(t, store.k, mode=mode, tune=True)
                assert np.array_equal(cl, ocl), (swap, spec, fb, ab, fork, mode)
                assert st.pairs == st.rej_length + st.rej_comp + st.rej_word + st.aligned
    finally:
        engine.set_tuning(None)
        store.free()


# shape masks over the twelve tile shapes (8x1 16x1 32x1 64x1 64x2 64x3 64x4
# 64x6 64x8 64x10 64x12 64x16): powers of two only, no tile narrower than a
# wave, everything but the 64x2 tile
ALL = (1 << 12) - 1
POW2 = ALL & ~((1 << 5) | (1 << 7) | (1 << 9) | (1 << 10))
NO_NARROW = ALL & ~7


@pytest.mark.parametrize("law,n,t", [(4, 4000, 0.95), (2, 3000, 0.90), (1, 2000, 0.80)])
def test_shape_tables_give_the_oracle_clusters(engine, tmp_path, law, n, t):
    """Per-length-bin tile shapes and the window's rebase period change which
    kernel runs a pair and how wide its window is, never a decision."""
    r, ocl = _oracle(tmp_path, gsdr.Records.generate(law, 1100 + law, n), t)
    store = engine.load(r, gsdr.auto_word_length(t))
    try:
        tables = [([POW2] * 4, [0] * 4, 0), ([NO_NARROW] * 4, [1] * 4, 1), ([0] * 4, [3] * 4, 8),
                  ([ALL & ~(1 << 4), POW2, NO_NARROW, 0], [2, 0, 1, 3], 16), ([0] * 4, [0] * 4, 1)]
        for masks, biases, period in tables:
            tu = gsdr.Tuning(swap_ratio=4, spec_max=64, filter_blocks=8, align_blocks=4, fork_shapes=1,
                             list_factor=64, win_period=period)
            for b in range(4):
                tu.shape_mask[b] = masks[b]
                tu.shape_bias[b] = biases[b]
            engine.set_tuning(tu)
            got = engine.tuning()
            assert list(got.shape_mask) == masks and list(got.shape_bias) == biases and got.win_period == period
            for mode in (gsdr.MODE_PLAIN, gsdr.MODE_BATCHED):
                cl, st = store.cluster(t, store.k, mode=mode, tune=True)
                assert np.array_equal(cl, ocl), (masks, biases, period, mode)
    finally:
        engine.set_tuning(None)
        store.free()


def test_autotune_then_tuned_calls_match(engine, tmp_path):
    r, ocl = _oracle(tmp_path, gsdr.Records.generate(2, 1002, 3000), 0.90)
    store = engine.load(r, 5)
    try:
        tu = store.autotune(0.90, 5, max_trials=6)
        assert 1 <= tu.trials <= 6
        assert tu.ms_default > 0 and 0 < tu.ms_tuned <= tu.ms_default
        got = engine.tuning()
        assert (got.swap_ratio, got.spec_max, got.filter_blocks, got.align_blocks) == (
            tu.swap_ratio, tu.spec_max, tu.filter_blocks, tu.align_blocks)
        assert list(got.shape_mask) == list(tu.shape_mask) and list(got.shape_bias) == list(tu.shape_bias)
        assert got.win_period == tu.win_period
        cl, _ = store.cluster(0.90, 5, mode=gsdr.MODE_BATCHED, tune=True)
        assert np.array_equal(cl, ocl)
    finally:
        engine.set_tuning(None)
        store.free()
    d = engine.tuning()
    assert (d.swap_ratio, d.spec_max, d.filter_blocks, d.align_blocks, d.fork_shapes, d.list_factor) == (4, 64, -1, -1, -1, 64)


def test_tuning_out_of_range_is_rejected(engine):
    for bad in [dict(swap_ratio=-1), dict(spec_max=0), dict(spec_max=65), dict(filter_blocks=0), dict(filter_blocks=-2),
                dict(align_blocks=65), dict(align_blocks=-2), dict(fork_shapes=2), dict(fork_shapes=-2), dict(list_factor=0),
                dict(win_period=17), dict(win_period=-1), dict(win_period=3), dict(shape_mask=4096), dict(shape_bias=4),
                dict(split_rows=63), dict(split_rows=32769), dict(split_rows=-2)]:
        tu = gsdr.Tuning(swap_ratio=4, spec_max=64, filter_blocks=8, align_blocks=4, fork_shapes=1, list_factor=64)
        for k, v in bad.items():
            if k in ("shape_mask", "shape_bias"):
                getattr(tu, k)[2] = v
            else:
                setattr(tu, k, v)
        with pytest.raises(gsdr.GsdrError):
            engine.set_tuning(tu)
    d = engine.tuning()
    assert (d.swap_ratio, d.spec_max, d.filter_blocks, d.align_blocks, d.fork_shapes, d.list_factor) == (4, 64, -1, -1, -1, 64)


@pytest.mark.parametrize("law,n,t", [(4, 5000, 0.95), (4, 4000, 0.80), (2, 3000, 0.99), (1, 2000, 0.90)])
def test_split_pairs_give_the_oracle_clusters(engine, tmp_path, law, n, t):
    """Pairs of both phases aligned by two waves (forward and backward halves
    joined at the middle row) from 64 rows up, from 300 rows up, and never:
    the oracle's clusters."""
    r, ocl = _oracle(tmp_path, gsdr.Records.generate(law, 1200 + law, n), t)
    store = engine.load(r, gsdr.auto_word_length(t))
    try:
        for rows in (64, 300, 0):
            engine.set_tuning(gsdr.Tuning(swap_ratio=4, spec_max=64, filter_blocks=8, align_blocks=4, fork_shapes=1,
                                          list_factor=64, split_rows=rows))
            assert engine.tuning().split_rows == rows
            cl, st = store.cluster(t, store.k, mode=gsdr.MODE_BATCHED, tune=True)
            assert np.array_equal(cl, ocl), rows
            assert st.pairs == st.rej_length + st.rej_comp + st.rej_word + st.aligned
        # concurrent and serial launch of the split kernel beside the shapes
        engine.set_tuning(gsdr.Tuning(swap_ratio=4, spec_max=64, filter_blocks=8, align_blocks=4, fork_shapes=0,
                                      list_factor=64, split_rows=64))
        cl, _ = store.cluster(t, store.k, mode=gsdr.MODE_BATCHED, tune=True)
        assert np.array_equal(cl, ocl)
    finally:
        engine.set_tuning(None)
        store.free()
