"""The per-level launch-configuration table (csrc/tune.cpp), host only: the choice follows the
level's size histogram (median Asn rows for the sparsification knobs, median pivot for the
trsm chunks) bucket by bucket, per factor precision."""
import pytest

import spand_public_amd as S


def choose(esz, **shape):
    return S.tune_choose(esz, shape)


def test_register_path_off_for_levels_of_small_clusters():
    for esz in (4, 8):
        assert choose(esz, rows_med=20, cols_med=190, n_spars=5000)["rrqr_reg"] == 0
        assert choose(esz, rows_med=31, cols_med=299)["rrqr_reg"] == 0
        assert choose(esz, rows_med=21, cols_med=322)["rrqr_reg"] == 1  # wider: the register path
        assert choose(esz, rows_med=32, cols_med=100)["rrqr_reg"] == 1
        assert choose(esz, rows_med=90, cols_med=800)["rrqr_reg"] == 1


def test_fp64_step_waves_follow_the_cluster_rows():
    assert choose(8, rows_med=40)["step_waves"] == 0
    assert choose(8, rows_med=64)["step_waves"] == 8
    assert choose(8, rows_med=159)["step_waves"] == 8
    assert choose(8, rows_med=160)["step_waves"] == 16
    assert choose(8, rows_med=399)["step_waves"] == 16
    assert choose(8, rows_med=400)["step_waves"] == 0
    for rows in (20, 100, 200, 500):  # FP32: the slab-count rule everywhere
        assert choose(4, rows_med=rows)["step_waves"] == 0


def test_defaults_and_source():
    c = choose(4, rows_med=100, elim_med=300)
    assert c["small_elems"] == 10000 and c["trsm_cw"] == 0 and c["source"] == 0


def test_bad_element_size_is_refused():
    with pytest.raises(S.SpandError):
        choose(2, rows_med=10)
