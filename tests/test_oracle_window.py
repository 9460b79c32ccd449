"""The oracle's row-window mode against its whole-image mode.

The full-size GPU parity tests compare a band of pixel rows of a large scene
with the oracle, which renders only that band (oracle_render_rows) and runs
the backward pass only over the samples that can see a band-limited d_render
(oracle_render_backward_rows).  The skipped work must contribute nothing:
here the band results are checked bit-for-bit against the whole-image oracle
on small scenes of every filter and sampling mode."""
import numpy as np
import pytest

from pydvg import scenes
from pydvg.scene import GRAD_FIELDS, SceneDesc


def _cases():
    return {
        "box_4x4": (scenes.cubic_paths_scene(48, 40, 12, seed=3), scenes.config1()[1]),
        "tent_mixed": (scenes.mixed_scene(40, 48, 24, seed=4), scenes.config2()[1]),
        "prefiltered": scenes.prefiltered_scene(48, 48, 16, 6, seed=5, step=6.0, radius=(3.0, 12.0)),
    }


@pytest.mark.parametrize("name", list(_cases().keys()))
@pytest.mark.parametrize("band", [(0, 7), (17, 26), (33, 99)])
def test_render_rows_equal_whole_image(oracle, name, band):
    sc, opts = _cases()[name]
    desc = SceneDesc(sc)
    full = oracle.render(desc, opts).reshape(sc.height, sc.width, 4)
    y0, y1 = band
    rows = oracle.render_rows(desc, opts, y0, y1).reshape(-1, sc.width, 4)
    assert rows.shape[0] == min(y1, sc.height) - y0
    assert np.array_equal(rows, full[y0:y1])


@pytest.mark.parametrize("name", list(_cases().keys()))
@pytest.mark.parametrize("band", [(0, 5), (19, 27)])
def test_backward_rows_equal_whole_image(oracle, name, band):
    sc, opts = _cases()[name]
    desc = SceneDesc(sc)
    y0, y1 = band
    rng = np.random.default_rng(11)
    dr = np.zeros((sc.height, sc.width, 4), np.float32)
    dr[y0:y1] = rng.normal(size=(y1 - y0, sc.width, 4)).astype(np.float32)
    g_full = oracle.render_backward(desc, opts, dr.reshape(-1))
    g_rows = oracle.render_backward_rows(desc, opts, dr.reshape(-1), y0, y1)
    nonzero = 0
    for k in GRAD_FIELDS:
        assert np.array_equal(g_rows[k], g_full[k]), k
        nonzero += int(np.count_nonzero(g_full[k]))
    assert nonzero > 0
