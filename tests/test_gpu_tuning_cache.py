"""The batches' tuning choices kept in a file (fp_tune_cache_save / _load).

Bar: a batch that tunes its contact lanes and launch shape writes them to
the file; after a reset and a load, a new batch of the same scenes runs
with the same choices from its first frame (no eager tuning frames, so
more graph replays), and both batches' results are bit-identical (no
choice changes any result).
"""
import json

import numpy as np
import pytest

from tests._gpu import log

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
SUB = 5
FRAMES = 24  # 4 contact-lane tuning frames, then the launch-shape candidates


@pytest.fixture(scope="module", autouse=True)
def _lib():
    from fluidphysics_amd import _abi
    _abi.load_library()


def test_tuning_cache_file_skips_tuning_frames(tmp_path):
    from fluidphysics_amd import Batch, from_scene, scenes, tune_cache_load, tune_cache_reset, tune_cache_save
    scs = [scenes.batched_scene(k, frames=FRAMES + 1) for k in range(70, 82)]
    log("case tuning cache: %d scenes, %d frames" % (len(scs), FRAMES))

    def run():
        ss = [from_scene(sc) for sc in scs]
        b = Batch()
        for s in ss:
            b.add(s)
        for _ in range(FRAMES):
            b.step(DT, SUB)
        pos = [s.read_positions().copy() for s in ss]
        inf = b.info()
        b.close()
        for s in ss:
            s.close()
        return inf, pos

    tune_cache_reset()
    try:
        inf1, pos1 = run()
        path = tmp_path / "tune.json"
        tune_cache_save(str(path))
        saved = json.loads(path.read_text())
        log("  tuned: lanes %d, staging %d, XCD run %d, %d graph launches; file %s"
            % (inf1.contact_lanes, inf1.density_stage, inf1.xcd_run, inf1.graph_launches, saved))
        assert inf1.shape_tuned == 1
        assert list(saved["contact_lanes"].values()) == [inf1.contact_lanes]
        assert len(saved["launch_shape"]) == 1
        tune_cache_reset()
        tune_cache_load(str(path))
        inf2, pos2 = run()
        log("  cached: lanes %d, staging %d, XCD run %d, %d graph launches"
            % (inf2.contact_lanes, inf2.density_stage, inf2.xcd_run, inf2.graph_launches))
        assert inf2.contact_lanes == inf1.contact_lanes
        assert (inf2.density_stage, inf2.xcd_run, inf2.shape_tuned) == (inf1.density_stage, inf1.xcd_run, 1)
        assert inf2.graph_launches > inf1.graph_launches
        for a, b in zip(pos1, pos2):
            assert np.array_equal(a, b)
    finally:
        tune_cache_reset()
