"""Contact-history checkpoints on the HIP path: the file the GPU writes is
byte-identical to the oracle's after the same run, a restart from it (state
and history restored, first evaluation with dt = 0, expiry after the next
evaluation) follows the oracle's restart bit for bit, and a GPU load + save
reproduces the file."""
import numpy as np
import pytest

from dem_mbd_cuda_amd import host
from helpers import run_frames
from test_checkpoint import agitated_box
from test_gpu_parity import compare_frames

pytestmark = pytest.mark.gpu


def test_gpu_file_equals_oracle_file(hip, oracle, tmp_path):
    sc = agitated_box(n=400, steps=60)
    g, o = str(tmp_path / "g.bt"), str(tmp_path / "o.bt")
    for lib, path in ((hip, g), (oracle, o)):
        s = lib.solver()
        s.initialize(sc)
        host.run_pure(s, sc, steps=60, history_out=path)
        s.close()
    assert open(g, "rb").read() == open(o, "rb").read()


def test_restart_matches_oracle(hip, oracle, tmp_path):
    sc = agitated_box(n=400, steps=40)
    path = str(tmp_path / "o.bt")
    s = oracle.solver()
    s.initialize(sc)
    host.run_pure(s, sc, steps=40, history_out=path)
    s.close()

    def restart(lib):
        s = lib.solver()
        s.initialize(sc)
        frames = {}

        def cb(step, t, extra):
            frames[step + 1] = dict(state=s.state(), forces=s.forces(), stats=s.stats(), extra=extra,
                                    history=s.debug_history(), grid=s.debug_grid())

        host.run_pure(s, sc, steps=30, output_interval=10, on_frame=cb, history_in=path)
        assert s.restored_leapfrog()
        s.close()
        return frames

    compare_frames(restart(hip), restart(oracle))


def test_gpu_load_save_round_trip(hip, oracle, tmp_path):
    sc = agitated_box(n=300, steps=20)
    a, b = str(tmp_path / "a.bt"), str(tmp_path / "b.bt")
    s = oracle.solver()
    s.initialize(sc)
    host.run_pure(s, sc, steps=20, history_out=a)
    s.close()
    s = hip.solver()
    s.initialize(sc)
    s.load_history(a)
    s.save_history(b)
    s.close()
    assert open(a, "rb").read() == open(b, "rb").read()
