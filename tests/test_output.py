"""Output writers (dem_mbd_cuda_amd.output / libdemio.so) against the text
the reference's own binary writes (oracle/_ref/dem_mbd_cpu on DATA1, frames 0
and 5, first 40 lines of each file; tests/golden/ref_output_text.npz): the
values parsed from the reference's file and written again must reproduce
the file byte for byte -- widths, precisions, separators, signed zeros and
the FI_pt columns of coupled runs.  The CSV rows follow the header literals
and %.17g conversions of the reference's CSV writers."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from dem_mbd_cuda_amd import output

TEXT = np.load(os.path.join(GOLDEN, "ref_output_text.npz"))


def rows(key):
    text = str(TEXT[key])
    return text, np.array([[float(v) for v in ln.split()] for ln in text.splitlines()])


def head(path, nlines):
    return "".join(open(path).read().splitlines(keepends=True)[:nlines])


@pytest.mark.parametrize("idx", [0, 5])
def test_particle_state_file(tmp_path, idx):
    text, a = rows(f"ph_{idx}")
    p = str(tmp_path / "ph.bt")
    output.write_particle_state(p, a[:, 0].astype(np.int32), a[:, 1].astype(np.int32), a[:, 2], a[:, 3],
                                a[:, 4:7], a[:, 7:10], a[:, 10:14], a[:, 14:17])
    assert head(p, len(a)) == text


@pytest.mark.parametrize("idx", [0, 5])
def test_particle_force_file_with_bodies(tmp_path, idx):
    text, a = rows(f"F_{idx}")
    n, cols = a.shape
    nb = (cols - 21) // 6
    assert nb == 2
    bp = np.stack([a[:, 21 + 6 * b:27 + 6 * b] for b in range(nb)])
    p = str(tmp_path / "F.bt")
    output.write_particle_force(p, a[:, 0:6], a[:, 6:12], a[:, 12:15], a[:, 15:21], bp)
    assert head(p, n) == text


def test_particle_force_file_without_bodies(tmp_path):
    _, a = rows("F_5")
    p = str(tmp_path / "F.bt")
    output.write_particle_force(p, a[:, 0:6], a[:, 6:12], a[:, 12:15], a[:, 15:21])
    got = open(p).read().splitlines()
    for ln in got:
        assert len(ln.split()) == 21 and ln == " ".join("%16.12e" % float(v) for v in ln.split())


@pytest.mark.parametrize("idx", [0, 5])
def test_body_force_file(tmp_path, idx):
    text, a = rows(f"FB_{idx}")
    p = str(tmp_path / "FB.bt")
    output.write_body_force(p, a[None, :, :])
    assert head(p, len(a)) == text


@pytest.mark.parametrize("idx", [0, 5])
def test_body_state_file(tmp_path, idx):
    text, a = rows(f"At_{idx}")
    p = str(tmp_path / "At.bt")
    t = a[0, 0]
    # the CPU binary writes no impulse columns; the coupled CUDA host adds six
    imp = a[:, 20:26] if a.shape[1] == 26 else None
    output.write_body_state(p, t, a[:, 1:4], a[:, 4:7], a[:, 7:10], a[:, 10:14], a[:, 14:20], imp)
    assert open(p).read() == text
    q = str(tmp_path / "At26.bt")
    output.write_body_state(q, t, a[:, 1:4], a[:, 4:7], a[:, 7:10], a[:, 10:14], a[:, 14:20], a[:, 14:20])
    for ln, ref in zip(open(q).read().splitlines(), text.splitlines()):
        assert ln == ref + ref[len(ref) - 6 * 26:]


def test_csv_rows():
    st = dict(regime_instant=np.array([[1, 2, 3, 4, 5, 6, 7]], np.uint64),
              regime_interval=np.array([[8, 9, 10, 11, 12, 13, 14]], np.uint64),
              micro_instant=np.arange(22, dtype=float)[None] * 0.1,
              micro_interval=np.arange(22, dtype=float)[None] * 1e-3,
              body_regime_interval_native_steps=5)
    r = output.regime_rows(st, 5, 5e-5)
    assert r == "native_pt_branch_v1,5,5.0000000000000002e-05,0,5,1,2,3,4,5,6,7,8,9,10,11,12,13,14\n"
    m = output.micro_rows(st, 5, 5e-5).rstrip("\n").split(",")
    assert m[:5] == ["native_pt_microstate_v1", "5", "5.0000000000000002e-05", "0", "5"]
    assert len(m) == 5 + 44 and m[5] == "0" and m[6] == "0.10000000000000001"
    assert output.REGIME_HEADER.count(",") == 18 and output.MICRO_HEADER.count(",") == 4 + 44
