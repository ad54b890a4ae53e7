"""Contact-history checkpoints (DMHIST2 / DMHIST1, the layout of the
reference's CudaDemSolver::saveContactHistory / loadContactHistory,
cuda_backend_cuda.cu:545-770) on the CPU oracle: the byte layout field by
field against the solver's own state and history, the closed-form file size,
round trips, the restart semantics (leapfrog state restored, history kept
through the next evaluation) and every rejection the reference makes."""
import os
import struct

import numpy as np
import pytest

from dem_mbd_cuda_amd import host, scene
from dem_mbd_cuda_amd._abi import DemError


def agitated_box(n=300, seed=4, steps=40):
    side = 0.02 * (n / 300.0) ** (1.0 / 3.0)
    sc = scene.box_scene(n=n, seed=seed, box=(side, side, 1.5 * side), steps=steps)
    rng = np.random.RandomState(seed)
    sc.vel = rng.uniform(-0.5, 0.5, (n, 3))
    sc.angvel = rng.uniform(-50, 50, (n, 3))
    return sc


def run_and_save(lib, sc, path, steps):
    s = lib.solver()
    s.initialize(sc)
    host.run_pure(s, sc, steps=steps, history_out=path)
    return s


def parse(path):
    b = open(path, "rb").read()
    assert b[:8] == b"DMHIST2\0"
    version, endian, owners, comps, s0, s1, s2, topo, rec = struct.unpack_from("<9I", b, 8)
    off = 8 + 36
    comp = []
    for _ in range(comps):
        cid, oid = struct.unpack_from("<2i", b, off)
        geo = struct.unpack_from("<4d", b, off + 8)
        comp.append((cid, oid) + geo)
        off += 40
    own = []
    for _ in range(owners):
        (oid,) = struct.unpack_from("<i", b, off)
        st = struct.unpack_from("<23d", b, off + 4)
        own.append((oid, st))
        off += 188
    keys = {}
    for name, slots in (("pp", s0), ("pw", s1), ("pt", s2)):
        keys[name] = np.frombuffer(b, "<u8", comps * slots, off).reshape(comps, slots)
        off += 8 * comps * slots
    vals = {}
    for name, slots in (("pp", s0), ("pw", s1), ("pt", s2)):
        vals[name] = np.frombuffer(b, "<f8", comps * slots * 9, off).reshape(comps, slots, 9)
        off += 72 * comps * slots
    assert off == len(b)
    return dict(version=version, endian=endian, owners=owners, comps=comps, slots=(s0, s1, s2), topo=topo,
                rec=rec, comp=comp, own=own, keys=keys, vals=vals, size=len(b))


def test_layout_matches_state_and_history(oracle, tmp_path):
    sc = agitated_box()
    path = str(tmp_path / "h.bt")
    s = run_and_save(oracle, sc, path, 40)
    f = parse(path)
    n = sc.n
    assert (f["version"], f["endian"], f["owners"], f["comps"]) == (2, 0x01020304, n, n)
    assert f["slots"] == (64, 16, 64) and (f["topo"], f["rec"]) == (40, 188)
    assert f["size"] == 44 + 40 * n + 188 * n + 8 * n * 144 + 72 * n * 144
    for i, c in enumerate(f["comp"]):
        assert c == (int(sc.comp_id[i]), int(sc.pid[i]), float(sc.comp_radius[i]), 0.0, 0.0, 0.0)
    st = s.state()
    own = np.array([o[1] for o in f["own"]])
    assert [o[0] for o in f["own"]] == [int(v) for v in sc.pid]
    assert np.array_equal(own[:, 0:3], st["pos"]) and np.array_equal(own[:, 3:6], st["vel"])
    assert np.array_equal(own[:, 6:10], st["quat"]) and np.array_equal(own[:, 10:13], st["angvel"])
    h = s.debug_history()
    assert np.array_equal(f["keys"]["pp"], h["pp_keys"]) and np.array_equal(f["keys"]["pw"], h["pw_keys"])
    assert np.array_equal(f["vals"]["pp"], h["pp_values"]) and np.array_equal(f["vals"]["pw"], h["pw_values"])
    assert (f["keys"]["pp"] != 0).sum() > 0
    s.close()


def test_round_trip_is_byte_identical_and_restores_leapfrog(oracle, tmp_path):
    sc = agitated_box()
    a, b = str(tmp_path / "a.bt"), str(tmp_path / "b.bt")
    run_and_save(oracle, sc, a, 30).close()
    s = oracle.solver()
    s.initialize(sc)
    assert not s.restored_leapfrog()
    s.load_history(a)
    assert s.restored_leapfrog()
    s.save_history(b)
    assert open(a, "rb").read() == open(b, "rb").read()
    s.close()


def test_restart_keeps_history_through_next_evaluation(oracle, tmp_path):
    """After a load the next evaluation runs with dt = 0 and expires nothing;
    the one after expires what that evaluation did not see."""
    sc = agitated_box(steps=30)
    a = str(tmp_path / "a.bt")
    s0 = run_and_save(oracle, sc, a, 30)
    keys0 = s0.debug_history()["pp_keys"]
    s0.close()
    s = oracle.solver()
    s.initialize(sc)
    s.load_history(a)
    s.compute_forces(collect=True)
    k1 = s.debug_history()["pp_keys"]
    assert np.array_equal(k1[keys0 != 0], keys0[keys0 != 0])  # nothing expired
    s.close()


def test_rejections(oracle, tmp_path):
    sc = agitated_box(n=120, steps=5)
    a = str(tmp_path / "a.bt")
    run_and_save(oracle, sc, a, 5).close()
    raw = open(a, "rb").read()

    def load(data, scn=sc, msg=None):
        p = str(tmp_path / "x.bt")
        open(p, "wb").write(data)
        s = oracle.solver()
        s.initialize(scn)
        try:
            with pytest.raises(DemError, match=msg):
                s.load_history(p)
        finally:
            s.close()

    load(raw[:-1], msg="truncated CUDA contact-history checkpoint")
    load(raw + b"\0", msg="oversized CUDA contact-history checkpoint")
    load(b"DMHIST3\0" + raw[8:], msg="unsupported contact-history checkpoint")
    bad = bytearray(raw)
    struct.pack_into("<I", bad, 12, 0x04030201)
    load(bytes(bad), msg="incompatible DMHIST2 header")
    other = sc.copy()
    other.comp_radius = other.comp_radius.copy()
    other.comp_radius[3] *= 1.01
    other.radius = other.comp_radius.copy()
    load(raw, scn=other, msg="DMHIST2 component topology mismatch")
    ids = sc.copy()
    ids.pid = ids.pid.copy()
    ids.pid[7] += 100000
    ids.comp_id = ids.comp_id.copy()
    load(raw, scn=ids, msg="DMHIST2 component topology mismatch")
    with pytest.raises(DemError, match="cannot open CUDA contact-history checkpoint"):
        s = oracle.solver()
        s.initialize(sc)
        s.load_history(str(tmp_path / "missing.bt"))


def test_dmhist1_legacy_only(oracle, tmp_path):
    sc = agitated_box(n=100, steps=5)
    a = str(tmp_path / "a.bt")
    run_and_save(oracle, sc, a, 5).close()
    f = parse(a)
    n = sc.n
    v1 = b"DMHIST1\0" + struct.pack("<5I", 1, n, 64, 16, 64)
    for name in ("pp", "pw", "pt"):
        v1 += f["keys"][name].tobytes()
    for name in ("pp", "pw", "pt"):
        v1 += f["vals"][name].tobytes()
    p = str(tmp_path / "v1.bt")
    open(p, "wb").write(v1)
    s = oracle.solver()
    s.initialize(sc)
    s.load_history(p)
    assert not s.restored_leapfrog()
    assert np.array_equal(s.debug_history()["pp_keys"], f["keys"]["pp"])
    s.close()
    # not the legacy sphere topology: component ids differ from particle ids
    other = sc.copy()
    other.comp_id = other.comp_id + 1000
    s = oracle.solver()
    s.initialize(other)
    with pytest.raises(DemError, match="legacy DMHIST1 checkpoint is forbidden for clump topology"):
        s.load_history(p)
    s.close()
