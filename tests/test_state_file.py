"""The reference's binary state file, field by field.

The layouts below restate read_state.cpp:5-23 (the 124-byte header: ints and
doubles, unpadded) and read_state.cpp:34-67 / write_state.cpp:27-62 (23
doubles per particle) as numpy structured types, independently of the
engine's C++ reader.  A file written from them must land in the engine's
arrays where the reference's CUDAsph::read_state puts it (range = 2 h, mu in
units of amu, ...), and a file the engine writes must parse back to the same
fields, with aa = bb = cc = -1 as write_state.cpp:55-59 writes them.
"""
import numpy as np
import pytest

AMU = 1.66054e-24   # read_state.cpp:61, write_state.cpp:53

HEADER = np.dtype([("n_bodies", "<i4"), ("nnopt", "<i4"), ("hmin", "<f8"), ("hmax", "<f8"), ("sep0", "<f8"),
                   ("t_end", "<f8"), ("dt_out", "<f8"), ("snapshot_number", "<i4"), ("iteration_number", "<i4"),
                   ("global_time", "<f8"), ("nav", "<i4"), ("alpha", "<f8"), ("beta", "<f8"), ("eta2", "<f8"),
                   ("ngr", "<i4"), ("nrelax", "<i4"), ("trelax", "<f8"), ("dt", "<f8"), ("omega2", "<f8")])
RECORD = np.dtype([(k, "<f8") for k in ("x", "y", "z", "m", "h", "rho", "vx", "vy", "vz", "ax", "ay", "az", "u",
                                        "udot", "gx", "gy", "gz", "gpot", "mu", "aa", "bb", "cc", "divv")])


def test_layouts_have_the_reference_sizes():
    assert HEADER.itemsize == 124 and RECORD.itemsize == 23 * 8


def test_header_count_on_the_host(tmp_path):
    """sph_state_file_count reads N from a header written by the layout above
    and refuses a header cut short (host code, no GPU)"""
    import cudasph_amd as cs
    h = np.zeros(1, HEADER)
    h["n_bodies"], h["nnopt"], h["dt"] = 12345, 58, 0.5
    good, short = tmp_path / "h.bin", tmp_path / "short.bin"
    good.write_bytes(h.tobytes())
    short.write_bytes(h.tobytes()[:-1])
    assert cs.lib().sph_state_file_count(str(good).encode()) == 12345
    assert cs.lib().sph_state_file_count(str(short).encode()) < 0


def make_file(path, n, seed):
    rng = np.random.default_rng(seed)
    h = np.zeros(1, HEADER)
    h["n_bodies"], h["nnopt"], h["snapshot_number"], h["iteration_number"] = n, 58, 7, 1234
    h["global_time"], h["dt"], h["alpha"], h["beta"] = 3.25, 0.0078125, 1.0, 2.0
    h["hmin"], h["hmax"], h["sep0"], h["t_end"], h["dt_out"] = 1e-3, 0.5, 5.0, 10.0, 0.4
    h["nav"], h["eta2"], h["ngr"], h["nrelax"], h["trelax"], h["omega2"] = 3, 0.01, 1, 0, 2.5, 0.125
    r = np.zeros(n, RECORD)
    for name in RECORD.names:
        # float32-representable values, so every form holds them exactly
        r[name] = rng.uniform(0.5, 2.0, n).astype(np.float32)
    r["mu"] = (rng.uniform(0.5, 1.5, n).astype(np.float32)).astype(np.float64) * AMU
    r["aa"] = r["bb"] = r["cc"] = -1.0
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(r.tobytes())
    return h[0], r


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "fp64", "mixed"])
def test_state_file_fields(cs, tmp_path, form):
    n = 300
    print(f"state file form={form} n={n}", flush=True)
    src = tmp_path / "in.bin"
    h, r = make_file(src, n, seed=3)
    e = cs.Engine(form, cs.preset_params(1, n))
    e.read_state(src)
    pos, vel, hyd = e.get("pos")[:n], e.get("vel")[:n], e.get("hydro")[:n]
    bd, gr, dd = e.get("bdots")[:n], e.get("grav")[:n], e.get("dots")[:n]
    expect = [(pos[:, 0], r["x"]), (pos[:, 1], r["y"]), (pos[:, 2], r["z"]), (pos[:, 3], 2 * r["h"]),
              (vel[:, 0], r["vx"]), (vel[:, 1], r["vy"]), (vel[:, 2], r["vz"]), (vel[:, 3], r["m"]),
              (hyd[:, 0], r["rho"]), (hyd[:, 2], r["u"]),
              (bd[:, 0], r["ax"]), (bd[:, 1], r["ay"]), (bd[:, 2], r["az"]), (bd[:, 3], r["udot"]),
              (gr[:, 0], r["gx"]), (gr[:, 1], r["gy"]), (gr[:, 2], r["gz"]), (gr[:, 3], r["gpot"]),
              (dd[:, 0], r["divv"])]
    for k, (got, want) in enumerate(expect):
        assert np.array_equal(got.astype(np.float64), want), f"field {k} misplaced"
    np.testing.assert_allclose(hyd[:, 3], r["mu"] / AMU, rtol=1e-7)
    assert e.scalar("dt") == h["dt"]
    assert e.scalar("global_time") == h["global_time"]
    assert e.scalar("iteration") == h["iteration_number"]
    if form == "mixed":   # the FP64 copies the mixed form integrates
        assert np.array_equal(e.get("pos64")[:n, :3], np.stack([r["x"], r["y"], r["z"]], 1))
        assert np.array_equal(e.get("vel64")[:n], np.stack([r["vx"], r["vy"], r["vz"], r["m"]], 1))
        assert np.array_equal(e.get("u64")[:n], r["u"])
    # and back out through the engine's writer
    out = tmp_path / "out.bin"
    e.write_state(out)
    raw = out.read_bytes()
    assert len(raw) == HEADER.itemsize + n * RECORD.itemsize
    h2 = np.frombuffer(raw[:HEADER.itemsize], HEADER)[0]
    r2 = np.frombuffer(raw[HEADER.itemsize:], RECORD)
    assert h2["n_bodies"] == n and h2["iteration_number"] == h["iteration_number"]
    assert h2["global_time"] == h["global_time"] and h2["dt"] == h["dt"]
    for name in RECORD.names:
        if name == "mu":
            np.testing.assert_allclose(r2[name], r[name], rtol=1e-7)
        else:
            assert np.array_equal(r2[name], r[name]), name
    e.close()


@pytest.mark.gpu
def test_state_file_of_another_size_is_refused(cs, tmp_path):
    src = tmp_path / "in.bin"
    make_file(src, 200, seed=4)
    e = cs.Engine("plain", cs.preset_params(1, 300))
    with pytest.raises(RuntimeError, match="differs"):
        e.read_state(src)
    e.close()


@pytest.mark.gpu
def test_output_and_header_follow_the_reference(cs, tmp_path):
    """output.cpp:3-15 with dt_out = 0.4, snapshot_number = 7 and
    global_time = 3.25: 3.25 > 7 * 0.4 and then > 8 * 0.4, so two calls write
    out0008.bin and out0009.bin, a third only restart.bin.  Every header field
    the run does not advance is written back as read."""
    n = 300
    print(f"state file output n={n}", flush=True)
    src = tmp_path / "in.bin"
    h, r = make_file(src, n, seed=5)
    e = cs.Engine("plain", cs.preset_params(1, n))
    e.read_state(src)
    out = tmp_path / "run"
    out.mkdir()
    assert e.output(out) == 8
    assert e.output(out) == 9
    assert e.output(out) == 0
    assert sorted(p.name for p in out.iterdir()) == ["out0008.bin", "out0009.bin", "restart.bin"]
    for name, snap in (("out0008.bin", 8), ("out0009.bin", 9), ("restart.bin", 9)):
        raw = (out / name).read_bytes()
        h2 = np.frombuffer(raw[:HEADER.itemsize], HEADER)[0]
        for field in HEADER.names:
            want = snap if field == "snapshot_number" else h[field]
            assert h2[field] == want, (name, field)
        assert np.array_equal(np.frombuffer(raw[HEADER.itemsize:], RECORD)["x"], r["x"])
    e.close()


@pytest.mark.gpu
def test_output_without_a_state_file_writes_the_restart_only(cs, tmp_path):
    e = cs.make_engine("plain", 1, 2000, 3)
    assert e.output(tmp_path) == 0
    assert sorted(p.name for p in tmp_path.iterdir()) == ["restart.bin"]
    h2 = np.frombuffer((tmp_path / "restart.bin").read_bytes()[:HEADER.itemsize], HEADER)[0]
    assert h2["n_bodies"] == 2000 and h2["alpha"] == 1.0 and h2["beta"] == 2.0
    e.close()
