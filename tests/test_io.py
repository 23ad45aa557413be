"""Host I/O (libvolprim_io.so): PLY loading with the reference's conventions,
RGBE round trips, EXR layout (B,G,R float, vertically flipped) -- CPU only."""
import ctypes
import os
import struct

import numpy as np
import pytest

from volprim_amd import scenes, vpio


def test_io_library_exports_declared_symbols():
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(root, "include", "vp_io.h")).read(), flags=re.S)
    decl = sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(vp_io_\w+)\s*\(", text, flags=re.M)))
    assert decl == sorted(vpio.EXPORTED)
    lib = ctypes.CDLL(vpio.LIB_PATH)
    for name in decl:
        assert hasattr(lib, name)


def _write_ply(path, rows_props, names=None, fmt="binary_little_endian", extra_elem=False):
    names = names or ["x", "y", "z", "rot_0", "rot_1", "rot_2", "rot_3", "scale_0", "scale_1", "scale_2",
                      "albedo_0", "albedo_1", "albedo_2", "sigma_t_0"]
    with open(path, "wb") as f:
        hdr = f"ply\nformat {fmt} 1.0\ncomment test\n"
        if extra_elem:
            hdr += "element meta 2\nproperty uchar a\nproperty float b\n"
        hdr += f"element vertex {len(rows_props)}\n" + "".join(f"property float {n}\n" for n in names)
        hdr += "element face 0\nproperty list uchar int vertex_indices\nend_header\n"
        f.write(hdr.encode())
        if extra_elem:
            f.write(struct.pack("<BfBf", 1, 2.0, 3, 4.0))
        for r in rows_props:
            if fmt == "ascii":
                f.write((" ".join(repr(float(v)) for v in r) + "\n").encode())
            else:
                f.write(struct.pack("<" + "f" * len(r), *r))


def test_ply_conventions(tmp_path):
    # log scale -> exp, mass = sigma_t_0 * multiplier, albedo clamped, NaN albedo -> 0
    r0 = [0.1, 0.2, 0.3, 2.0, 0.0, 0.0, 0.0, np.log(0.5), np.log(0.25), 0.0, 1.5, -0.2, 0.4, 3.0]
    r1 = [1.0, -1.0, 0.5, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, np.nan, 0.5, 1.0, 0.5]
    p = str(tmp_path / "a.ply")
    _write_ply(p, [r0, r1], extra_elem=True)
    rows = vpio.load_ply(p, sigma_multiplier=2.0)
    assert rows.shape == (2, 14)
    assert np.allclose(rows[0, 7:10], [0.5, 0.25, 1.0], rtol=1e-6)
    assert np.allclose(rows[:, 13], [6.0, 1.0])
    assert rows[0, 10:13].tolist() == [1.0, 0.0, np.float32(0.4)]
    assert rows[1, 10] == 0.0
    assert np.allclose(rows[0, 3:7], [2.0, 0, 0, 0])  # the library normalises quaternions
    ov = vpio.load_ply(p, 1.0, albedo_override=[0.9, 0.9, 0.9])
    assert np.allclose(ov[:, 10:13], 0.9)
    pa = str(tmp_path / "b.ply")
    _write_ply(pa, [r0, r1], fmt="ascii")
    assert np.allclose(vpio.load_ply(pa, 2.0), rows, equal_nan=True)


@pytest.mark.parametrize("bad,msg", [(7, "scale"), (13, "optical_thickness"), (0, "center")])
def test_ply_validation_errors(tmp_path, bad, msg):
    r = [0.0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0.5, 0.5, 0.5, 1.0]
    r[bad] = np.nan if bad == 0 else (-np.inf if bad == 7 else -1.0)
    p = str(tmp_path / "bad.ply")
    _write_ply(p, [r, r])
    with pytest.raises(vpio.IoError, match=msg):
        vpio.load_ply(p, 1.0)


def test_ply_missing_property(tmp_path):
    p = str(tmp_path / "m.ply")
    names = ["x", "y", "z", "rot_0", "rot_1", "rot_2", "rot_3", "scale_0", "scale_1", "scale_2", "albedo_0",
             "albedo_1", "albedo_2", "density"]
    _write_ply(p, [[0.0] * 14], names=names)
    with pytest.raises(vpio.IoError, match="sigma_t_0"):
        vpio.load_ply(p)


def test_ply_round_trip_of_a_generated_scene(tmp_path):
    rows = scenes.c1_scattering(n=257, seed=3)
    p = str(tmp_path / "c1.ply")
    vpio.save_ply(p, rows, sigma_multiplier=10.0)
    back = vpio.load_ply(p, sigma_multiplier=10.0)
    assert np.allclose(back[:, 0:7], rows[:, 0:7]) and np.allclose(back[:, 7:10], rows[:, 7:10], rtol=1e-6)
    assert np.allclose(back[:, 13], rows[:, 13], rtol=1e-6)


@pytest.mark.parametrize("w,h", [(64, 32), (5, 3)])
def test_rgbe_round_trip(tmp_path, w, h):
    rng = np.random.default_rng(1)
    img = np.ones((h, w, 4), np.float32)
    img[..., :3] = np.exp(rng.uniform(-6, 6, (h, w, 3)))
    img[0, 0, :3] = 0.0
    p = str(tmp_path / "e.hdr")
    vpio.save_hdr(p, img)
    back = vpio.load_hdr(p)
    assert back.shape == img.shape and np.all(back[..., 3] == 1.0)
    m = img[..., :3].max(axis=2, keepdims=True)
    # RGBE keeps an 8-bit mantissa relative to the largest channel
    assert np.all(np.abs(back[..., :3] - img[..., :3]) <= m / 128.0 + 1e-30)
    assert np.all(back[0, 0, :3] == 0.0)


def test_rgbe_reader_value_convention(tmp_path):
    # byte * 2^(e - 136) (hdr.cpp:32-44): (128, 64, 1, e=129) -> (1, 0.5, 1/128)
    p = str(tmp_path / "r.hdr")
    with open(p, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 2\n")
        f.write(bytes([128, 64, 1, 129, 0, 0, 0, 0]))
    img = vpio.load_hdr(p)
    assert img[0, 0, :3].tolist() == [1.0, 0.5, 1.0 / 128.0] and img[0, 1, :3].tolist() == [0, 0, 0]


def test_exr_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    img = rng.random((7, 9, 4)).astype(np.float32)
    p = str(tmp_path / "o.exr")
    vpio.save_exr(p, img)
    data = open(p, "rb").read()
    assert data[:4] == bytes([0x76, 0x2F, 0x31, 0x01])
    i = data.index(b"channels\x00chlist\x00") + len(b"channels\x00chlist\x00") + 4
    assert data[i:i + 2] == b"B\x00" and data[i + 18:i + 20] == b"G\x00" and data[i + 36:i + 38] == b"R\x00"
    # first scanline block holds the LAST image row (always flipped, exr.cpp:55)
    hdr_end = data.index(b"screenWindowWidth\x00float\x00") + len(b"screenWindowWidth\x00float\x00") + 8 + 1
    first = struct.unpack("<Q", data[hdr_end:hdr_end + 8])[0]
    y, size = struct.unpack("<ii", data[first:first + 8])
    assert y == 0 and size == 9 * 12
    b0 = np.frombuffer(data[first + 8:first + 8 + 36], np.float32)
    assert np.array_equal(b0, img[6, :, 2])
    back = vpio.load_exr(p)
    assert np.array_equal(back[..., :3], img[..., :3])


def test_ply_parallel_ingest_at_cloud_scale(tmp_path):
    # 10^6 rows (the cloud's size), decoded, converted and validated in
    # parallel chunks (ply.cpp:97,123 use std::execution::par): every row
    # equals the per-row conversion, binary and ascii agree, and the error
    # reported is the lowest failing row even when later chunks fail too
    import time
    n = 1_000_000
    rng = np.random.default_rng(9)
    props = np.empty((n, 14), np.float32)
    props[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    props[:, 3:7] = rng.normal(size=(n, 4))
    props[:, 7:10] = np.log(rng.uniform(0.005, 0.03, (n, 3)))
    props[:, 10:13] = rng.uniform(-0.2, 1.2, (n, 3))
    props[:, 13] = rng.uniform(0.1, 2.0, n)
    p = str(tmp_path / "big.ply")
    names = ["x", "y", "z", "rot_0", "rot_1", "rot_2", "rot_3", "scale_0", "scale_1", "scale_2",
             "albedo_0", "albedo_1", "albedo_2", "sigma_t_0"]
    with open(p, "wb") as f:
        f.write(("ply\nformat binary_little_endian 1.0\n" + f"element vertex {n}\n" +
                 "".join(f"property float {k}\n" for k in names) + "end_header\n").encode())
        f.write(props.astype("<f4").tobytes())
    t0 = time.perf_counter()
    rows = vpio.load_ply(p, sigma_multiplier=7.5)
    dt = time.perf_counter() - t0
    print(f"\n[ply] 10^6 rows loaded in {dt * 1e3:.0f} ms ({os.cpu_count()} cores)")
    assert rows.shape == (n, 14)
    assert np.array_equal(rows[:, 0:3], props[:, 0:3])
    assert np.array_equal(rows[:, 7:10], np.exp(props[:, 7:10]).astype(np.float32)) or \
        np.allclose(rows[:, 7:10], np.exp(props[:, 7:10]), rtol=1e-6, atol=0)
    assert np.array_equal(rows[:, 10:13], np.clip(props[:, 10:13], 0, 1))
    assert np.array_equal(rows[:, 13], props[:, 13] * np.float32(7.5))
    # ascii of a slice parses to the same rows
    m = 20_000
    pa = str(tmp_path / "slice.ply")
    _write_ply(pa, props[:m].tolist(), fmt="ascii")
    assert np.array_equal(vpio.load_ply(pa, 7.5), rows[:m])
    # two bad rows in different chunks: the lower one is reported
    bad = props.copy()
    bad[700_001, 13] = -1.0
    bad[123_457, 7] = np.inf
    pb = str(tmp_path / "bad.ply")
    with open(pb, "wb") as f:
        f.write(("ply\nformat binary_little_endian 1.0\n" + f"element vertex {n}\n" +
                 "".join(f"property float {k}\n" for k in names) + "end_header\n").encode())
        f.write(bad.astype("<f4").tobytes())
    with pytest.raises(RuntimeError, match="Primitive 123457: NaN/Inf in scale"):
        vpio.load_ply(pb, 1.0)
