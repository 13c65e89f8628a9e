"""The exp_runner / dataset_loader boundary on the CPU: the HOCON subset parser
on a config with config/diligent.conf's key layout, the command line of
exp_runner.py:639-645, the Runner's dataset selection (exp_runner.py:45-48)
and the on-disk dataset format (models/dataset_loader.py:39-145).

The EXR files are assembled here byte by byte from the published OpenEXR
layout (magic + version, attribute header, chunk offset table, per-line
channel planes; ZIP = zlib of the byte-interleaved, delta-predicted lines),
independently of the reader under test."""
import struct
import zlib

import numpy as np
import pytest
import torch

DILIGENT_STYLE = """
general {
    dataset_class = models.dataset_loader.Dataset
    renderer_class = models.renderer.NeuSRenderer

    base_exp_dir = ./exp/diligent_mv/CASE_NAME
    recording = [
        ./,
        ./models
    ]
}

dataset {
    data_dir = data/diligent_mv_normals/CASE_NAME/
    normal_dir = normal_world_space_sdmunips # world space normals
    cameras_name = cameras_sphere.npz
    exclude_views = [0, 4, 8, 12, 16]  # 0-based
    upsample_factor = 1
}

train {
    learning_rate = 5e-4
    learning_rate_alpha = 0.05
    end_iter = 5000
    increase_bindwidth_every = 350  // a comment of the other kind
    gradient_method = dfd
    batch_size = 2048
    patch_size = 3
    warm_up_end = 50
    use_white_bkgd = False
    loss_type = l2
    normal_weight = 1
    eikonal_weight = 1
    mask_weight = 1
}

model {
    sdf_network {
        d_out = 1
        d_in = 3
        d_hidden = 64
        n_layers = 1
        skip_in = [-1]
        bias = 0.6
        geometric_init = True
        weight_norm = True
        input_concat = True
    }
    variance_network {
        init_val = 0.5
    }
    ray_marching {
        start_step_size = 1e-2
        end_step_size = 1e-3
        occ_threshold = 0.1
        occ_resolution = 128
        occ_update_freq = 8  # batches
    }
    encoding{
        otype=HashGrid,
\t\tn_levels=14
\t\tn_features_per_level=2
\t\tlog2_hashmap_size=19
\t\tbase_resolution=32
\t\tper_level_scale=1.3195079107728942
   }
}
"""


def test_hocon_diligent_layout():
    from supernormal_amd.conf import parse
    c = parse(DILIGENT_STYLE.replace("CASE_NAME", "bear"))
    assert c["general.dataset_class"] == "models.dataset_loader.Dataset"
    assert c["general.base_exp_dir"] == "./exp/diligent_mv/bear"
    assert c["general.recording"] == ["./", "./models"]
    assert c["dataset.normal_dir"] == "normal_world_space_sdmunips"
    assert c["dataset.exclude_views"] == [0, 4, 8, 12, 16]
    assert c.get_float("train.learning_rate") == 5e-4
    assert c.get_int("train.end_iter") == 5000
    assert c.get_int("train.increase_bindwidth_every") == 350
    assert c["train.gradient_method"] == "dfd"
    assert c.get_bool("train.use_white_bkgd") is False
    assert c.get_int("train.missing", 7) == 7
    sdf = c["model.sdf_network"]
    assert sdf["skip_in"] == [-1] and sdf["geometric_init"] is True and sdf["bias"] == 0.6
    enc = c["model.encoding"]
    assert enc["otype"] == "HashGrid" and enc["n_levels"] == 14
    assert enc["per_level_scale"] == 1.3195079107728942
    assert c.get_float("model.ray_marching.end_step_size") == 1e-3


def test_command_line_flags(tmp_path):
    from supernormal_amd.runner import build_parser, read_conf_text
    a = build_parser().parse_args([])
    assert (a.conf, a.mode, a.mcube_threshold, a.is_continue, a.gpu, a.obj_name) == \
        ("./confs/base.conf", "eval_normal", 0.0, False, 0, "")
    a = build_parser().parse_args(["--conf", "x.conf", "--mode", "train", "--is_continue",
                                   "--gpu", "3", "--obj_name", "cow", "--mcube_threshold", "0.1"])
    assert (a.conf, a.mode, a.is_continue, a.gpu, a.obj_name, a.mcube_threshold) == \
        ("x.conf", "train", True, 3, "cow", 0.1)
    p = tmp_path / "c.conf"
    p.write_text(DILIGENT_STYLE)
    assert "./exp/diligent_mv/cow" in read_conf_text(str(p), "cow")


def test_runner_dataset_selection_raises_instead_of_substituting():
    """No dataset block, an unknown loader class or a missing scene directory is
    an error; nothing falls back to a synthetic scene."""
    from supernormal_amd.conf import parse
    from supernormal_amd.runner import Runner
    r = Runner.__new__(Runner)
    r.device = torch.device("cpu")
    with pytest.raises(KeyError):
        r._config_dataset(parse("train { end_iter = 1 }"), 0)
    with pytest.raises(ValueError):
        r._config_dataset(parse("general { dataset_class = foo.Bar }\ndataset { data_dir = x }"), 0)
    with pytest.raises(FileNotFoundError):
        r._config_dataset(parse(DILIGENT_STYLE.replace("CASE_NAME", "none")), 0)


# ------------------------------------------------------------ on-disk format
def _attr(name, typ, payload):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(payload)) + payload


def _write_exr(path, chans, compression):
    """Scanline EXR of float32 planes {name: [H, W]} stored as HALF (R, G) or
    FLOAT (others); compression 0 (NONE) or 3 (ZIP, 16 lines per chunk)."""
    names = sorted(chans)
    H, W = chans[names[0]].shape
    ptype = {n: (1 if n in ("R", "G") else 2) for n in names}
    chl = b"".join(n.encode() + b"\0" + struct.pack("<iB3sii", ptype[n], 0, b"\0\0\0", 1, 1)
                   for n in names) + b"\0"
    hdr = struct.pack("<ii", 20000630, 2)
    hdr += _attr("channels", "chlist", chl)
    hdr += _attr("compression", "compression", bytes([compression]))
    hdr += _attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, W - 1, H - 1))
    hdr += _attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, W - 1, H - 1))
    hdr += _attr("lineOrder", "lineOrder", b"\0")
    hdr += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    hdr += b"\0"
    lpb = 16 if compression == 3 else 1
    chunks = []
    for y in range(0, H, lpb):
        raw = b""
        for row in range(y, min(y + lpb, H)):
            for n in names:
                dt = "<f2" if ptype[n] == 1 else "<f4"
                raw += chans[n][row].astype(dt).tobytes()
        if compression == 3:
            b = np.frombuffer(raw, np.uint8)
            t = np.concatenate([b[0::2], b[1::2]]).astype(np.int64)
            d = t.copy()
            d[1:] = (t[1:] - t[:-1] + 128 + 256) & 0xFF
            data = zlib.compress(d.astype(np.uint8).tobytes())
            if len(data) >= len(raw):  # OpenEXR stores such a chunk uncompressed
                data = raw
        else:
            data = raw
        chunks.append(struct.pack("<ii", y, len(data)) + data)
    table_at = len(hdr)
    offs, pos = [], table_at + 8 * len(chunks)
    for ch in chunks:
        offs.append(pos)
        pos += len(ch)
    with open(path, "wb") as f:
        f.write(hdr + struct.pack(f"<{len(offs)}Q", *offs) + b"".join(chunks))


def test_exr_reader_none_and_zip(tmp_path):
    from supernormal_amd.dataset_loader import exr_rgb, read_exr
    rng = np.random.default_rng(0)
    H, W = 37, 21
    yy, xx = np.mgrid[0:H, 0:W] / 8.0
    planes = {"R": np.sin(xx), "G": np.cos(yy) * 0.5, "B": xx * yy,
              "Z": rng.normal(size=(H, W))}  # smooth planes compress, noise does not
    for comp in (0, 3):
        p = tmp_path / f"n{comp}.exr"
        _write_exr(p, planes, comp)
        if comp == 3:  # some chunks really are compressed
            assert p.stat().st_size < 8 + 37 * 21 * (2 + 2 + 4 + 4)
        got = read_exr(p)
        for n in "RG":
            np.testing.assert_array_equal(got[n], planes[n].astype(np.float16).astype(np.float32))
        for n in "BZ":
            np.testing.assert_array_equal(got[n], planes[n].astype(np.float32))
        rgb = exr_rgb(p)
        assert rgb.shape == (H, W, 3)
        np.testing.assert_array_equal(rgb[..., 2], planes["B"].astype(np.float32))


def test_decompose_projection_recovers_camera():
    from supernormal_amd.dataset_loader import decompose_projection
    rng = np.random.default_rng(1)
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    R = q * np.sign(np.diag(r))
    if np.linalg.det(R) < 0:
        R[:, 0] *= -1
    K = np.array([[800.0, 0.5, 300.0], [0, 790.0, 260.0], [0, 0, 1]])
    C = np.array([0.4, -1.9, 2.2])
    P = 3.7 * K @ np.concatenate([R, (-R @ C)[:, None]], 1)  # any scale
    intr, pose = decompose_projection(P)
    np.testing.assert_allclose(intr[:3, :3], K, rtol=1e-9, atol=1e-7)
    np.testing.assert_allclose(pose[:3, :3], R.T, atol=1e-6)  # camera-to-world
    np.testing.assert_allclose(pose[:3, 3], C, atol=1e-5)
    intr2, pose2 = decompose_projection(-P)  # the sign of P does not matter
    np.testing.assert_allclose(pose2, pose, atol=1e-6)


def test_load_dataset_from_disk(tmp_path):
    """A 3-view scene in the reference's layout, loaded through the Runner's
    config path: cameras, normals (background zeroed), masks, exclusions,
    per-pixel V^-1."""
    from PIL import Image

    from supernormal_amd.conf import parse
    from supernormal_amd.runner import Runner
    from supernormal_amd.scene import look_at_pose
    H, W = 24, 32
    K = np.array([[40.0, 0, 16.0], [0, 42.0, 12.0], [0, 0, 1]])
    d = tmp_path / "obj"
    (d / "normals").mkdir(parents=True)
    (d / "mask").mkdir()
    cams, poses, ns, ms = {}, [], [], []
    rng = np.random.default_rng(2)
    for i, eye in enumerate([(0, 0, 2.5), (2.5, 0, 0), (0, 2.0, 1.5)]):
        pose = look_at_pose(np.array(eye, float))
        Rw = pose[:3, :3].T  # world-to-camera
        P = np.eye(4)
        P[:3] = K @ np.concatenate([Rw, (-Rw @ pose[:3, 3])[:, None]], 1)
        S = np.diag([2.0, 2.0, 2.0, 1.0])  # scale_mat: unit sphere -> world
        cams[f"world_mat_{i}"] = P.astype(np.float32)  # projects world points
        cams[f"scale_mat_{i}"] = S.astype(np.float32)
        poses.append(pose)
        yy, xx = np.mgrid[0:H, 0:W] / 5.0
        n = np.stack([np.sin(xx + i), np.cos(yy), xx * yy / 10], -1) + rng.normal(size=(H, W, 3)) * 1e-3
        m = (rng.random((H, W)) > 0.3).astype(np.uint8) * 255
        _write_exr(d / "normals" / f"{i:03d}.exr", {"R": n[..., 0], "G": n[..., 1],
                                                  "B": n[..., 2]}, 3)
        Image.fromarray(np.stack([m] * 3, -1)).save(d / "mask" / f"{i:03d}.png")
        ns.append(n)
        ms.append(m)
    np.savez(d / "cameras_sphere.npz", **cams)
    conf = parse(f"""general {{ dataset_class = models.dataset_loader.Dataset }}
dataset {{
    data_dir = {d}
    normal_dir = normals
    cameras_name = cameras_sphere.npz
    exclude_views = [1]
}}""")
    r = Runner.__new__(Runner)
    r.device = torch.device("cpu")
    ds = r._config_dataset(conf, 0)
    assert (ds.n_images, ds.H, ds.W) == (3, H, W)
    assert list(ds.train_images) == [0, 2]
    for i in range(3):
        np.testing.assert_allclose(ds.intrinsics_all[i, :3, :3].numpy(), K, rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(ds.pose_all[i, :3, :3].numpy(), poses[i][:3, :3], atol=1e-5)
        np.testing.assert_allclose(ds.pose_all[i, :3, 3].numpy(), poses[i][:3, 3] / 2.0,
                                   atol=1e-5)  # in the unit sphere's coordinates
        m = ms[i] > 0
        np.testing.assert_array_equal(ds.masks[i].numpy() > 0.5, m)
        stored = np.concatenate([ns[i][..., :2].astype(np.float16).astype(np.float32),
                                 ns[i][..., 2:].astype(np.float32)], -1)  # R, G half; B float
        expect = np.where(m[..., None], stored, 0.0)
        np.testing.assert_array_equal(ds.normals[i].numpy(), expect)
    # V^-1 inverts V = [ray direction, camera x, camera y] at every pixel
    V = torch.linalg.inv(ds.V_inverse_all[2])
    np.testing.assert_allclose(V[..., 1, :].numpy(), np.broadcast_to(poses[2][:3, 0], (H, W, 3)),
                               atol=1e-5)


def test_profiling_reductions(tmp_path):
    """The rocprofv3 CSV reductions behind profiles/: per-kernel means over
    all and over the last dispatches, counter means, the probe means."""
    from supernormal_amd import profiling
    d = tmp_path / "run"
    d.mkdir()
    with open(d / "x_kernel_trace.csv", "w") as f:
        f.write("Kernel_Name,Start_Timestamp,End_Timestamp,Queue_Id\n")
        for i in range(10):
            f.write(f"kA,{i * 1000},{i * 1000 + 100 + 10 * i},1\n")
        f.write("kB,50000,52000,2\n")
    out = tmp_path / "k.tsv"
    profiling.kernel_means(str(d), str(out), last=2)
    lines = out.read_text().splitlines()
    assert lines[0].startswith("kernel\tcalls")
    rows = {l.split("\t")[0]: l.split("\t") for l in lines[1:]}
    assert rows["kA"][1] == "10" and rows["kA"][2] == "0.1" and rows["kA"][3] == "0.2"
    assert rows["kB"][3] == "2.0"
    with open(d / "x_counter_collection.csv", "w") as f:
        f.write("Kernel_Name,Counter_Name,Dispatch_Id,Counter_Value\n")
        for i in range(6):
            f.write(f"kA,FETCH_SIZE,{i},{i}\n")
    out2 = tmp_path / "c.tsv"
    profiling.counter_means(str(d), str(out2), last=2)
    assert out2.read_text().split("\t")[-1].strip() == "4.5"
    out3 = tmp_path / "p.tsv"
    profiling.probe_means(str(d), str(out3))
    assert out3.read_text().split("\t")[-1].strip() == "3.0"
    out4 = tmp_path / "t.tsv"
    profiling.timeline(str(d), str(out4), last=3)
    assert len(out4.read_text().splitlines()) == 3
