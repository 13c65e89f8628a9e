"""The reference's on-disk dataset format (models/dataset_loader.py:39-145).

A scene directory holds
  <data_dir>/<cameras_name>            npz with world_mat_<i> (3x4 or 4x4 world ->
                                       image projection) and scale_mat_<i> (the
                                       normalisation into the unit sphere) per view
  <data_dir>/<normal_dir>/*.exr        world-space normal maps (first 3 channels)
  <data_dir>/mask/*.png                masks (first channel / 255 > 0)
and the config block `dataset { data_dir, normal_dir, cameras_name,
exclude_views, upsample_factor }` (config/diligent.conf:13-19).  View i is
the i-th file of the sorted listing; its cameras are world_mat_<k> /
scale_mat_<k> with k the file name's integer stem (dataset_loader.py:59-92).

The reference reads EXR with pyexr, PNG with OpenCV and splits the projection
with cv2.decomposeProjectionMatrix; none of them is in this image, so:
  * read_exr: OpenEXR scanline files (single part) with NONE, RLE, ZIPS or ZIP
    compression and HALF / FLOAT / UINT channels, restated from the published
    OpenEXR file layout (magic, attribute header, chunk offset table, per-line
    channel planes; zlib + byte interleave + delta predictor for ZIP / ZIPS);
    other compressions raise NotImplementedError;
  * masks through Pillow;
  * decompose_projection: P[:, :3] = K R by an RQ decomposition with K's
    diagonal made positive, camera centre = the null vector of P -- what
    load_K_Rt_from_P (dataset_loader.py:13-35) returns.
The result is a supernormal_amd.scene.Dataset (the same interface as the
synthetic scenes).
"""
from __future__ import annotations

import os
import struct
import zlib
from glob import glob

import numpy as np

EXR_MAGIC = 20000630
_LINES_PER_BLOCK = {0: 1, 1: 1, 2: 1, 3: 16}  # NONE, RLE, ZIPS, ZIP
_PIXEL_BYTES = {0: 4, 1: 2, 2: 4}            # UINT, HALF, FLOAT
_PIXEL_DTYPE = {0: "<u4", 1: "<f2", 2: "<f4"}


def _cstr(buf, pos):
    end = buf.index(b"\0", pos)
    return buf[pos:end].decode("latin-1"), end + 1


def _exr_header(buf):
    magic, version = struct.unpack_from("<ii", buf, 0)
    if magic != EXR_MAGIC:
        raise ValueError("not an OpenEXR file")
    if version & 0x1a00:  # tiled (0x200), deep (0x800), multi-part (0x1000)
        raise NotImplementedError("only single-part scanline OpenEXR files are supported")
    pos = 8
    attrs = {}
    while buf[pos] != 0:
        name, pos = _cstr(buf, pos)
        typ, pos = _cstr(buf, pos)
        size, = struct.unpack_from("<i", buf, pos)
        pos += 4
        attrs[name] = (typ, buf[pos:pos + size])
        pos += size
    return attrs, pos + 1


def _channels(raw):
    out, pos = [], 0
    while raw[pos] != 0:
        name, pos = _cstr(raw, pos)
        ptype, _, _, xs, ys = struct.unpack_from("<iB3sii", raw, pos)
        pos += 16
        if xs != 1 or ys != 1:
            raise NotImplementedError("sub-sampled EXR channels are not supported")
        out.append((name, ptype))
    return out


def _unpredict_deinterleave(t):
    """ZIP / RLE post-processing of OpenEXR: undo the delta predictor, then
    re-interleave the two halves (even bytes first, odd bytes second)."""
    t = np.frombuffer(t, np.uint8)
    if t.size == 0:
        return b""
    d = np.cumsum(t.astype(np.int64) - 128, dtype=np.int64)
    d += 128  # t[0] is stored as is: cumsum added -128 to it once
    t = (d & 0xFF).astype(np.uint8)
    half = (t.size + 1) // 2
    out = np.empty_like(t)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def _rle_decode(src, expect):
    out = bytearray()
    i = 0
    while i < len(src) and len(out) < expect:
        c = struct.unpack_from("b", src, i)[0]
        i += 1
        if c < 0:
            out += src[i:i - c]
            i -= c
        else:
            out += src[i:i + 1] * (c + 1)
            i += 1
    return bytes(out)


def read_exr(path):
    """{channel name: float32 [H, W]} of an OpenEXR scanline file."""
    with open(path, "rb") as f:
        buf = f.read()
    attrs, pos = _exr_header(buf)
    chans = _channels(attrs["channels"][1])
    comp = attrs["compression"][1][0]
    if comp not in _LINES_PER_BLOCK:
        raise NotImplementedError(f"EXR compression {comp} is not supported "
                                  "(NONE, RLE, ZIPS and ZIP are)")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    lpb = _LINES_PER_BLOCK[comp]
    n_chunks = (H + lpb - 1) // lpb
    offsets = struct.unpack_from(f"<{n_chunks}Q", buf, pos)
    line_bytes = sum(W * _PIXEL_BYTES[t] for _, t in chans)
    planes = {name: np.zeros((H, W), np.float32) for name, _ in chans}
    for off in offsets:
        y, size = struct.unpack_from("<ii", buf, off)
        data = buf[off + 8: off + 8 + size]
        n_lines = min(lpb, y1 - y + 1)
        expect = n_lines * line_bytes
        if size < expect:
            if comp in (2, 3):
                data = _unpredict_deinterleave(zlib.decompress(data))
            elif comp == 1:
                data = _unpredict_deinterleave(_rle_decode(data, expect))
        if len(data) != expect:
            raise ValueError(f"{path}: chunk at line {y} has {len(data)} bytes, expected {expect}")
        p = 0
        for k in range(n_lines):
            row = y - y0 + k
            for name, t in chans:
                nb = W * _PIXEL_BYTES[t]
                planes[name][row] = np.frombuffer(data, _PIXEL_DTYPE[t], W, p).astype(np.float32)
                p += nb
    return planes


def exr_rgb(path):
    """The first three channels in R, G, B order (pyexr's read()[..., :3]);
    files without R / G / B give their first three channels in stored order."""
    planes = read_exr(path)
    names = [n for n in ("R", "G", "B") if n in planes]
    if len(names) < 3:
        names = list(planes)[:3]
    return np.stack([planes[n] for n in names], -1)


def read_mask(path):
    """First channel of a PNG mask scaled to [0, 1] (cv.imread(...) / 255, [..., 0])."""
    from PIL import Image
    a = np.asarray(Image.open(path))
    if a.ndim == 3:
        a = a[..., 0]
    return a.astype(np.float32) / (65535.0 if a.dtype == np.uint16 else 255.0)


def decompose_projection(P):
    """(intrinsics 4x4, camera-to-world pose 4x4) of a 3x4 projection
    P = K [R | -R C]: load_K_Rt_from_P (dataset_loader.py:13-35)."""
    from scipy.linalg import rq
    P = np.asarray(P, np.float64)[:3, :4]
    M = P[:, :3]
    if np.linalg.det(M) < 0:  # P and -P project alike
        P = -P
        M = -M
    K, R = rq(M)
    D = np.diag(np.sign(np.diag(K)))
    K, R = K @ D, D @ R
    C = -np.linalg.solve(M, P[:, 3])
    intr = np.eye(4)
    intr[:3, :3] = K / K[2, 2]
    pose = np.eye(4, dtype=np.float32)
    pose[:3, :3] = R.T
    pose[:3, 3] = C
    return intr, pose


def load_dataset(conf, device):
    """The scene directory named by a `dataset` config block as a
    supernormal_amd.scene.Dataset (dataset_loader.py:39-145)."""
    from .scene import Dataset
    data_dir = conf.get("data_dir")
    normal_dir = conf.get("normal_dir")
    cameras_name = conf.get("cameras_name")
    if not data_dir or not normal_dir or not cameras_name:
        raise ValueError("dataset: data_dir, normal_dir and cameras_name are required")
    exclude = list(conf.get("exclude_views", []) or [])
    up = int(conf.get("upsample_factor", 1) or 1)
    normal_files = sorted(glob(os.path.join(data_dir, normal_dir, "*.exr")))
    mask_files = sorted(glob(os.path.join(data_dir, "mask", "*.png")))
    if not normal_files:
        raise FileNotFoundError(f"no normal maps in {os.path.join(data_dir, normal_dir)}")
    if len(mask_files) != len(normal_files):
        raise ValueError(f"{len(normal_files)} normal maps but {len(mask_files)} masks")
    cams = np.load(os.path.join(data_dir, cameras_name))
    idx = [int(os.path.basename(f).split(".")[0]) for f in normal_files]
    normals = np.stack([exr_rgb(f) for f in normal_files]).astype(np.float32)
    masks = np.stack([read_mask(f) for f in mask_files])
    if up > 1:
        import torch
        import torch.nn.functional as F
        normals = F.interpolate(torch.from_numpy(normals).permute(0, 3, 1, 2), scale_factor=up,
                                mode="bilinear", align_corners=False).permute(0, 2, 3, 1).numpy()
        masks = F.interpolate(torch.from_numpy(masks)[:, None], scale_factor=up,
                              mode="nearest")[:, 0].numpy()
    normals[masks == 0] = 0.0
    Ks, poses = [], []
    for k in idx:
        P = cams[f"world_mat_{k}"].astype(np.float32) @ cams[f"scale_mat_{k}"].astype(np.float32)
        K, pose = decompose_projection(P[:3, :4])
        if up > 1:
            K[0, 0] *= up
            K[1, 1] *= up
            K[0, 2] *= up
            K[1, 2] *= up
        Ks.append(K)
        poses.append(pose)
    # masks keep their values; the training loop binarises at 0.5 (exp_runner.py:169-172)
    return Dataset.from_arrays(normals, masks.astype(np.float32), np.stack(Ks),
                               np.stack(poses), device, exclude_views=exclude)
