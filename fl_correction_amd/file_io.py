"""Attenuation-map files, interchangeable with the reference's.

The reference writes one float32 TIFF per element and angle
(FL/file_io.py:69-126) and reads them back per angle
(read_attenuation_at_all_angle, FL/file_io.py:37-63):

    atten_{elem}_prj_{angle_id:04d}.tiff                total attenuation
    atten_fl_{elem}_prj_{angle_id:04d}.tiff             fluorescence path only
    atten_incident_xray_{elem}_prj_{angle_id:04d}.tiff  incident beam only

each holding the (B, H, W) map of that angle as B pages of (H, W), or one page
for a single slice.  Here the maps normally stay in HBM as one
(K, A, B, H, W) stack; these functions are the optional bridge to files with
the same names and layout (read and written with PIL, float32 pages).
"""
from __future__ import annotations

import os

import numpy as np
from PIL import Image, ImageSequence

_PREFIX = {"all": "atten_", "fl": "atten_fl_", "xray": "atten_incident_xray_"}


def attenuation_file(file_path, elem, angle_id, mode="all"):
    """Path of one map, named as the reference names it (FL/file_io.py:86,106,125)."""
    if mode not in _PREFIX:
        raise ValueError(f"mode must be one of {sorted(_PREFIX)}")
    return os.path.join(file_path, f"{_PREFIX[mode]}{elem}_prj_{int(angle_id):04d}.tiff")


def write_tiff(fname, data):
    """float32 TIFF: a 2-D array as one page, a 3-D array (B, H, W) as B pages."""
    data = np.ascontiguousarray(np.asarray(data, dtype=np.float32))
    if data.ndim == 2:
        data = data[None]
    if data.ndim != 3:
        raise ValueError("write_tiff needs a 2-D or 3-D array")
    pages = [Image.fromarray(p, mode="F") for p in data]
    pages[0].save(fname, format="TIFF", save_all=True, append_images=pages[1:])


def read_tiff(fname):
    """float32 array: (H, W) for one page, (B, H, W) for several."""
    with Image.open(fname) as im:
        pages = [np.array(p, dtype=np.float32) for p in ImageSequence.Iterator(im)]
    return pages[0] if len(pages) == 1 else np.stack(pages)


def _write(elem, data, angle_id, file_path, mode):
    os.makedirs(file_path, exist_ok=True)
    write_tiff(attenuation_file(file_path, elem, angle_id, mode), data)


def write_attenuation(elem, data, angle_id, file_path="./Angle_prj"):
    """Total attenuation of one angle (FL/file_io.py:69-87)."""
    _write(elem, data, angle_id, file_path, "all")


def write_attenuation_fl(elem, data, angle_id, file_path="./Angle_prj"):
    """Fluorescence-path attenuation of one angle (FL/file_io.py:89-107)."""
    _write(elem, data, angle_id, file_path, "fl")


def write_attenuation_xray(elem, data, angle_id, file_path="./Angle_prj"):
    """Incident-beam attenuation of one angle (FL/file_io.py:110-126)."""
    _write(elem, data, angle_id, file_path, "xray")


def read_attenuation(angle_id, fpath_atten, elem):
    return read_tiff(attenuation_file(fpath_atten, elem, angle_id, "all"))


def read_attenuation_fl(angle_id, fpath_atten, elem):
    return read_tiff(attenuation_file(fpath_atten, elem, angle_id, "fl"))


def read_attenuation_xray(angle_id, fpath_atten, elem):
    return read_tiff(attenuation_file(fpath_atten, elem, angle_id, "xray"))


def read_attenuation_at_all_angle(angle_list, fpath_atten, elem, mode="all"):
    """(n_angle, *map shape) for one element, float64 like the reference's
    (FL/file_io.py:37-63; mode 'all', 'fl' or 'xray')."""
    n = len(angle_list)
    out = None
    for i in range(n):
        tmp = read_tiff(attenuation_file(fpath_atten, elem, i, mode))
        if out is None:
            out = np.zeros((n, *tmp.shape))
        out[i] = tmp
    return out


def write_attenuation_stack(atten, elem_list, file_path="./Angle_prj", mode="all"):
    """Every map of a (K, A, B, H, W) stack (device tensor or array) as per-angle
    files for elements elem_list[k]."""
    K, A = int(atten.shape[0]), int(atten.shape[1])
    if len(elem_list) != K:
        raise ValueError("one element name per map")
    for k in range(K):
        maps = atten[k].detach().cpu().numpy() if hasattr(atten, "detach") else np.asarray(atten[k])
        for a in range(A):
            _write(elem_list[k], maps[a], a, file_path, mode)


def read_attenuation_stack(n_angle, elem_list, file_path="./Angle_prj", mode="all", device=None):
    """(K, A, B, H, W) float32 stack from per-angle files, as a device tensor when
    `device` is given (the layout the projector streams), else an array."""
    maps = np.stack([np.stack([read_tiff(attenuation_file(file_path, e, a, mode)) for a in range(int(n_angle))])
                     for e in elem_list]).astype(np.float32)
    if maps.ndim == 4:  # single-slice maps: (K, A, H, W) -> (K, A, 1, H, W)
        maps = maps[:, :, None]
    if device is None:
        return maps
    import torch
    return torch.from_numpy(maps).to(device)
