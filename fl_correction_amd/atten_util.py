"""Drop-in for FL/atten_util.py: parameter plumbing for the attenuation maps.

The ``param`` dictionary has the layout ``load_param`` produces (FL/core.py:213-341):
  param['elem_type']      ['Ni', 'Co', ...]           emission lines, one per element
  param['elem_compound']  ['LiNiO2', 'LiCoO2', ...]   compound carrying each element
  param['rho']            {compound: density g/cm^3}
  param['pix']            voxel size in cm
  param['cs']             {'elem_type', 'elem_compound', f'{comp}-x': cross-section(s) at
                           the incident energy (one value or one per angle/energy),
                           f'{comp}-{elem}': cross-section at elem's emission line}
  param['em_cs']          {elem: emission cross-section(s) (one value or one per angle)}
``make_param`` builds one from plain arrays (the workloads seed cross-sections
directly; xraylib is not available here).
"""
from __future__ import annotations

from copy import deepcopy

import numpy as np
import torch


def get_atten_coef(elem_type, elem_compound, XEng_list, em_E):
    """Cross-section table from xraylib (FL/atten_util.py:14-39).  xraylib is not part of
    this environment; build the table with make_param from known values instead."""
    try:
        import xraylib  # noqa: F401
    except ImportError as err:
        raise ImportError("get_atten_coef needs xraylib, which is not installed; "
                          "use atten_util.make_param with tabulated cross-sections") from err
    import xraylib
    cs = {"elem_type": elem_type, "elem_compound": elem_compound, "x": XEng_list}
    for i in range(len(elem_type)):
        cs[f"{elem_compound[i]}-x"] = np.array([xraylib.CS_Total_CP(elem_compound[i], e) for e in XEng_list])
        for j in range(len(elem_type)):
            cs[f"{elem_compound[i]}-{elem_type[j]}"] = xraylib.CS_Total_CP(elem_compound[i], em_E[elem_type[j]])
    return cs


def make_param(elem_type, elem_compound, cs_x, cs_line, rho=1.0, pix=1.0e-4, em_cs=None):
    """param dict from arrays: cs_x[k] (scalar or per-angle array) for compound k at the incident
    energy, cs_line[k][e] for compound k at element e's line, rho scalar or per-compound."""
    K = len(elem_type)
    cs = {"elem_type": list(elem_type), "elem_compound": list(elem_compound)}
    for k, comp in enumerate(elem_compound):
        cs[f"{comp}-x"] = np.atleast_1d(np.asarray(cs_x[k], np.float64))
        for e, el in enumerate(elem_type):
            cs[f"{comp}-{el}"] = float(cs_line[k][e])
    if np.isscalar(rho):
        rho = {comp: float(rho) for comp in elem_compound}
    if em_cs is None:
        em_cs = {el: np.ones(1) for el in elem_type}
    return {"elem_type": list(elem_type), "elem_compound": list(elem_compound), "rho": dict(rho),
            "pix": float(pix), "cs": cs, "em_cs": em_cs, "nelem": K}


def extract_em_cs_elem(param, elem, use_ref=False):
    """Emission cross-sections of one element (FL/atten_util.py:63-76): param['em_cs'][elem]
    as given, or with use_ref an (n, n_ref) array whose column i holds
    param['em_cs'][f'{elem}_ref{i+1}'] (n = len(param['em_cs'][elem]), n_ref =
    param['em_cs']['n_ref']).  A missing key raises KeyError naming it; the reference
    prints the error and returns a partly filled array instead."""
    table = param["em_cs"]
    base = table[elem]
    if not use_ref:
        return base
    wanted = ["n_ref"] + [f"{elem}_ref{i + 1}" for i in range(int(table.get("n_ref", 0)))]
    missing = [k for k in wanted if k not in table]
    if missing:
        raise KeyError(f"em_cs of {elem} with references needs param['em_cs'] keys {missing}")
    cols = [np.broadcast_to(np.asarray(table[k], np.float64), (len(base),)) for k in wanted[1:]]
    return np.stack(cols, axis=1) if cols else np.zeros((len(base), 0))


def update_dict_use_ref_at_angle(dict_use_ref, current_angle):
    """Rotating reference spectra with the sample (FL/atten_util.py:79-87).  The reference's
    reference-spectrum composition calls an undefined function (cross_section_util.py:62),
    so only the empty / use_ref=False form is supported."""
    d = deepcopy(dict_use_ref)
    if len(d) > 0 and d.get("use_ref", False):
        raise NotImplementedError("reference-spectrum (use_ref) attenuation is not supported")
    return d


def extract_single_value_param(param, energy_angle_id, dict_use_ref={}):
    """Cross-sections at one angle/energy index (FL/atten_util.py:90-136): per-angle arrays are
    indexed, length-1 arrays kept."""
    elem_type = param["elem_type"]
    elem_comp = param["elem_compound"]
    out = param.copy()
    cs = param["cs"]
    cs_a = deepcopy(cs)
    for comp in elem_comp:
        v = np.atleast_1d(cs[f"{comp}-x"])
        cs_a[f"{comp}-x"] = v if len(v) == 1 else v[energy_angle_id]
    em = param["em_cs"]
    em_a = deepcopy(em)
    for el in elem_type:
        v = np.atleast_1d(em[el])
        em_a[el] = v if len(v) == 1 else v[energy_angle_id]
    if len(dict_use_ref) > 0 and dict_use_ref.get("use_ref", False):
        raise NotImplementedError("reference-spectrum (use_ref) attenuation is not supported")
    out["cs"] = cs_a
    out["em_cs"] = em_a
    return out


def coef_xray(param_angle_energy):
    """coef_x[j] = cs[comp_j-x] * rho[comp_j] * pix, rounded to float32
    (torch_extract_3D_mu_xray, FL/torch_functions.py:48-49)."""
    cs = param_angle_energy["cs"]
    rho, pix = param_angle_energy["rho"], param_angle_energy["pix"]
    vals = []
    for comp in cs["elem_compound"]:
        v = np.asarray(cs[f"{comp}-x"], np.float64).reshape(-1)
        if v.size != 1:
            raise ValueError(f"cs['{comp}-x'] has {v.size} values: extract the angle's value first")
        vals.append(v[0] * rho[comp] * pix)
    return np.asarray(vals, np.float64).astype(np.float32)


def coef_xrf(param_angle_energy):
    """coef[j, i] = cs[comp_j-elem_i] * rho[comp_j] * pix, rounded to float32
    (torch_extract_3D_mu_xrf, FL/torch_functions.py:72-80)."""
    cs = param_angle_energy["cs"]
    rho, pix = param_angle_energy["rho"], param_angle_energy["pix"]
    elem_type, comps = cs["elem_type"], cs["elem_compound"]
    out = np.zeros((len(comps), len(elem_type)), np.float64)
    for j, comp in enumerate(comps):
        for i, el in enumerate(elem_type):
            out[j, i] = float(np.asarray(cs[f"{comp}-{el}"]).reshape(-1)[0]) * rho[comp] * pix
    return out.astype(np.float32)


def cal_total_atten_at_angles(fsave, param):
    """Total attenuation files from the fluorescence and incident file sets
    (FL/atten_util.py:41-58, the same function as FL/core.py:84-101)."""
    from .core import cal_total_atten_at_angles as total
    return total(fsave, param)
