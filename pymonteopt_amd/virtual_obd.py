"""The virtual OBD data-set pipeline (pymopt/virtualOBD.py) on the MI355X.

One iteration of `calc` makes one synthetic subject:
1. `determining_params_range` picks the population from one np.random.rand(): osteoporosis
   at a <= 0.4, normal at a >= 0.6, osteopenia between (:47-60);
2. `generate_variable_params` draws the subject's skin, subcutis and bone parameters
   (:200-203);
3. `generate_bone_model` grows the Turing-pattern trabecular volume on the device
   (TuringPattern: hipFFT steps, threshold, peak count, binarisation) and returns the
   labels and the realised BV/TV (:113-125);
4. `calc_montecalro` maps the subject onto TuringModel_Rectangular's keys (mu_s = mu_s' /
   (1 - g); the marrow space takes the subcutis optics), walks nPh packets and saves the
   run (:127-156);
5. `calc_ray_tracing` writes the three detector curves as CSV (:158-195):
   * `_B`: top-face Rd(r) of v_z < 0 exits with z < 1e-8;
   * `_F`: Tt(r) of every v_z > 0 exit. The reference's `v_z > 0 & (z > l - margin)`
     parses as `v_z > (0 & ...)`, so its exit-face test drops out; kept as is;
   * `_L`: the side line profile over z of v_x < 0 exits with |y| < 5 mm.

The three curves are the walk's own FP64 tallies (`rd_top_r`, `tt_r`, `line_z`: 300 and
400 bins of 0.1 mm). No per-packet dump crosses PCIe, where the reference copies 28 B per
packet back and bins it on the host. `calc_ray_tracing` also takes a dump dict
{'p', 'v', 'w', 'nPh'}, as the reference's does.

Differences:
* `monte_params` has 'symmetrization' and 'enclosure'. These are keys of the unverified
  voxel classes (pymopt/voxel_unverified/_classes.py:1152). The TuringModel the script
  builds has neither, and its set_params raises KeyError on them
  (pymopt/voxel_gpu/_classes.py:554-575). `calc_montecalro` leaves both keys out.
* The reference writes under ./model_result, ./monte_result and ./opt_result; `calc`
  takes the root directory as `root`.
* nPh, the device and the root are arguments; their defaults are the script's values.
"""
from __future__ import annotations

import datetime
import gc
import os
import time

import numpy as np

from .param_generator import generate_variable_params

__all__ = ["range_params_norm", "range_params_osteoporosis", "range_params_osteopenia", "model_params",
           "th_coef", "monte_params", "determining_params_range", "generate_bone_model", "calc_montecalro",
           "calc_ray_tracing", "calc", "run"]

repetitions = 1000
nPh = 1e7

range_params_norm = {
    "th_dermis": [1, 2],
    "ma_dermis": [0.00633, 0.08560],
    "msd_dermis": [1.420, 2.506],
    "th_subcutaneus": [1, 6],
    "ma_subcutaneus": [0.00485, 0.01239],
    "msd_subcutaneus": [0.83, 1.396],
    "bv_tv": [0.134, 0.028],
    "th_cortical": [0.804, 0.149],
    "corr": 0.54,
}
range_params_osteoporosis = dict(range_params_norm, bv_tv=[0.085, 0.022], th_cortical=[0.487, 0.138])
range_params_osteopenia = dict(range_params_norm, bv_tv=[0.103, 0.030], th_cortical=[0.571, 0.173])

model_params = {
    "grid": 30, "dx": 1 / 30, "dt": 1, "du": 0.0002, "dv": 0.01, "length": 10, "repetition": 100,
    "voxelsize": 0.0306, "seed": False, "ct_coef": 4.5e4, "tile_num_xz": 2, "tile_num_y": 4,
}

th_coef = np.array([-10.93021385, 2.62630274, -0.50913966, 0.60371039])

monte_params = {
    "voxel_space": model_params["voxelsize"],
    "xz_size": 17.15,
    "symmetrization": True,
    "enclosure": True,
    "n_space": 1.4, "n_trabecular": 1.55, "n_cortical": 1.55, "n_subcutaneus": 1.4, "n_dermis": 1.4, "n_air": 1.,
    "ma_trabecular": 0.02374, "ma_cortical": 0.02374,
    "ms_trabecular": 20.588, "ms_cortical": 20.588,
    "g_space": 0.90, "g_trabecular": 0.90, "g_cortical": 0.90, "g_subcutaneus": 0.90, "g_dermis": 0.90,
}

# keys of the unverified voxel classes that the Turing geometries do not take
_UNVERIFIED_KEYS = ("symmetrization", "enclosure")

# the detector binning of calc_ray_tracing (:159-172)
RD_BINS, RD_WIDTH = 300, 30 / 300
LINE_BINS, LINE_WIDTH, LINE_Y = 400, 40 / 400, 5


def determining_params_range():
    """:47-60."""
    a = np.random.rand()
    if a <= 0.4:
        print("Osteoporosis")
        return range_params_osteoporosis
    if a >= 0.6:
        print("Normal")
        return range_params_norm
    print("Osteoprnia")
    return range_params_osteopenia


def generate_bone_model(bv_tv, path, model_params, device="cuda:0"):
    """:113-125: the Turing-pattern label volume and its realised BV/TV."""
    from .modeling import TuringPattern
    model_params["bv_tv"] = bv_tv
    tp = TuringPattern(device=device)
    tp.set_params(model_params)
    tp.set_threshold_func_coef(th_coef)
    os.makedirs(path, exist_ok=True)
    u = tp.modeling(path, save_dicom=False)
    bvtv_ = tp.bv_tv_real
    del tp
    gc.collect()
    return u, bvtv_


def subject_params(vp, params, nn=0):
    """Writes subject nn of `vp` into the model keys in place (:127-145) and returns params."""
    params["th_dermis"] = vp["th_dermis"][nn]
    params["ma_dermis"] = vp["ma_dermis"][nn]
    params["ms_dermis"] = vp["msd_dermis"][nn] / (1 - params["g_dermis"])
    params["th_subcutaneus"] = vp["th_subcutaneus"][nn]
    params["ma_subcutaneus"] = vp["ma_subcutaneus"][nn]
    params["ms_subcutaneus"] = vp["msd_subcutaneus"][nn] / (1 - params["g_subcutaneus"])
    params["ma_space"] = vp["ma_subcutaneus"][nn]
    params["ms_space"] = vp["msd_subcutaneus"][nn] / (1 - params["g_space"])
    params["bv_tv"] = vp["bv_tv"][nn]
    params["th_cortical"] = vp["th_cortical"][nn]
    return params


def _subject_model(vp, params, u, nPh, device):
    """The subject's VoxelTuringModel, built on the host (the layers wrapped around u)."""
    from .voxel import VoxelTuringModel
    subject_params(vp, params)
    model = VoxelTuringModel(nPh=nPh, model_name="TuringModel", nr=RD_BINS, dr=RD_WIDTH, nl=LINE_BINS,
                             dl=LINE_WIDTH, line_y=LINE_Y, fluence=False, device=device)
    model.set_model(u)
    model.build(**{k: v for k, v in params.items() if k not in _UNVERIFIED_KEYS})
    return model


def _walk_and_save(model, iteral, path):
    start = time.time()
    model = model.start(iteral)
    print("%s sec" % (time.time() - start))
    print("# %s" % iteral)
    print("Save -> %s" % path)
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    model.save_result(path, coment="for machine learning")
    return model.get_result()


def calc_montecalro(vp, iteral, params, path, u, nPh=nPh, device="cuda:0"):
    """:127-156: walks nPh packets through the subject's bone model with the walk seeded by
    `iteral`, saves <path>_LID.pkl.bz2 and <path>_info.json, and returns (result, params)."""
    print()
    print("###################################")
    print("# %s" % iteral)
    model = _subject_model(vp, params, u, nPh, device)
    res = _walk_and_save(model, iteral, path)
    del model
    gc.collect()
    return res, params


def _device_curve(res, key, axis_key, nn, dr):
    """The walk's tally `key` when it has the detector's binning, else None."""
    if key not in res or axis_key not in res:
        return None
    y, x = np.asarray(res[key]), np.asarray(res[axis_key])
    if y.size != nn or x.size != nn or (nn > 1 and not np.isclose(x[1] - x[0], dr)):
        return None
    return np.array([i * dr for i in range(nn + 1)])[:-1], y


def detector_curves(res):
    """(alpha, Rd) top face, (alpha, Tt), (alpha, line) side, from the walk's tallies or
    from an exit dump (:159-172)."""
    from . import metrics as met
    B = _device_curve(res, "Rd_top_r", "r", RD_BINS, RD_WIDTH)
    F = _device_curve(res, "Tt_r", "r", RD_BINS, RD_WIDTH)
    L = _device_curve(res, "line_z", "z", LINE_BINS, LINE_WIDTH)
    if B is not None and F is not None and L is not None:
        return B, F, L
    if not all(k in res for k in ("p", "v", "w")):
        raise ValueError("the result holds neither the detector tallies (300 x 0.1 mm radial, "
                         "400 x 0.1 mm line) nor an exit dump")
    p, v, w, n = np.asarray(res["p"]), np.asarray(res["v"]), np.asarray(res["w"]), res["nPh"]
    margin = 1e-8
    ind = np.where((v[2] < 0) & (p[2] < margin))[0]
    B = met.radialDistance(p[:, ind], w[ind], RD_BINS, RD_WIDTH, n)
    ind = np.where(v[2] > 0)[0]
    F = met.radialDistance(p[:, ind], w[ind], RD_BINS, RD_WIDTH, n)
    ind = np.where(v[0] < 0)[0]
    L = met.lineDistance(p[:, ind], w[ind], LINE_BINS, LINE_WIDTH, n, y_range=LINE_Y)
    return B, F, L


def calc_ray_tracing(res, monte_params, path, alias_name):
    """:158-195: writes <path>_B.csv, <path>_F.csv and <path>_L.csv, one column named
    <alias_name>_B/_F/_L indexed by the bins' inner edges."""
    import pandas as pd
    curves = detector_curves(res)
    print("# Ray Tracing save -> %s" % path)
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    for suffix, (alpha, y) in zip(("_B", "_F", "_L"), curves):
        df = pd.DataFrame()
        df[alias_name + suffix] = y
        df.index = alpha
        df.to_csv(path + suffix + ".csv")


def _draw_subject(iteral):
    """The subject's parameters (:199-203), or None for an invalid draw."""
    gvp = generate_variable_params()
    range_params = determining_params_range()
    gvp.set_params(range_params)
    vp = gvp.generate(1)
    if not (vp["bv_tv"][0] > 0 and vp["th_cortical"][0] > 0):
        print("Invalid parameter was generated")
        print("BV/TV : ", vp["bv_tv"][0])
        print("th_cortical : ", vp["th_cortical"][0])
        return None
    return vp


def _paths(root, alias_name):
    return tuple(os.path.join(root, d, "") + alias_name for d in ("model_result", "monte_result", "opt_result"))


def calc(iteral, *, root=".", nPh=nPh, device="cuda:0", model_params=model_params, monte_params=monte_params):
    """:198-237: one subject. Returns the alias name, or None when the draw is invalid
    (BV/TV or cortical thickness <= 0)."""
    vp = _draw_subject(iteral)
    if vp is None:
        return None
    alias_name = "-".join((str(datetime.datetime.now().isoformat()).split(".")[0]).split(":")) + "_it" + f"{iteral:04}"
    print("### iteral number ", iteral)
    print("Alias name: ", alias_name)
    model_path, monte_path, opt_path = _paths(root, alias_name)
    u, bv_tv = generate_bone_model(vp["bv_tv"][0], model_path + "_dicom", model_params, device)
    print("it: ", iteral, ", change bvtv: ", vp["bv_tv"][0], "-->", bv_tv)
    vp["bv_tv"][0] = bv_tv
    res, params_ = calc_montecalro(vp, iteral, monte_params, monte_path, u, nPh=nPh, device=device)
    print("###### end monte calro in it: ", iteral)
    calc_ray_tracing(res, params_, opt_path, alias_name)
    print("")
    print("############### End %s it ###################" % iteral)
    print("")
    return alias_name


def _prepare(iteral, root, nPh, device, model_params, monte_params):
    """The host side of one subject and its bone growth: draws, Turing volume, wrapped model."""
    vp = _draw_subject(iteral)
    if vp is None:
        return None
    alias_name = "-".join((str(datetime.datetime.now().isoformat()).split(".")[0]).split(":")) + "_it" + f"{iteral:04}"
    print("### iteral number ", iteral, " alias ", alias_name)
    model_path, monte_path, opt_path = _paths(root, alias_name)
    u, bv_tv = generate_bone_model(vp["bv_tv"][0], model_path + "_dicom", model_params, device)
    vp["bv_tv"][0] = bv_tv
    params = dict(monte_params)
    model = _subject_model(vp, params, u, nPh, device)
    return iteral, alias_name, model, params, monte_path, opt_path


def _finish(prep, stream):
    """The device side of one subject on `stream`: upload, walk, result files, detector CSVs."""
    import torch
    iteral, alias_name, model, params, monte_path, opt_path = prep
    with torch.cuda.stream(stream):
        res = _walk_and_save(model, iteral, monte_path)
        calc_ray_tracing(res, params, opt_path, alias_name)
    print("############### End %s it ###################" % iteral)
    return alias_name


def run(repetitions=repetitions, *, overlap=True, root=".", nPh=nPh, device="cuda:0", model_params=model_params,
        monte_params=monte_params):
    """The script's main loop (:241-247). Returns the alias names of the valid subjects.

    With overlap (the default), subject i walks on its own HIP stream in a worker thread
    while the main thread draws subject i + 1, grows its bone and wraps its layers. The
    draws from np.random keep the script's order, and each walk is seeded by its
    iteration number, so the files are those of the sequential loop (overlap=False, which
    calls `calc` as the script does). The subjects' parameters go to copies of
    monte_params instead of the shared dict."""
    aliases = []
    if not overlap:
        for iteral in range(repetitions):
            a = calc(iteral, root=root, nPh=nPh, device=device, model_params=model_params, monte_params=monte_params)
            if a is not None:
                aliases.append(a)
    else:
        from concurrent.futures import ThreadPoolExecutor

        import torch
        stream = torch.cuda.Stream(device=device)
        with ThreadPoolExecutor(max_workers=1) as pool:
            pending = None
            for iteral in range(repetitions):
                prep = _prepare(iteral, root, nPh, device, model_params, monte_params)
                if pending is not None:
                    aliases.append(pending.result())    # at most one subject in flight
                    pending = None
                if prep is not None:
                    pending = pool.submit(_finish, prep, stream)
                del prep
            if pending is not None:
                aliases.append(pending.result())
    print()
    print("######################")
    print(datetime.datetime.now())
    return aliases


if __name__ == "__main__":
    run()
